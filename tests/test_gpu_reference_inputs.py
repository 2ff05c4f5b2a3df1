"""The reference's own 2D inputs (tests/golden: its inputs/input-dam.grd and
inputs/input-drop.grd) run with the reference's constants through the GPU
library and through the oracle. Bar as for the other trajectories: positions,
velocities and pressure within 1e-6 of their scale after 30 steps (the solver
tolerance is the only difference), neighbour lists bit-exact at the end."""
import dataclasses
import os

import numpy as np
import pytest

import voxarmps_amd as vx

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DL = 0.01


def _dam():
    """input-dam.grd with main.cu's constants and box (bounds +- DL, 692-715)."""
    case = dataclasses.replace(vx.cases.read_grd(os.path.join(GOLDEN, "input-dam.grd"), 2), l0=DL)
    p = vx.set_box(vx.default_params(2, DL), case)
    return case, p


def _drop():
    """input-drop.grd as main.cu reads it for the drop test: u = -x, v = y
    (645-651) and the box widened by 100 DL, and by 10000 DL below (692-700)."""
    case = dataclasses.replace(vx.cases.read_grd(os.path.join(GOLDEN, "input-drop.grd"), 2), l0=DL)
    case.u[:] = -1 * case.x
    case.v[:] = 1 * case.y
    p = vx.default_params(2, DL)
    p.test_kind = vx._abi.TEST_DROP
    p.xmin, p.xmax = float(case.x.min()) - 100 * DL, float(case.x.max()) + 100 * DL
    p.ymin, p.ymax = float(case.y.min()) - 10000 * DL, float(case.y.max()) + 100 * DL
    return case, p


@pytest.mark.parametrize("make", [_dam, _drop], ids=["input-dam", "input-drop"])
def test_reference_input_trajectory(hip, oracle, make):
    case, p = make()
    p.cg_tol = 1e-10
    g = vx.Simulation(p, case, backend=hip)
    c = vx.Simulation(p, case, backend=oracle)
    g.init()
    c.init()
    assert g.n0() == c.n0()
    g.step(30)
    c.step(30)
    for n in ("X", "Y", "U", "V", "P"):
        a, b = g.get(n), c.get(n)
        scale = max(np.max(np.abs(b)), 1.0)
        err = np.max(np.abs(a - b))
        print(f"{case.name.split('/')[-1]} {n}: max |gpu - cpu| {err:.3e} (scale {scale:.3e})")
        assert err <= 1e-6 * scale, n
    for w in (vx.LIST_S, vx.LIST_L):
        cg, ig = g.neighbors(w)
        cc, ic = c.neighbors(w)
        assert np.array_equal(cg, cc) and np.array_equal(ig, ic)
    sg, sc = g.status(), c.status()
    assert (sg.nan_count, sg.out_of_box) == (sc.nan_count, sc.out_of_box) == (0, 0)
    assert np.max(np.abs(c.get("P"))) > 0
    g.close()
    c.close()
