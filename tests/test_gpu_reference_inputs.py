"""The reference's own inputs (tests/golden: its inputs/input-dam.grd,
input-drop.grd, input-shear.grd, input-turb.grd and input-dam3d.grd) run with the reference's
constants through the GPU library and through the oracle. Bar as for the other trajectories: positions,
velocities and pressure within 1e-6 of their scale after 30 steps (100 for
the periodic case; the solver tolerance is the only difference), neighbour
lists bit-exact at the end."""
import dataclasses
import gzip
import hashlib
import os

import numpy as np
import pytest

import voxarmps_amd as vx

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DL = 0.01
SHA256 = {"input-turb.grd": "8ba1584f12cc47e89c4d49f65d17fa42f0d3c66995e660e96cad69bd5ece9378",
          "input-dam3d.grd": "ec882df02442f15e81f5943dceed3469c6383afabf2cc74c1473c6288eaa1744"}


def read_golden(name, dim, tmp):
    """A fixture as a Case at spacing DL; .gz fixtures are checked against the
    reference file's hash and unpacked first."""
    path = os.path.join(GOLDEN, name)
    if not os.path.exists(path):
        with gzip.open(path + ".gz", "rb") as f:
            data = f.read()
        assert hashlib.sha256(data).hexdigest() == SHA256[name]
        path = os.path.join(tmp, name)
        with open(path, "wb") as f:
            f.write(data)
    return dataclasses.replace(vx.cases.read_grd(path, dim), l0=DL)


def _dam(tmp):
    """input-dam.grd with main.cu's constants and box (bounds +- DL, 692-715)."""
    case = read_golden("input-dam.grd", 2, tmp)
    p = vx.set_box(vx.default_params(2, DL), case)
    return case, p


def _shear(tmp):
    """input-shear.grd, the shear-driven cavity: no gravity (main.cu:319-328);
    the lid is the wall rows the file gives u = 0.5 and 1.0."""
    case = read_golden("input-shear.grd", 2, tmp)
    p = vx.set_box(vx.default_params(2, DL), case)
    p.gx = p.gy = 0.0
    return case, p


def _turb(tmp):
    """input-turb.grd as main.cu runs the Poiseuille test: SPS on, no gravity,
    velCor 0.8 (305-322, 509-512), the periodic boundary of BOUNDARIE."""
    case = read_golden("input-turb.grd", 2, tmp)
    p = vx.set_box(vx.default_params(2, DL), case)
    p.test_kind = vx._abi.TEST_TURB
    p.turb = 1
    p.gx = p.gy = 0.0
    p.vel_cor = 0.8
    return case, p


def _dam3d(tmp):
    """input-dam3d.grd: 3D, one radius of 3.2 DL with the particle in its own
    list (main.cu:252-260), box bounds +- DL."""
    case = read_golden("input-dam3d.grd", 3, tmp)
    p = vx.set_box(vx.default_params(3, DL), case)
    return case, p


def _drop(tmp):
    """input-drop.grd as main.cu reads it for the drop test: u = -x, v = y
    (645-651) and the box widened by 100 DL, and by 10000 DL below (692-700)."""
    case = read_golden("input-drop.grd", 2, tmp)
    case.u[:] = -1 * case.x
    case.v[:] = 1 * case.y
    p = vx.default_params(2, DL)
    p.test_kind = vx._abi.TEST_DROP
    p.xmin, p.xmax = float(case.x.min()) - 100 * DL, float(case.x.max()) + 100 * DL
    p.ymin, p.ymax = float(case.y.min()) - 10000 * DL, float(case.y.max()) + 100 * DL
    return case, p


# the Poiseuille case runs 100 steps: its fastest fluid column (0.5 m/s, 2 DL
# before Xmax) then crosses Xmax and wraps to Xmin
@pytest.mark.parametrize("make,steps",
                         [(_dam, 30), (_drop, 30), (_shear, 30), (_turb, 100), (_dam3d, 30)],
                         ids=["input-dam", "input-drop", "input-shear", "input-turb", "input-dam3d"])
def test_reference_input_trajectory(hip, oracle, make, steps, tmp_path):
    case, p = make(str(tmp_path))
    p.cg_tol = 1e-10
    g = vx.Simulation(p, case, backend=hip)
    c = vx.Simulation(p, case, backend=oracle)
    g.init()
    c.init()
    assert g.n0() == c.n0()
    x0 = c.get("X")
    g.step(steps)
    c.step(steps)
    for n in ("X", "Y", "Z", "U", "V", "W", "P"):
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
    if p.gy != 0.0:
        assert np.max(np.abs(c.get("P"))) > 0
    if make is _shear:
        f = case.ptype > 0
        assert np.max(np.abs(c.get("U")[f])) > 0                # the lid drives the fluid
    if p.test_kind == vx._abi.TEST_TURB:
        assert np.count_nonzero(c.get("X") < x0 - 0.5) > 0      # periodic wrap exercised
    g.close()
    c.close()
