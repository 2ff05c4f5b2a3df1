"""Degenerate inputs on the GPU, against the oracle where it applies: a single
fluid particle (every PPE row empty, no unknown or a 1-row system), a column
so thin that every fluid particle is free surface (no unknowns at all), in
each precision and with each solver."""
import numpy as np
import pytest

import voxarmps_amd as vx

pytestmark = pytest.mark.gpu


def _tiny(water):
    # a one-layer-wall tank of 12 x 10 cells around a `water` column
    return vx.dam_break(2, 0.01, water, (12, 10), 1, 2, name=f"tiny{water}")


@pytest.mark.parametrize("water", [(1, 1), (2, 1), (1, 3)])
def test_tiny_columns_match_oracle(hip, oracle, water):
    case = _tiny(water)
    p = vx.dam2d_params(case, tol=1e-10)
    g = vx.Simulation(p, case, backend=hip)
    c = vx.Simulation(p, case, backend=oracle)
    for s in (g, c):
        s.init()
        s.step(10)
    for n in ("X", "Y", "U", "V", "P"):
        a, b = g.get(n), c.get(n)
        assert np.allclose(a, b, rtol=1e-9, atol=1e-12), n
    assert g.status().nan_count == 0
    g.close()
    c.close()


@pytest.mark.parametrize("precision", [vx._abi.PREC_FP64, vx._abi.PREC_MIXED])
@pytest.mark.parametrize("solver", [vx._abi.SOLVER_JACOBI_CG, vx._abi.SOLVER_LEARNED_FCG])
def test_no_unknowns(hip, precision, solver):
    """A single fluid particle in mid-air, more than a radius from every
    wall, is below the free-surface threshold: the PPE has no unknowns, every
    solver returns at once and the pressure is 0."""
    case = _tiny((1, 1))
    w = case.ptype <= 0
    f = np.nonzero(case.ptype > 0)[0][0]
    case.x[f] = 0.5 * (case.x[w].min() + case.x[w].max())
    case.y[f] = 0.5 * (case.y[w].min() + case.y[w].max())
    p = vx.dam2d_params(case)
    p.precision = precision
    p.solver = solver
    s = vx.Simulation(p, case, backend=hip)
    net = None
    if solver == vx._abi.SOLVER_LEARNED_FCG:
        net = vx.LearnedPreconditioner(backend=hip)
        net.set_weights(np.zeros(vx.lpc.NPARAM, np.float32))
        s.attach_lpc(net)
    s.init()
    s.step(3)
    for st in ("PNUM", "NEIGHBOR", "V_FRACTION", "VISCOSITY", "PREDICTION", "PNUMSTAR", "BCON",
               "MATRIX"):
        getattr(s, st)()
    info = s.CGM()
    assert info.unknowns == 0 and info.converged
    assert np.all(s.get("PNEW")[case.ptype > 0] == 0)
    assert s.status().nan_count == 0
    s.close()
    if net is not None:
        net.close()


def test_volume_fraction_single_and_two_phase(hip, oracle):
    """With no PTYPE 2 particle the library writes C = +0 without the walk
    (every fluid-2 sum is +0); the values equal the oracle's walk bit for bit.
    Uploading types with a second fluid brings the walk back."""
    case = vx.jitter(vx.dam_break_2d_c0(), 0.1 * 0.01, seed=8)
    p = vx.dam2d_params(case)
    g = vx.Simulation(p, case, backend=hip)
    c = vx.Simulation(p, case, backend=oracle)
    for s in (g, c):
        s.init()
        s.set("C", np.full(case.tp, 7.0))
        s.V_FRACTION()
    cg, cc = g.get("C"), c.get("C")
    assert np.array_equal(cg, cc) and np.all(cg == 0) and not np.any(np.signbit(cg))
    pt = case.ptype.copy()
    f = pt > 0
    pt[f & (case.y > np.median(case.y[f]))] = 2
    for s in (g, c):
        s.set("PTYPE", pt)
        s.V_FRACTION()
    cg, cc = g.get("C"), c.get("C")
    assert np.array_equal(cg, cc) and np.count_nonzero(cg) > 0
    g.close()
    c.close()


def test_profiling_one_class(hip):
    """vox_prof_only limits the events to one kernel class; NULL brings every
    class back; an unknown class name is refused."""
    import ctypes as C
    case = vx.dam_break_2d_c0()
    p = vx.dam2d_params(case)
    lib = hip.lib

    def launches(s, name):
        t, n = C.c_double(0.0), C.c_int64(0)
        assert lib.vox_prof_get(s.ctx, name, C.byref(t), C.byref(n)) == 0
        return n.value

    with vx.Simulation(p, case, backend=hip) as s:
        s.init()
        assert lib.vox_prof_only(s.ctx, b"no-such-class") != 0
        for only, expect_pred in ((b"spmv", False), (None, True)):
            assert lib.vox_prof_reset(s.ctx) == 0
            assert lib.vox_prof_only(s.ctx, only) == 0
            assert lib.vox_prof_enable(s.ctx, 1) == 0
            s.step(2)
            assert lib.vox_prof_enable(s.ctx, 0) == 0
            assert launches(s, b"spmv") > 0
            assert (launches(s, b"pred") > 0) == expect_pred
