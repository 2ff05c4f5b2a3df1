"""The autotuner's kernel variants give the same bits: every predictor
variant (records gathered per group x threads per block, gradient terms in
the row walk or deferred to a queue) and every PPE
product variant (entries per load batch; columns as 16-bit steps or 32-bit
ids) against the default on the same state, and a tuned run against an
untuned one."""
import numpy as np
import pytest

import voxarmps_amd as vx

pytestmark = pytest.mark.gpu


def _case():
    return vx.dam_break(3, 0.005, (16, 16, 12), (48, 20, 13))


@pytest.mark.parametrize("precision", [vx._abi.PREC_FP64, vx._abi.PREC_MIXED])
def test_kernel_variants_are_bit_identical(hip, precision):
    case = _case()
    p = vx.dam3d_params(case, tol=1e-10)
    p.precision = precision
    s = vx.Simulation(p, case, backend=hip)
    s.init()
    s.step(3)
    for st in ("PNUM", "NEIGHBOR", "V_FRACTION", "VISCOSITY"):
        getattr(s, st)()
    ref = None
    for v in range(9):
        s.set_tuning([0, 0, v, 0])
        s.PREDICTION()
        out = [s.get(n) for n in ("XSTAR", "YSTAR", "ZSTAR", "USTAR", "VSTAR", "WSTAR")]
        if ref is None:
            ref = out
        for a, b in zip(out, ref):
            assert np.array_equal(a, b), f"predictor variant {v}"
    s.set_tuning([0, 0, 0, 0])
    s.PREDICTION()
    for st in ("PNUMSTAR", "BCON", "MATRIX"):
        getattr(s, st)()
    ref = None
    for v in range(5):
        s.set_tuning([0, 0, 0, v])
        info = s.CGM()
        out = (s.get("PNEW"), info.iterations)
        # the unknowns are numbered in the list's cell order, so every row's
        # columns ascend: variant 0 reads them as ranks in its blocks' LDS
        # windows, variant 4 as 16-bit steps
        st = s.status()
        assert st.product_col_bytes == (2 if v in (0, 4) else 4), v
        assert (st.product_window_cols > 0) == (v == 0), v
        if ref is None:
            ref = out
        assert np.array_equal(out[0], ref[0]) and out[1] == ref[1], f"product variant {v}"
    with pytest.raises(vx.VoxError):
        s.set_tuning([0, 0, 9, 0])
    with pytest.raises(vx.VoxError):
        s.set_tuning([0, 0, 0, 5])
    s.close()


def test_autotuned_run_matches_untuned(hip):
    case = _case()
    p = vx.dam3d_params(case, tol=1e-10)
    a = vx.Simulation(p, case, backend=hip)
    b = vx.Simulation(p, case, backend=hip)
    a.init()
    b.init()
    b.autotune(reps=2)
    t = b.tuning()
    print("tuning (walk S, walk L, predictor, product):", t)
    assert t[0] in (1, 2) and t[1] in (1, 2) and 0 <= t[2] < 9 and 0 <= t[3] < 5
    a.step(10)
    b.step(10)
    for n in ("X", "Y", "Z", "U", "V", "W", "P"):
        assert np.array_equal(a.get(n), b.get(n)), n
    a.close()
    b.close()
