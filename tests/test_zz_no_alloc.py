"""vox_cgm_slot_noalloc: the allocation-free variant of vox_cgm_slot, whose
CG workspace is a static per-GPU array. Two calls in a row, the second on a
smaller system, each bit-identical to vox_cgm_slot on the same input; sizes
beyond the array are refused."""
import ctypes as C

import numpy as np
import pytest

import voxarmps_amd as vx

from helpers import HipMem, slot_system
from test_gpu_cgm_slot import _assembled, cgm_slot

pytestmark = pytest.mark.gpu


def _small(hip):
    case = vx.dam_break(2, 0.01, (10, 20), (30, 25))
    p = vx.dam2d_params(case, tol=1e-10)
    s = vx.Simulation(p, case, backend=hip)
    s.init()
    s.step(3)
    for st in ("PNUM", "NEIGHBOR", "V_FRACTION", "VISCOSITY", "PREDICTION", "PNUMSTAR", "BCON",
               "MATRIX"):
        getattr(s, st)()
    return s


def test_noalloc_matches_allocating_entry(hip):
    mem = HipMem(hip)
    sims = [_assembled(hip, 3), _small(hip)]
    assert sims[1].tp < sims[0].tp
    for s in sims:
        sysm = slot_system(s)
        a = cgm_slot(hip, mem, sysm, s.tp, 1e-10)
        b = cgm_slot(hip, mem, sysm, s.tp, 1e-10, fn="vox_cgm_slot_noalloc")
        assert a[0] == 0 and b[0] == 0, hip.lib.vox_last_error()
        assert np.array_equal(a[2], b[2])
        for k in ("iterations", "converged", "relres", "bnorm"):
            assert getattr(a[1], k) == getattr(b[1], k), k
        A, rhs, ids = sysm[5]["A"], sysm[5]["rhs"], sysm[5]["ids"]
        rel = np.linalg.norm(rhs - A @ b[2][1:][ids]) / np.linalg.norm(rhs)
        assert rel <= 1e-10 * 1.01
        print(f"tp {s.tp}: {b[1].iterations} iterations, relres {rel:.2e}, identical")
        s.close()
    mem.free()


def test_noalloc_refuses_sizes_beyond_its_array(hip):
    f = hip.lib.vox_cgm_slot_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                  C.c_double, C.c_int, C.c_void_p]
    # host-side check: no kernel runs, so the (non-null) addresses are never used
    rc = f(1 << 20, 8, 16, 16, 16, 16, 16, 1e-8, 10, None)
    assert rc == vx._abi.VOX_ERR_ARG
    assert b"exceeds the static workspace" in hip.lib.vox_last_error()
