"""ctypes binding of the CPU oracle (test infrastructure only).

Returns a backend with the same stage surface as libvoxarmps.so, so the
package's Simulation class drives either one; only tests, smoke() and the
CPU-baseline leg of bench.py use it.
"""
import ctypes as C
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from voxarmps_amd import _abi  # noqa: E402
from voxarmps_amd.sim import Backend  # noqa: E402

LIB_PATH = os.path.join(HERE, "liboracle.so")

_P = C.c_void_p
_i32p = C.POINTER(C.c_int32)
_dp = C.POINTER(C.c_double)
_pp = C.POINTER(_abi.VoxParams)

PROTOTYPES = {
    "orc_create": (C.c_int, [C.POINTER(_P), _pp, C.c_int, C.c_int, C.c_int, C.c_int]),
    "orc_destroy": (None, [_P]),
    "orc_set_params": (C.c_int, [_P, _pp]),
    "orc_get_params": (C.c_int, [_P, _pp]),
    "orc_upload": (C.c_int, [_P, C.c_int, _P]),
    "orc_download": (C.c_int, [_P, C.c_int, _P]),
    "orc_density": (C.c_int, [_P]),
    "orc_neighbor": (C.c_int, [_P]),
    "orc_pnum": (C.c_int, [_P, C.c_int]),
    "orc_pnumstar": (C.c_int, [_P]),
    "orc_sps": (C.c_int, [_P]),
    "orc_vfraction": (C.c_int, [_P]),
    "orc_viscosity": (C.c_int, [_P]),
    "orc_prediction": (C.c_int, [_P]),
    "orc_bcon": (C.c_int, [_P]),
    "orc_ppe_assemble": (C.c_int, [_P]),
    "orc_ppe_solve": (C.c_int, [_P, C.POINTER(_abi.VoxSolveInfo)]),
    "orc_ppe_solve_iccg": (C.c_int, [_P, C.c_int, C.c_int, C.c_double, C.POINTER(_abi.VoxSolveInfo)]),
    "orc_pressurecalc": (C.c_int, [_P, C.POINTER(_abi.VoxSolveInfo)]),
    "orc_phat": (C.c_int, [_P]),
    "orc_pressgrad": (C.c_int, [_P]),
    "orc_euler": (C.c_int, [_P]),
    "orc_collision": (C.c_int, [_P]),
    "orc_prepdata": (C.c_int, [_P]),
    "orc_dtcalc": (C.c_int, [_P]),
    "orc_threads": (C.c_int, []),
    "orc_set_threads": (C.c_int, [C.c_int]),
    "orc_init": (C.c_int, [_P]),
    "orc_step": (C.c_int, [_P, C.c_int]),
    "orc_get_status": (C.c_int, [_P, C.POINTER(_abi.VoxStatus)]),
    "orc_get_n0": (C.c_int, [_P, _dp, _dp]),
    "orc_set_n0": (C.c_int, [_P, C.c_double, C.c_double]),
    "orc_get_neighbors": (C.c_int, [_P, C.c_int, _i32p, _i32p, C.c_int]),
    "orc_max_count": (C.c_int, [_P, C.c_int]),
    "orc_get_cells": (C.c_int, [_P, C.c_int, _i32p]),
    "orc_ppe_export": (C.c_int, [_P, _i32p, _i32p, _i32p, _i32p, _dp, _dp, _i32p]),
    "orc_weight": (C.c_double, [C.c_double, C.c_int, C.c_int, C.c_double]),
    "orc_neighbor_bruteforce": (C.c_int, [_pp, C.c_double, C.c_int, _dp, _dp, _dp, _i32p, _i32p,
                                          C.c_int]),
}

_ORC = None


def oracle_backend():
    global _ORC
    if _ORC is None:
        if not os.path.exists(LIB_PATH):
            from oracle.build import build_oracle
            build_oracle()
        lib = C.CDLL(LIB_PATH)
        _abi.bind(lib, PROTOTYPES)
        _ORC = Backend(lib, "orc_", "oracle")
    return _ORC


def weight(r, ktype, dim, re):
    return oracle_backend().lib.orc_weight(r, ktype, dim, re)


def bruteforce_neighbors(params, re, x, y, z, width):
    import numpy as np
    tp = len(x)
    cnt = np.zeros(tp, np.int32)
    ids = np.full((tp, width), -1, np.int32)
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    z = np.ascontiguousarray(z if z is not None else np.zeros(tp), np.float64)
    rc = oracle_backend().lib.orc_neighbor_bruteforce(
        C.byref(params), re, tp, x.ctypes.data_as(_dp), y.ctypes.data_as(_dp),
        z.ctypes.data_as(_dp), cnt.ctypes.data_as(_i32p), ids.ctypes.data_as(_i32p), width)
    if rc != 0:
        raise RuntimeError(f"bruteforce neighbour search failed: {rc}")
    return cnt, ids
