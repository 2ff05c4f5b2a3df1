"""ctypes mirror of include/voxarmps.h (structs, constants, prototypes).

The struct layouts follow the header member by member; ``check_sizes`` compares
them with the library's own ``vox_abi_sizes``.
"""
import ctypes as C

VOX_OK = 0
VOX_ERR_ARG = 1
VOX_ERR_HIP = 2
VOX_ERR_OVERFLOW = 3
VOX_ERR_DOMAIN = 4
VOX_ERR_SOLVER = 5
VOX_ERR_NOLIB = 6
VOX_ERR_STATE = 7

FIELDS = [
    "X", "Y", "Z", "U", "V", "W", "P", "PTYPE",
    "XSTAR", "YSTAR", "ZSTAR", "USTAR", "VSTAR", "WSTAR",
    "UNEW", "VNEW", "WNEW", "PNEW", "PHAT", "N", "NSTAR",
    "NEUT", "TURB1", "TURB2", "TURB3", "C", "MEU", "RHO", "BCON", "SOURCE",
]
F = {name: k for k, name in enumerate(FIELDS)}
INT_FIELDS = {"PTYPE", "BCON"}

LIST_S = 0
LIST_L = 1
PREC_FP64 = 0
PREC_MIXED = 1
SOLVER_JACOBI_CG = 0
SOLVER_LEARNED_FCG = 1
LPC_NPARAM = 19330
TEST_DAM, TEST_TURB, TEST_DROP = 0, 1, 2

_i32 = C.c_int32
_f64 = C.c_double


class VoxParams(C.Structure):
    _fields_ = [
        ("dim", _i32), ("ktype", _i32),
        ("dl", _f64), ("re_s", _f64), ("re_l", _f64), ("delta", _f64), ("lambda_", _f64),
        ("rho1", _f64), ("nu1", _f64), ("rho2", _f64), ("nu2", _f64),
        ("gx", _f64), ("gy", _f64), ("gz", _f64),
        ("dt", _f64),
        ("dt_mode", _i32), ("turb", _i32),
        ("courant", _f64), ("c01", _f64), ("c02", _f64), ("dt_max", _f64),
        ("dirichlet", _f64), ("relaxp", _f64), ("coll", _f64), ("cc", _f64),
        ("vel_cor", _f64), ("cs", _f64), ("ncorrection", _f64), ("vmax", _f64),
        ("pmin", _f64), ("pmax", _f64),
        ("fluid2_type", _i32), ("kh_correction", _i32),
        ("n_pl", _f64), ("meu0", _f64), ("phi", _f64), ("cohes", _f64),
        ("xmin", _f64), ("xmax", _f64), ("ymin", _f64), ("ymax", _f64),
        ("zmin", _f64), ("zmax", _f64),
        ("test_kind", _i32), ("coef2d", _i32), ("precision", _i32), ("solver", _i32),
        ("cg_tol", _f64),
        ("cg_maxit", _i32), ("cg_check", _i32), ("cap_s", _i32), ("cap_l", _i32),
        ("ir_eta", _f64), ("ir_maxouter", _i32), ("warm_start", _i32),
        ("list_mode", _i32), ("lpc_cap_pct", _i32),
        ("method", _i32), ("matopt", _i32), ("srcopt", _i32), ("ref_gpu", _i32), ("wc_beta", _f64),
    ]

    def copy(self):
        out = VoxParams()
        C.memmove(C.byref(out), C.byref(self), C.sizeof(VoxParams))
        return out

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class VoxSolveInfo(C.Structure):
    _fields_ = [
        ("iterations", _i32), ("converged", _i32), ("unknowns", _i32), ("nnz", _i32),
        ("relres", _f64), ("bnorm", _f64),
    ]


class VoxStatus(C.Structure):
    _fields_ = [
        ("nan_count", _i32), ("out_of_box", _i32), ("max_count_s", _i32), ("max_count_l", _i32),
        ("collisions", _i32), ("collision_cand", _i32), ("steps", _i32), ("last_iters", _i32),
        ("last_relres", _f64), ("n0_s", _f64), ("n0_l", _f64), ("time", _f64),
        ("lpc_fallbacks", _i32), ("lpc_baseline", _i32), ("product_col_bytes", _i32),
        ("product_window_cols", _i32),
    ]


_P = C.c_void_p
_ip = C.POINTER(C.c_int)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_pp = C.POINTER(VoxParams)

# name -> (restype, argtypes)
PROTOTYPES = {
    "vox_last_error": (C.c_char_p, []),
    "vox_version": (C.c_char_p, []),
    "vox_abi_sizes": (C.c_int, [_i32p, _i32p, _i32p]),
    "vox_device_count": (C.c_int, [_ip]),
    "vox_params_default": (C.c_int, [_pp, C.c_int, C.c_double]),
    "vox_params_box": (C.c_int, [_pp, C.c_int, _dp, _dp, _dp]),
    "vox_params_main": (C.c_int, [_pp, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                  C.c_char_p, C.c_int, C.c_double, C.c_double, C.c_double,
                                  C.c_double, C.c_double, C.c_int, _dp, _dp, _dp]),
    "vox_main_input": (C.c_char_p, [C.c_char_p, C.c_int]),
    "vox_create": (C.c_int, [C.POINTER(_P), _pp, C.c_int, C.c_int, C.c_int, C.c_int]),
    "vox_destroy": (C.c_int, [_P]),
    "vox_get_params": (C.c_int, [_P, _pp]),
    "vox_set_params": (C.c_int, [_P, _pp]),
    "vox_upload": (C.c_int, [_P, C.c_int, _P]),
    "vox_download": (C.c_int, [_P, C.c_int, _P]),
    "vox_upload_state": (C.c_int, [_P, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _i32p]),
    "vox_density": (C.c_int, [_P]),
    "vox_neighbor": (C.c_int, [_P]),
    "vox_pnum": (C.c_int, [_P, C.c_int]),
    "vox_pnumstar": (C.c_int, [_P]),
    "vox_sps": (C.c_int, [_P]),
    "vox_vfraction": (C.c_int, [_P]),
    "vox_viscosity": (C.c_int, [_P]),
    "vox_prediction": (C.c_int, [_P]),
    "vox_bcon": (C.c_int, [_P]),
    "vox_ppe_assemble": (C.c_int, [_P]),
    "vox_ppe_solve": (C.c_int, [_P, C.POINTER(VoxSolveInfo)]),
    "vox_pressurecalc": (C.c_int, [_P, C.POINTER(VoxSolveInfo)]),
    "vox_phat": (C.c_int, [_P]),
    "vox_pressgrad": (C.c_int, [_P]),
    "vox_euler": (C.c_int, [_P]),
    "vox_collision": (C.c_int, [_P]),
    "vox_prepdata": (C.c_int, [_P]),
    "vox_dtcalc": (C.c_int, [_P]),
    "vox_init": (C.c_int, [_P]),
    "vox_step": (C.c_int, [_P, C.c_int]),
    "vox_sync": (C.c_int, [_P]),
    "vox_get_status": (C.c_int, [_P, C.POINTER(VoxStatus)]),
    "vox_get_n0": (C.c_int, [_P, _dp, _dp]),
    "vox_set_regions": (C.c_int, [_P, C.c_int, _i32p, _dp]),
    "vox_autotune": (C.c_int, [_P, C.c_int, _i32p]),
    "vox_get_tuning": (C.c_int, [_P, _i32p]),
    "vox_set_tuning": (C.c_int, [_P, _i32p]),
    "vox_set_n0": (C.c_int, [_P, C.c_double, C.c_double]),
    "vox_get_neighbors": (C.c_int, [_P, C.c_int, _i32p, _i32p, C.c_int]),
    "vox_get_neighbors_slot": (C.c_int, [_P, C.c_int, _i32p, C.c_int]),
    "vox_get_cells": (C.c_int, [_P, C.c_int, _i32p]),
    "vox_ppe_export": (C.c_int, [_P, _i32p, _i32p, _i32p, _i32p, _dp, _dp, _i32p]),
    "vox_ppe_load": (C.c_int, [_P, C.c_int32, C.c_int32, _i32p, _i32p, _dp, _dp, _i32p]),
    "vox_vtu_count": (C.c_int, [C.c_char_p, _ip]),
    "vox_vtu_read": (C.c_int, [C.c_char_p, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                               _i32p, _ip, _ip, _ip]),
    "vox_prof_enable": (C.c_int, [_P, C.c_int]),
    "vox_prof_only": (C.c_int, [_P, C.c_char_p]),
    "vox_prof_get": (C.c_int, [_P, C.c_char_p, _dp, _i64p]),
    "vox_prof_reset": (C.c_int, [_P]),
    "vox_grd_count": (C.c_int, [C.c_char_p, _ip]),
    "vox_grd_read": (C.c_int, [C.c_char_p, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                               _i32p, _ip, _ip, _ip]),
    "vox_grd_write": (C.c_int, [C.c_char_p, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                _i32p]),
    "vox_write_vtu": (C.c_int, [_P, C.c_char_p, C.c_int]),
    "vox_lpc_create": (C.c_int, [C.POINTER(_P), C.c_uint64]),
    "vox_lpc_destroy": (C.c_int, [_P]),
    "vox_lpc_get_weights": (C.c_int, [_P, _fp]),
    "vox_lpc_set_weights": (C.c_int, [_P, _fp]),
    "vox_lpc_dump": (C.c_int, [_P, _P, _i32p]),
    "vox_lpc_count": (C.c_int, [_P, _i32p, _i64p]),
    "vox_lpc_eval": (C.c_int, [_P, C.c_int32, _fp, _fp, _fp]),
    "vox_lpc_loss_grad": (C.c_int, [_P, C.c_int32, C.c_uint64, _dp, _fp]),
    "vox_lpc_train": (C.c_int, [_P, C.c_int32, C.c_double, C.c_uint64, _dp, _dp]),
    "vox_attach_lpc": (C.c_int, [_P, _P]),
    "vox_cgm_slot": (C.c_int, [C.c_int, C.c_int, _P, _P, _P, _P, _P, C.c_double, C.c_int,
                               C.POINTER(VoxSolveInfo)]),
    "vox_cgm_slot_noalloc": (C.c_int, [C.c_int, C.c_int, _P, _P, _P, _P, _P, C.c_double, C.c_int,
                                       C.POINTER(VoxSolveInfo)]),
}


def bind(lib, prototypes=PROTOTYPES, prefix_map=None):
    """Attach restype/argtypes to every function of ``prototypes`` found in lib."""
    for name, (res, args) in prototypes.items():
        fname = prefix_map(name) if prefix_map else name
        fn = getattr(lib, fname)
        fn.restype = res
        fn.argtypes = args
    return lib


def check_sizes(lib):
    a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
    lib.vox_abi_sizes(C.byref(a), C.byref(b), C.byref(c))
    got = (a.value, b.value, c.value)
    want = (C.sizeof(VoxParams), C.sizeof(VoxSolveInfo), C.sizeof(VoxStatus))
    if got != want:
        raise RuntimeError(f"ABI struct size mismatch: library {got}, python {want}")
    return got
