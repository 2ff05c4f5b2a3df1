"""Host-side mirror of the reference's stage interface over the C ABI.

`Simulation` keeps the particle state in the backend (device memory for the
HIP library) and exposes one method per reference routine, named as in
functions.cuh:56-111 (NEIGHBOR, PNUM, ..., PREPDATA). Any backend exporting the
same C entry points under a prefix can drive it; the product backend is
libvoxarmps.so (prefix ``vox_``).
"""
import ctypes as C
import os

import numpy as np

from . import _abi
from ._abi import F, INT_FIELDS, VoxSolveInfo, VoxStatus

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvoxarmps.so")


class VoxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Backend:
    """A loaded library exposing the stage ABI under `prefix`."""

    def __init__(self, lib, prefix, name, last_error=None):
        self.lib = lib
        self.prefix = prefix
        self.name = name
        self._last_error = last_error

    def fn(self, name):
        return getattr(self.lib, self.prefix + name)

    def check(self, rc, what):
        if rc != 0:
            msg = self._last_error() if self._last_error else ""
            raise VoxError(rc, f"{self.name}.{what}: {msg}")


_HIP = None


def hip_backend(path=None):
    """Load libvoxarmps.so (raises if it has not been built)."""
    global _HIP
    if _HIP is not None and path is None:
        return _HIP
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise VoxError(_abi.VOX_ERR_NOLIB,
                       f"HIP library {p} is not built; run __graft_entry__.build()")
    lib = C.CDLL(p)
    _abi.bind(lib)
    _abi.check_sizes(lib)
    be = Backend(lib, "vox_", "hip", lambda: lib.vox_last_error().decode())
    if path is None:
        _HIP = be
    return be


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class Simulation:
    """MPS particle state plus the per-stage operations of the reference."""

    def __init__(self, params, case, backend=None):
        self.be = backend or hip_backend()
        self.params = params.copy()
        self.case = case
        self.dim = case.dim
        self.tp, self.gp, self.wp, self.fp = case.counts()
        self.ctx = C.c_void_p()
        rc = self.be.fn("create")(C.byref(self.ctx), C.byref(self.params), self.tp, self.gp,
                                  self.wp, self.fp)
        self.be.check(rc, "create")
        for name in ("X", "Y", "Z", "U", "V", "W", "P"):
            self.set(name, getattr(case, name.lower()))
        self.set("PTYPE", case.ptype)

    # ---- lifetime -----------------------------------------------------------
    def close(self):
        if self.ctx:
            self._lpc = None
            self.be.fn("destroy")(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- fields ------------------------------------------------------------------
    def set(self, name, arr):
        if name in INT_FIELDS:
            a = np.ascontiguousarray(arr, dtype=np.int32)
        else:
            a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != (self.tp,):
            raise ValueError(f"{name}: expected {self.tp} values")
        self.be.check(self.be.fn("upload")(self.ctx, F[name], a.ctypes.data_as(C.c_void_p)),
                      f"upload {name}")

    def get(self, name):
        a = np.empty(self.tp, dtype=np.int32 if name in INT_FIELDS else np.float64)
        self.be.check(self.be.fn("download")(self.ctx, F[name], a.ctypes.data_as(C.c_void_p)),
                      f"download {name}")
        return a

    def set_params(self, params):
        self.be.check(self.be.fn("set_params")(self.ctx, C.byref(params)), "set_params")
        self.params = params.copy()

    def get_params(self):
        """The parameters in effect (DTcalculation updates dt in place)."""
        q = self.params.copy()
        self.be.check(self.be.fn("get_params")(self.ctx, C.byref(q)), "get_params")
        return q

    # ---- stages, named after the reference routines ---------------------------------
    def _stage(self, name, *args):
        self.be.check(self.be.fn(name)(self.ctx, *args), name)

    def density(self):
        self._stage("density")

    def NEIGHBOR(self):
        self._stage("neighbor")

    def PNUM(self, loopstarted=True):
        self._stage("pnum", int(bool(loopstarted)))

    def PNUMSTAR(self):
        self._stage("pnumstar")

    def SPS(self):
        self._stage("sps")

    def V_FRACTION(self):
        self._stage("vfraction")

    def VISCOSITY(self):
        self._stage("viscosity")

    def PREDICTION(self):
        self._stage("prediction")

    def BCON(self):
        self._stage("bcon")

    def MATRIX(self):
        self._stage("ppe_assemble")

    def CGM(self):
        info = VoxSolveInfo()
        self.be.check(self.be.fn("ppe_solve")(self.ctx, C.byref(info)), "ppe_solve")
        return info

    def PRESSURECALC(self):
        info = VoxSolveInfo()
        self.be.check(self.be.fn("pressurecalc")(self.ctx, C.byref(info)), "pressurecalc")
        return info

    def PHATCALC(self):
        self._stage("phat")

    def PRESSGRAD(self):
        self._stage("pressgrad")

    def EULERINTEGRATION(self):
        self._stage("euler")

    def COLLISION2(self):
        self._stage("collision")

    def PREPDATA(self):
        self._stage("prepdata")

    def DTcalculation(self):
        self._stage("dtcalc")

    def init(self):
        self._stage("init")

    def step(self, n=1):
        self._stage("step", int(n))

    # ---- queries -------------------------------------------------------------------------
    def status(self):
        st = VoxStatus()
        self.be.check(self.be.fn("get_status")(self.ctx, C.byref(st)), "get_status")
        return st

    def n0(self):
        a, b = C.c_double(), C.c_double()
        self.be.check(self.be.fn("get_n0")(self.ctx, C.byref(a), C.byref(b)), "get_n0")
        return a.value, b.value

    def set_n0(self, n0s, n0l):
        self.be.check(self.be.fn("set_n0")(self.ctx, n0s, n0l), "set_n0")

    def neighbors(self, which=_abi.LIST_S):
        """(counts, ids) with ids row-major, 0-based, width = longest row."""
        cnt = np.empty(self.tp, np.int32)
        self.be.check(self.be.fn("get_neighbors")(self.ctx, which, _ptr(cnt, C.c_int32), None, 0),
                      "get_neighbors")
        width = max(int(cnt.max()) if self.tp else 0, 1)
        ids = np.full((self.tp, width), -1, np.int32)
        self.be.check(self.be.fn("get_neighbors")(self.ctx, which, _ptr(cnt, C.c_int32),
                                                  _ptr(ids, C.c_int32), width), "get_neighbors")
        return cnt, ids

    def cells(self, which=_abi.LIST_S):
        a = np.empty(self.tp, np.int32)
        self.be.check(self.be.fn("get_cells")(self.ctx, which, _ptr(a, C.c_int32)), "get_cells")
        return a

    def set_regions(self, region, boxes):
        """Batched scenarios: region[i] is particle i's scenario, boxes[k] the
        6 box bounds scenario k uses for BOUNDARIE and the PREPDATA check."""
        region = np.ascontiguousarray(region, np.int32)
        boxes = np.ascontiguousarray(boxes, np.float64).reshape(-1, 6)
        self._keep = (region, boxes)
        self.be.check(self.be.fn("set_regions")(self.ctx, boxes.shape[0], _ptr(region, C.c_int32),
                                                _ptr(boxes, C.c_double)), "set_regions")

    def autotune(self, reps=3):
        """Times both neighbour-search walks and keeps the faster per list,
        then the predictor and PPE product variants; returns the chosen walk
        (1 per particle, 2 per cell) for lists S, L (tuning() has all)."""
        m = np.zeros(2, np.int32)
        self.be.check(self.be.fn("autotune")(self.ctx, int(reps), _ptr(m, C.c_int32)), "autotune")
        return int(m[0]), int(m[1])

    def tuning(self):
        """(walk S, walk L, predictor variant, product variant)."""
        t = np.zeros(4, np.int32)
        self.be.check(self.be.fn("get_tuning")(self.ctx, _ptr(t, C.c_int32)), "get_tuning")
        return tuple(int(v) for v in t)

    def set_tuning(self, choice):
        t = np.ascontiguousarray(choice, dtype=np.int32)
        self.be.check(self.be.fn("set_tuning")(self.ctx, _ptr(t, C.c_int32)), "set_tuning")

    def attach_lpc(self, lpc):
        """Use a LearnedPreconditioner when the solver is the learned FCG
        (None detaches). The simulation keeps a reference to it."""
        self.be.check(self.be.fn("attach_lpc")(self.ctx, lpc.h if lpc is not None else None),
                      "attach_lpc")
        self._lpc = lpc

    def ppe(self):
        """Assembled PPE as CSR over the bcon == 0 rows: dict with rowptr, cols,
        vals (diagonal first in each row), rhs and the particle id of each row."""
        nu, nnz = C.c_int32(), C.c_int32()
        fn = self.be.fn("ppe_export")
        self.be.check(fn(self.ctx, C.byref(nu), C.byref(nnz), None, None, None, None, None),
                      "ppe_export")
        rowptr = np.empty(nu.value + 1, np.int32)
        cols = np.empty(nnz.value, np.int32)
        vals = np.empty(nnz.value, np.float64)
        rhs = np.empty(nu.value, np.float64)
        ids = np.empty(nu.value, np.int32)
        self.be.check(fn(self.ctx, C.byref(nu), C.byref(nnz), _ptr(rowptr, C.c_int32),
                         _ptr(cols, C.c_int32), _ptr(vals, C.c_double), _ptr(rhs, C.c_double),
                         _ptr(ids, C.c_int32)), "ppe_export")
        return {"rowptr": rowptr, "cols": cols, "vals": vals, "rhs": rhs, "ids": ids}


def _sim_write_vtu(self, path, fluid_only=False):
    """vox_write_vtu: every particle (or the fluid only) at full precision."""
    self.be.check(self.be.fn("write_vtu")(self.ctx, str(path).encode(), 1 if fluid_only else 0),
                  "write_vtu")


def _sim_load_ppe(self, system):
    """vox_ppe_load: a CSR system in the form ppe() returns replaces the
    assembled PPE (the next CGM solves it)."""
    nu = int(system["rhs"].shape[0])
    i32 = {k: np.ascontiguousarray(system[k], dtype=np.int32) for k in ("rowptr", "cols", "ids")}
    f64 = {k: np.ascontiguousarray(system[k], dtype=np.float64) for k in ("vals", "rhs")}
    self.be.check(self.be.fn("ppe_load")(
        self.ctx, nu, int(i32["rowptr"][-1]) if nu else 0, _ptr(i32["rowptr"], C.c_int32),
        _ptr(i32["cols"], C.c_int32), _ptr(f64["vals"], C.c_double), _ptr(f64["rhs"], C.c_double),
        _ptr(i32["ids"], C.c_int32)), "ppe_load")


Simulation.write_vtu = _sim_write_vtu
Simulation.load_ppe = _sim_load_ppe


def read_vtu(path, dim, l0, backend=None):
    """A Case from a VTU restart file (vox_vtu_count / vox_vtu_read, the
    reference's getTP / getVTU of inOut.cpp:25-102)."""
    from .cases import Case
    be = backend or hip_backend()
    n = C.c_int(0)
    be.check(be.lib.vox_vtu_count(str(path).encode(), C.byref(n)), "vtu_count")
    tp = n.value
    a = {k: np.zeros(tp) for k in ("x", "y", "z", "u", "v", "w", "p")}
    t = np.zeros(tp, np.int32)
    g, w, f = C.c_int(), C.c_int(), C.c_int()
    be.check(be.lib.vox_vtu_read(str(path).encode(), dim, tp, *(_ptr(a[k], C.c_double) for k in
                                                                ("x", "y", "z", "u", "v", "w", "p")),
                                 _ptr(t, C.c_int32), C.byref(g), C.byref(w), C.byref(f)), "vtu_read")
    return Case(dim, l0, a["x"], a["y"], a["z"], a["u"], a["v"], a["w"], a["p"], t,
                name=f"vtu:{path}", meta={"counts": (g.value, w.value, f.value)})
