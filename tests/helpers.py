"""Shared builders for the parity tests: small lattices, random states and
the mathematics the oracle is pinned against."""
import numpy as np

import voxarmps_amd as vx


def lattice(dim, l0, half, ptype=1):
    """A (2*half+1)^dim block of particles on a square lattice centred at the
    origin, all of one type (the centre particle has a full stencil when
    half*l0 > radius)."""
    r = np.arange(-half, half + 1)
    if dim == 2:
        I, J = np.meshgrid(r, r, indexing="ij")
        pts = np.stack([I.ravel(), J.ravel()], 1)
    else:
        I, J, K = np.meshgrid(r, r, r, indexing="ij")
        pts = np.stack([I.ravel(), J.ravel(), K.ravel()], 1)
    n = pts.shape[0]
    x = pts[:, 0] * l0
    y = pts[:, 1] * l0
    z = pts[:, 2] * l0 if dim == 3 else np.zeros(n)
    zeros = np.zeros(n)
    return vx.Case(dim, l0, x.astype(np.float64), y.astype(np.float64), z.astype(np.float64),
                   zeros.copy(), zeros.copy(), zeros.copy(), zeros.copy(),
                   np.full(n, ptype, np.int32), name=f"lattice{dim}d")


def centre_index(case):
    return int(np.argmin(case.x ** 2 + case.y ** 2 + case.z ** 2))


def lattice_offsets(dim, k):
    """Integer offsets (excluding 0) with |offset| < k, as exact integers."""
    m = int(np.ceil(k)) + 1
    r = np.arange(-m, m + 1)
    if dim == 2:
        A, B = np.meshgrid(r, r, indexing="ij")
        off = np.stack([A.ravel(), B.ravel()], 1)
    else:
        A, B, Cc = np.meshgrid(r, r, r, indexing="ij")
        off = np.stack([A.ravel(), B.ravel(), Cc.ravel()], 1)
    d2 = (off ** 2).sum(1)
    keep = (d2 > 0) & (d2 < k * k)
    return off[keep]


def n0_closed_form(dim, k):
    """sum over lattice offsets of W = k/|o| - 1 (ktype 2, radius k spacings)."""
    off = lattice_offsets(dim, k)
    r = np.sqrt((off ** 2).sum(1).astype(np.float64))
    return float(np.sum(k / r - 1.0))


def lambda_closed_form(dim, k):
    """sum r^2 W / sum W on the lattice (in units of the spacing squared)."""
    off = lattice_offsets(dim, k)
    r2 = (off ** 2).sum(1).astype(np.float64)
    w = k / np.sqrt(r2) - 1.0
    return float(np.sum(r2 * w) / np.sum(w))


def params_for(case, re_s_factor, re_l_factor=None, dt=1e-4):
    p = vx.default_params(case.dim, case.l0)
    p.re_s = re_s_factor * case.l0
    p.re_l = (re_l_factor or re_s_factor) * case.l0
    p.lambda_ = 0.16832 * p.re_l * p.re_l
    p.dt = dt
    return vx.set_box(p, case)


def random_state(sim, seed, with_pressure=True):
    """Smooth-ish random fields for stage-level parity tests."""
    rng = np.random.default_rng(seed)
    n = sim.tp
    for name, scale in (("U", 0.3), ("V", 0.3), ("W", 0.3), ("UNEW", 0.3), ("VNEW", 0.3),
                        ("WNEW", 0.3)):
        sim.set(name, rng.normal(0, scale, n))
    if with_pressure:
        sim.set("PNEW", rng.uniform(0, 3000, n))
        sim.set("PHAT", rng.uniform(0, 1000, n))
        sim.set("P", rng.uniform(0, 3000, n))
    return rng


class HipMem:
    """Device buffers through the HIP runtime that libvoxarmps itself loaded
    (found in this process's mappings, so both share one runtime instance)."""

    def __init__(self, backend):
        import ctypes as C
        path = None
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64.so" in line:
                    path = line.split()[-1]
                    break
        if path is None:
            raise RuntimeError("the HIP runtime is not loaded")
        self.C = C
        self.rt = C.CDLL(path)
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipDeviceSynchronize.argtypes = []
        self.live = []

    def put(self, arr):
        """Copy a contiguous numpy array to a new device buffer; returns its address."""
        C = self.C
        a = np.ascontiguousarray(arr)
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0      # host to device
        self.live.append(p.value)
        return p.value

    def get(self, ptr, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.rt.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.live:
            self.rt.hipFree(p)
        self.live = []


def slot_system(sim, rng=None):
    """The context's assembled PPE (after MATRIX) in the reference's 1-based
    slot layout of the CUDA path (functions.cu:1007-1060): neighb and poiss of
    (tp+1)*stride elements over list L, the diagonal in slot 1, the source and
    bcon of tp+1 elements. Neighbours with bcon == 1 get a coefficient as the
    reference writes one (the solve must treat them as p = 0); dummy
    neighbours and the particle itself get 0. Returns (stride, neighb, poiss,
    source, bcon, csr) with csr the exported system."""
    import scipy.sparse as sp
    tp = sim.tp
    cnt, ids = sim.neighbors(vx.LIST_L)
    stride = int(cnt.max()) + 2
    neighb = np.zeros((tp + 1) * stride, np.int32)
    poiss = np.zeros((tp + 1) * stride, np.float64)
    source = np.zeros(tp + 1, np.float64)
    bcon = np.zeros(tp + 1, np.int32)
    b = sim.get("BCON")
    bcon[1:] = b
    csr = sim.ppe()
    nu = len(csr["rhs"])
    A = sp.csr_matrix((csr["vals"], csr["cols"], csr["rowptr"]), shape=(nu, nu))
    uid = np.full(tp, -1, np.int64)
    uid[csr["ids"]] = np.arange(nu)
    rng = rng or np.random.default_rng(0)
    coo = A.tocoo()
    keys = coo.row.astype(np.int64) * nu + coo.col
    order = np.argsort(keys)
    keys, kval = keys[order], coo.data[order]
    rows = np.nonzero(uid >= 0)[0]
    base = (np.arange(tp) + 1) * stride
    neighb[base + 1] = cnt + 1
    lane = np.arange(ids.shape[1])[None, :]
    valid = lane < cnt[:, None]
    I = np.broadcast_to(np.arange(tp)[:, None], ids.shape)[valid]
    L = np.broadcast_to(lane, ids.shape)[valid]
    J = ids[valid].astype(np.int64)
    neighb[base[I] + 2 + L] = J + 1
    r = uid[rows]
    poiss[base[rows] + 1] = A.diagonal()[r]
    source[rows + 1] = csr["rhs"][r]
    m = (uid[I] >= 0) & (J != I) & (b[J] != -1)
    I, L, J = I[m], L[m], J[m]
    unk = uid[J] >= 0
    k = uid[I[unk]] * nu + uid[J[unk]]
    pos = np.clip(np.searchsorted(keys, k), 0, len(keys) - 1)
    hit = keys[pos] == k
    v = np.zeros(len(k))
    v[hit] = kval[pos[hit]]
    poiss[base[I[unk]] + 2 + L[unk]] = v
    ns = int(np.count_nonzero(~unk))
    poiss[base[I[~unk]] + 2 + L[~unk]] = -rng.uniform(1e3, 1e5, ns)
    return stride, neighb, poiss, source, bcon, {"A": A, "rhs": csr["rhs"], "ids": csr["ids"]}
