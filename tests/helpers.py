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
