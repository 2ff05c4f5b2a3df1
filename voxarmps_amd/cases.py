"""Particle-case generation and the .grd text format.

The dam-break lattices are built on integer lattice coordinates and scaled by
the spacing l0, then sorted dummy (PTYPE -1) -> wall (0) -> fluid (1), the
order the stage index ranges of the reference rely on (main.cu:635-682).
Within each type, particles are numbered block by block (4^3 lattice blocks in
3D, 8x8 in 2D) so that the 64 consecutive ids of one wavefront are spatial
neighbours.

Boundary layers: every lattice point of the tank walls up to the tank height
(open top) within `wall_layers + dummy_layers` of the inner faces; the first
`wall_layers` are walls, the rest dummies.
"""
from dataclasses import dataclass, field

import numpy as np

SPLITMIX_GAMMA = np.uint64(0x9E3779B97F4A7C15)


@dataclass
class Case:
    dim: int
    l0: float
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    u: np.ndarray
    v: np.ndarray
    w: np.ndarray
    p: np.ndarray
    ptype: np.ndarray
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def tp(self):
        return int(self.x.shape[0])

    @property
    def gp(self):
        return int(np.count_nonzero(self.ptype < 0))

    @property
    def wp(self):
        return int(np.count_nonzero(self.ptype == 0))

    @property
    def fp(self):
        return int(np.count_nonzero(self.ptype > 0))

    def counts(self):
        return self.tp, self.gp, self.wp, self.fp

    def copy(self):
        return Case(self.dim, self.l0, self.x.copy(), self.y.copy(), self.z.copy(), self.u.copy(),
                    self.v.copy(), self.w.copy(), self.p.copy(), self.ptype.copy(), self.name,
                    dict(self.meta))


def splitmix64(seed, n):
    """First n outputs of SplitMix64 started at `seed` (uint64 array)."""
    k = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + k * SPLITMIX_GAMMA
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def uniform01(seed, n):
    """n doubles in [0, 1) from the top 53 bits of SplitMix64."""
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def jitter(case, amplitude, seed=0):
    """Add U[-a, a) to every coordinate of every particle; draws are taken in
    particle-id order, coordinates x, y(, z)."""
    out = case.copy()
    d = case.dim
    r = uniform01(seed, case.tp * d).reshape(case.tp, d)
    delta = (2.0 * r - 1.0) * amplitude
    out.x = out.x + delta[:, 0]
    out.y = out.y + delta[:, 1]
    if d == 3:
        out.z = out.z + delta[:, 2]
    out.meta["jitter"] = (amplitude, seed)
    return out


def _blocked_order(ij, block):
    """Permutation sorting lattice points by block (slowest axis last) then
    by position inside the block."""
    ij = np.asarray(ij)
    b = np.floor_divide(ij, block)
    r = ij - b * block
    keys = [r[:, a] for a in range(ij.shape[1])] + [b[:, a] for a in range(ij.shape[1])]
    return np.lexsort(keys)


def dam_break(dim, l0, water, tank_cells, wall_layers=1, dummy_layers=2, name="", x_offset=0):
    """Dam break: a water block of `water` particles in the corner of an open
    tank whose inner faces are at lattice 0 and `tank_cells` along each axis
    (shifted by `x_offset` lattice cells along x). y is vertical (gravity along
    -y)."""
    nl = wall_layers + dummy_layers
    if dim == 2:
        nx, ny = tank_cells
        I, J = np.meshgrid(np.arange(-(nl - 1), nx + nl), np.arange(-(nl - 1), ny + 1),
                           indexing="ij")
        I, J = I.ravel(), J.ravel()
        bnd = (I <= 0) | (I >= nx) | (J <= 0)
        depth = np.maximum.reduce([-I, I - nx, -J])
        bi = np.stack([I[bnd], J[bnd]], 1)
        bdepth = depth[bnd]
        wx, wy = water
        FI, FJ = np.meshgrid(np.arange(1, wx + 1), np.arange(1, wy + 1), indexing="ij")
        fi = np.stack([FI.ravel(), FJ.ravel()], 1)
        block = 8
    else:
        nx, ny, nz = tank_cells
        I, J, K = np.meshgrid(np.arange(-(nl - 1), nx + nl), np.arange(-(nl - 1), ny + 1),
                              np.arange(-(nl - 1), nz + nl), indexing="ij")
        I, J, K = I.ravel(), J.ravel(), K.ravel()
        bnd = (I <= 0) | (I >= nx) | (J <= 0) | (K <= 0) | (K >= nz)
        depth = np.maximum.reduce([-I, I - nx, -J, -K, K - nz])
        bi = np.stack([I[bnd], J[bnd], K[bnd]], 1)
        bdepth = depth[bnd]
        wx, wy, wz = water
        FI, FJ, FK = np.meshgrid(np.arange(1, wx + 1), np.arange(1, wy + 1), np.arange(1, wz + 1),
                                 indexing="ij")
        fi = np.stack([FI.ravel(), FJ.ravel(), FK.ravel()], 1)
        block = 4
    dummy = bi[bdepth >= wall_layers]
    wall = bi[bdepth < wall_layers]
    groups = []
    for pts, t in ((dummy, -1), (wall, 0), (fi, 1)):
        pts = pts[_blocked_order(pts, block)]
        groups.append((pts, t))
    lat = np.concatenate([g[0] for g in groups], 0)
    ptype = np.concatenate([np.full(len(g[0]), g[1], np.int32) for g in groups])
    n = lat.shape[0]
    x = (lat[:, 0] + x_offset).astype(np.float64) * l0
    y = lat[:, 1].astype(np.float64) * l0
    z = lat[:, 2].astype(np.float64) * l0 if dim == 3 else np.zeros(n)
    zeros = np.zeros(n)
    return Case(dim, l0, x, y, z, zeros.copy(), zeros.copy(), zeros.copy(), zeros.copy(), ptype,
                name=name, meta={"water": tuple(water), "tank_cells": tuple(tank_cells),
                                 "wall_layers": wall_layers, "dummy_layers": dummy_layers})


def dam_break_2d_c0():
    """2D dam break: l0 = 0.01 m, water 25 x 50 in a 1.6 x 0.6 m tank, one wall
    layer plus two dummy layers."""
    return dam_break(2, 0.01, (25, 50), (160, 60), 1, 2, name="dam2d_25x50_l0.01")


def dam_break_3d_c1(scale=1):
    """3D dam break: l0 = 0.005 m, water 100 (x) x 100 (y) x 80 (z) in a tank
    1.6 (x) x 0.6 (y) x 0.405 (z) m (81 spacings across z so that 80 fluid
    layers fit between the walls), one wall layer plus two dummy layers.
    `scale` > 1 shrinks every count by that factor (parity-test sizes)."""
    wx, wy, wz = 100 // scale, 100 // scale, 80 // scale
    return dam_break(3, 0.005, (wx, wy, wz), (320 // scale, 120 // scale, wz + 1), 1, 2,
                     name=f"dam3d_{wx}x{wy}x{wz}_l0.005")


def two_phase(case, fraction=0.5):
    """Fluid particles above `fraction` of the water column's height become
    fluid 2 (PTYPE 2): a layered two-phase column."""
    out = case.copy()
    f = out.ptype > 0
    y0, y1 = out.y[f].min(), out.y[f].max()
    out.ptype[f & (out.y > y0 + fraction * (y1 - y0))] = 2
    return out


def box_of(case):
    """Grid box of a case run on its own: bounding box +- l0 (main.cu:692-715)."""
    l0 = case.l0
    if case.dim == 3:
        return (case.x.min() - l0, case.x.max() + l0, case.y.min() - l0, case.y.max() + l0,
                case.z.min() - l0, case.z.max() + l0)
    return (case.x.min() - l0, case.x.max() + l0, case.y.min() - l0, case.y.max() + l0, 0.0, 0.0)


def batch_cases(scen, name="batch"):
    """Concatenate independent cases of one dimension, already placed apart
    from each other, into one particle set ordered dummy -> wall -> fluid over
    all of them, each keeping its own relative order. Returns (case, region,
    boxes): region[i] is particle i's case, boxes the 6 box bounds of each case
    run alone."""
    groups = []
    region = []
    for sel in (lambda t: t < 0, lambda t: t == 0, lambda t: t > 0):
        for k, c in enumerate(scen):
            m = sel(c.ptype)
            groups.append((c, m))
            region.append(np.full(int(m.sum()), k, np.int32))
    cat = lambda attr: np.concatenate([getattr(c, attr)[m] for c, m in groups])
    case = Case(scen[0].dim, scen[0].l0, cat("x"), cat("y"), cat("z"), cat("u"), cat("v"),
                cat("w"), cat("p"), cat("ptype").astype(np.int32), name=name)
    boxes = np.array([box_of(c) for c in scen], np.float64)
    return case, np.concatenate(region), boxes


def ensemble_2d(widths, heights, l0=0.01, gap_cells=16):
    """Independent 2D two-phase dam breaks laid out along x in one particle set.

    Scenario k is a water column of widths[k] x heights[k] particles (lower
    half fluid 1, upper half fluid 2) in an open tank 4 * width wide and
    height + 20 high, with one wall and two dummy layers. Tanks are placed
    `gap_cells` lattice cells apart (more than twice the largest neighbour
    radius, so no particle of one scenario neighbours another). Particles are
    ordered dummy -> wall -> fluid over all scenarios, each scenario keeping its
    own relative order. Returns (case, region, boxes, scenarios): region[i] is
    particle i's scenario, boxes the 6 box bounds of each scenario run alone,
    and scenarios the cases each scenario forms alone (same coordinates)."""
    scen = []
    off = 0
    for w, h in zip(widths, heights):
        w, h = int(w), int(h)
        tank = (4 * w, h + 20)
        c = two_phase(dam_break(2, l0, (w, h), tank, 1, 2, name=f"col{w}x{h}", x_offset=off))
        scen.append(c)
        off += tank[0] + 2 + gap_cells
    case, region, boxes = batch_cases(scen, name=f"ensemble{len(scen)}")
    case.meta.update({"widths": list(map(int, widths)), "heights": list(map(int, heights))})
    return case, region, boxes, scen


def jittered_dam_breaks_2d(seeds, amplitude=0.1, gap_cells=16):
    """The 2D dam break (l0 = 0.01, water 25 x 50, tank 160 x 60) once per
    seed, each copy jittered by U[-amplitude*l0, amplitude*l0) from its seed
    and placed `gap_cells` cells after the previous tank. Returns (case,
    region, boxes, scenarios) as ensemble_2d does."""
    scen = []
    for k, seed in enumerate(seeds):
        c = dam_break(2, 0.01, (25, 50), (160, 60), 1, 2, name=f"dam2d_seed{seed}",
                      x_offset=k * (160 + 2 + gap_cells))
        scen.append(jitter(c, amplitude * c.l0, seed=int(seed)))
    case, region, boxes = batch_cases(scen, name=f"dam2d_x{len(scen)}")
    return case, region, boxes, scen


def ensemble_c4(count=64, seed=0, l0=0.01):
    """The batched ensemble: `count` two-phase 2D dam breaks with column
    widths uniform in 20..200 and heights in 40..200 particles, drawn from
    SplitMix64(seed) as (width, height) per scenario in order."""
    r = uniform01(seed, 2 * count).reshape(count, 2)
    widths = 20 + np.floor(r[:, 0] * 181).astype(np.int64)
    heights = 40 + np.floor(r[:, 1] * 161).astype(np.int64)
    return ensemble_2d(widths, heights, l0)


def write_grd(path, case):
    """Full-precision .grd: TP, then x y [z] PTYPE u v [w] p per row."""
    with open(path, "w") as f:
        f.write(f"{case.tp}\n")
        if case.dim == 3:
            for k in range(case.tp):
                f.write(f"{case.x[k]:.17g} {case.y[k]:.17g} {case.z[k]:.17g} {int(case.ptype[k])} "
                        f"{case.u[k]:.17g} {case.v[k]:.17g} {case.w[k]:.17g} {case.p[k]:.17g}\n")
        else:
            for k in range(case.tp):
                f.write(f"{case.x[k]:.17g} {case.y[k]:.17g} {int(case.ptype[k])} "
                        f"{case.u[k]:.17g} {case.v[k]:.17g} {case.p[k]:.17g}\n")


def read_grd(path, dim):
    """Parse a .grd file (main.cu:496-682 layout) into a Case."""
    with open(path) as f:
        tp = int(f.readline().split()[0])
        rows = np.loadtxt(f, max_rows=tp, ndmin=2)
    if rows.shape[0] != tp:
        raise ValueError("row count differs from the header")
    if dim == 3:
        x, y, z, t, u, v, w, p = (rows[:, k] for k in range(8))
    else:
        x, y, t, u, v, p = (rows[:, k] for k in range(6))
        z = np.zeros(tp)
        w = np.zeros(tp)
    return Case(dim, 0.0, x.copy(), y.copy(), z.copy(), u.copy(), v.copy(), w.copy(), p.copy(),
                t.astype(np.int32), name=path)
