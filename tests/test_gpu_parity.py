"""GPU parity of every stage against the CPU oracle, through the C ABI.

Bar: neighbour cells/lists, per-particle sums, the assembled PPE, the exact
collision replay and prepdata are bit-identical (integer results exactly,
FP64 results to the last bit: both sides evaluate the same IEEE operations in
the same order without contraction). Solver outputs agree to the solver
tolerance: |p_gpu - p_cpu| <= 1e-6 * max|p| for cg_tol 1e-10. Trajectories
agree to 1e-6 relative over the stated horizon.
"""
import ctypes as C

import numpy as np
import pytest

import voxarmps_amd as vx

from helpers import random_state

pytestmark = pytest.mark.gpu

FLOAT_OUT = {
    "PNUM": ["N"], "PNUMSTAR": ["NSTAR"], "SPS": ["NEUT", "TURB1", "TURB2", "TURB3"],
    "V_FRACTION": ["C"], "VISCOSITY": ["MEU"],
    "PREDICTION": ["XSTAR", "YSTAR", "ZSTAR", "USTAR", "VSTAR", "WSTAR"],
    "PHATCALC": ["PHAT"], "PRESSGRAD": ["UNEW", "VNEW", "WNEW"],
    "EULERINTEGRATION": ["X", "Y", "Z", "UNEW", "VNEW", "WNEW"],
    "COLLISION2": ["X", "Y", "Z", "UNEW", "VNEW", "WNEW"],
    "PREPDATA": ["P", "U", "V", "W"],
}


def _pair(case, p, hip, oracle):
    g = vx.Simulation(p, case, backend=hip)
    c = vx.Simulation(p, case, backend=oracle)
    return g, c


def _sync_fields(src, dst, names):
    for n in names:
        dst.set(n, src.get(n))


def _assert_same(g, c, names, what):
    for n in names:
        a, b = g.get(n), c.get(n)
        ok = np.array_equal(a, b, equal_nan=True) if a.dtype.kind == "f" else np.array_equal(a, b)
        if not ok:
            bad = np.nonzero(~((a == b) | (np.isnan(a) & np.isnan(b))))[0] if a.dtype.kind == "f" \
                else np.nonzero(a != b)[0]
            k = bad[0]
            raise AssertionError(f"{what}: {n} differs at {len(bad)} particles, first {k}: "
                                 f"gpu {a[k]!r} cpu {b[k]!r}")


def _assert_lists(g, c, what):
    for w in (vx.LIST_S, vx.LIST_L):
        assert np.array_equal(g.cells(w), c.cells(w)), f"{what}: cells {w}"
        cg, ig = g.neighbors(w)
        cc, ic = c.neighbors(w)
        assert np.array_equal(cg, cc), f"{what}: counts {w}"
        assert np.array_equal(ig, ic), f"{what}: ids {w}"


def _cases():
    c0 = vx.dam_break_2d_c0()
    yield "2d-lattice", c0, vx.dam2d_params(c0)
    j0 = vx.jitter(c0, 0.1 * 0.01, seed=0)
    yield "2d-jitter", j0, vx.dam2d_params(j0)
    c3 = vx.jitter(vx.dam_break(3, 0.005, (16, 16, 12), (48, 20, 13)), 0.1 * 0.005, seed=0)
    yield "3d-jitter", c3, vx.dam3d_params(c3)
    r3 = vx.dam_break(3, 0.01, (10, 12, 8), (24, 16, 9))
    p3 = vx.set_box(vx.default_params(3, 0.01), r3)      # reference single radius 3.2 DL
    yield "3d-ref-radius", r3, p3
    pg = vx.dam2d_params(j0)
    pg.ref_gpu = 1                                       # the reference's CUDA search
    yield "2d-ref-gpu", j0, pg


@pytest.mark.parametrize("mode", [0, 1, 2], ids=["auto", "per-particle", "per-cell"])
@pytest.mark.parametrize("name,case,p", list(_cases()), ids=lambda v: v if isinstance(v, str) else "")
def test_neighbor_lists_bit_exact(hip, oracle, name, case, p, mode):
    """Both search kernels (one thread per particle, one wave per 64 sorted particles) give
    the oracle's lists."""
    p = p.copy()
    p.list_mode = mode
    g, c = _pair(case, p, hip, oracle)
    g.NEIGHBOR()
    c.NEIGHBOR()
    _assert_lists(g, c, name)
    slot = np.zeros((case.tp + 1) * 400, np.int32)
    import ctypes as C
    assert hip.lib.vox_get_neighbors_slot(g.ctx, vx.LIST_S, slot.ctypes.data_as(C.POINTER(C.c_int32)),
                                          400) == 0
    slot = slot.reshape(case.tp + 1, 400)
    cnt, ids = c.neighbors(vx.LIST_S)
    i = case.tp // 2
    assert slot[i + 1, 1] == cnt[i] + 1
    assert np.array_equal(slot[i + 1, 2:2 + cnt[i]], ids[i, :cnt[i]] + 1)
    g.close()
    c.close()


def test_neighbor_lists_full_3d_jitter(hip, oracle):
    """Full-size 3D dam break (1,175,450 particles) plus U[-0.1 l0, 0.1 l0)
    jitter from SplitMix64 seed 0: both lists bit-exact."""
    case = vx.jitter(vx.dam_break_3d_c1(), 0.1 * 0.005, seed=0)
    p = vx.dam3d_params(case)
    g, c = _pair(case, p, hip, oracle)
    g.NEIGHBOR()
    c.NEIGHBOR()
    for w in (vx.LIST_S, vx.LIST_L):
        cg, ig = g.neighbors(w)
        cc, ic = c.neighbors(w)
        assert np.array_equal(cg, cc)
        assert np.array_equal(ig, ic)
        del ig, ic
    g.close()
    c.close()


def _stage_cases():
    j0 = vx.jitter(vx.dam_break_2d_c0(), 0.1 * 0.01, seed=1)
    yield "2d", j0, vx.dam2d_params(j0, tol=1e-10)
    c3 = vx.jitter(vx.dam_break(3, 0.005, (16, 16, 12), (48, 20, 13)), 0.1 * 0.005, seed=2)
    yield "3d", c3, vx.dam3d_params(c3, tol=1e-10)
    pg = vx.dam2d_params(j0, tol=1e-10)
    pg.ref_gpu = 1                                       # the reference's CUDA path
    yield "2d-ref-gpu", j0, pg


@pytest.mark.parametrize("name,case,p", list(_stage_cases())[:2], ids=["2d", "3d"])
def test_prediction_single_fluid_bit_exact(hip, oracle, name, case, p):
    """One fluid with Newtonian viscosity: every MEU is nu1 rho1, and the GPU
    predictor then skips the neighbour's type (the harmonic mean of equal
    values rounds back to the value). PREDICTION after VISCOSITY is still
    bit-exact, walls and dummies included; after MEU is uploaded by hand the
    general path runs and is bit-exact as well."""
    g, c = _pair(case, p, hip, oracle)
    g.init()
    c.init()
    for sim in (g, c):
        random_state(sim, seed=11)
    _sync_fields(c, g, ["U", "V", "W", "UNEW", "VNEW", "WNEW", "PNEW", "PHAT", "P"])
    for st in ("PNUM", "VISCOSITY", "PREDICTION"):
        getattr(g, st)()
        getattr(c, st)()
        _assert_same(g, c, FLOAT_OUT[st], f"{name} {st}")
    rng = np.random.default_rng(3)
    meu = rng.uniform(0.5e-3, 2e-3, case.tp)
    for sim in (g, c):
        sim.set("MEU", meu)
        sim.PREDICTION()
    _assert_same(g, c, FLOAT_OUT["PREDICTION"], f"{name} PREDICTION, MEU uploaded")
    g.close()
    c.close()


def test_compact_list_l_3d(hip, oracle):
    """3D list L feeds PREDICTION, PNUM on list L (n0) and MATRIX with only the
    entries of nonzero weight (d < re_l, the particle itself left out); the
    reference rows hold d <= re_l + DELTA and the particle itself. Consumers on
    the compact rows, never exported first, are bit-exact; an export then
    gives the reference rows; and after EULER moves the particles, PREDICTION
    and MATRIX on the list of the earlier positions (no NEIGHBOR in between)
    read the reference rows, as the reference does, and are bit-exact too."""
    case = vx.jitter(vx.dam_break(3, 0.005, (16, 16, 12), (48, 20, 13)), 0.1 * 0.005, seed=5)
    p = vx.dam3d_params(case, tol=1e-10)
    g, c = _pair(case, p, hip, oracle)
    g.init()
    c.init()
    assert g.n0() == c.n0(), "n0 on the compact list L"
    for sim in (g, c):
        random_state(sim, seed=13)
    _sync_fields(c, g, ["U", "V", "W", "UNEW", "VNEW", "WNEW", "PNEW", "PHAT", "P"])
    for sim in (g, c):
        sim.NEIGHBOR()
    for st in ("PNUM", "VISCOSITY", "PREDICTION", "PNUMSTAR"):
        getattr(g, st)()
        getattr(c, st)()
        _assert_same(g, c, FLOAT_OUT[st], f"compact {st}")
    for sim in (g, c):
        sim.BCON()
        sim.MATRIX()
    _assert_same(g, c, ["SOURCE"], "compact MATRIX source")
    Ag, Ac = g.ppe(), c.ppe()
    for k in ("rowptr", "cols", "vals", "rhs", "ids"):
        assert np.array_equal(Ag[k], Ac[k]), f"compact MATRIX {k}"
    _assert_lists(g, c, "export after compact consumers")
    # a fresh compact list, then motion without a new search
    for sim in (g, c):
        sim.NEIGHBOR()
        sim.PREDICTION()
        sim.set("UNEW", sim.get("U") + 0.5)
        sim.EULERINTEGRATION()
        sim.PREDICTION()
        sim.MATRIX()
    _assert_same(g, c, FLOAT_OUT["PREDICTION"], "stale list PREDICTION")
    Ag, Ac = g.ppe(), c.ppe()
    for k in ("rowptr", "cols", "vals", "rhs", "ids"):
        assert np.array_equal(Ag[k], Ac[k]), f"stale list MATRIX {k}"
    g.close()
    c.close()


@pytest.mark.parametrize("name,case,p", list(_stage_cases()), ids=lambda v: v if isinstance(v, str) else "")
def test_stages_bit_exact(hip, oracle, name, case, p):
    case = case.copy()
    # a second fluid in the upper half exercises the two-phase branches
    f = case.ptype > 0
    case.ptype[f & (case.y > np.median(case.y[f]))] = 2
    p = p.copy()
    p.turb = 1
    g, c = _pair(case, p, hip, oracle)
    g.init()
    c.init()
    assert g.n0() == c.n0(), "n0"
    _assert_lists(g, c, f"{name} init")
    for sim in (g, c):
        random_state(sim, seed=7)
    _sync_fields(c, g, ["U", "V", "W", "UNEW", "VNEW", "WNEW", "PNEW", "PHAT", "P"])
    stages = ["PNUM", "SPS", "V_FRACTION", "VISCOSITY", "PREDICTION", "PNUMSTAR"]
    for st in stages:
        getattr(g, st)()
        getattr(c, st)()
        _assert_same(g, c, FLOAT_OUT[st], f"{name} {st}")
    # Herschel-Bulkley branch (pow: last-ulp differences allowed)
    q = p.copy()
    q.fluid2_type = 1
    q.cohes = 1.0
    q.n_pl = 1.0
    q.meu0 = 0.85
    for sim in (g, c):
        sim.set_params(q)
        sim.VISCOSITY()
    mg, mc = g.get("MEU"), c.get("MEU")
    assert np.allclose(mg, mc, rtol=1e-13, atol=0)
    for sim in (g, c):
        sim.set_params(p)
    g.BCON()
    c.BCON()
    _assert_same(g, c, ["BCON"], f"{name} BCON")
    g.MATRIX()
    c.MATRIX()
    _assert_same(g, c, ["SOURCE"], f"{name} MATRIX source")
    Ag, Ac = g.ppe(), c.ppe()
    for k in ("rowptr", "cols", "vals", "rhs", "ids"):
        assert np.array_equal(Ag[k], Ac[k]), f"{name} MATRIX {k}"
    # MATRIX options (functions.cu:901-914, 928-988): higher-order Laplacian
    # and the higher-order, error-compensated source, bit-exact as well
    for opt in ("matopt", "srcopt"):
        q = p.copy()
        setattr(q, opt, 1)
        for sim in (g, c):
            sim.set_params(q)
            sim.MATRIX()
        _assert_same(g, c, ["SOURCE"], f"{name} MATRIX {opt} source")
        Ag, Ac = g.ppe(), c.ppe()
        for k in ("rowptr", "cols", "vals", "rhs", "ids"):
            assert np.array_equal(Ag[k], Ac[k]), f"{name} MATRIX {opt} {k}"
    src = g.get("SOURCE")
    assert np.count_nonzero(src) > 0
    for sim in (g, c):
        sim.set_params(p)
        sim.MATRIX()
    ig = g.CGM()
    ic = c.CGM()
    assert ig.converged and ic.converged
    assert ig.unknowns == ic.unknowns and ig.nnz == ic.nnz
    assert abs(ig.iterations - ic.iterations) <= 1
    pg, pc = g.get("PNEW"), c.get("PNEW")
    assert np.max(np.abs(pg - pc)) <= 1e-6 * max(np.max(np.abs(pc)), 1.0)
    # the reference's own solver on the same (bit-identical) system: IC(0)-CG
    # with its stop rule (functions.cu:506-716, 1139-1200; imin 10, imax 50,
    # eps 1e-9, main.cu:229-230)
    info = vx.VoxSolveInfo()
    assert oracle.lib.orc_ppe_solve_iccg(c.ctx, 10, 50, 1e-9, C.byref(info)) == 0
    pi = c.get("PNEW")
    assert info.converged
    assert np.max(np.abs(pg - pi)) <= 1e-8 * max(np.max(np.abs(pi)), 1.0), f"{name} vs ICCG"
    # PRESSURECALC method 3, the equation of state on n* (functions.cu:2471-2500):
    # pow(n*/n0, 7) may differ in its last bits between the device and host
    # libraries; (n*/n0)^7 - 1 cancels near n* = n0, so the bound is absolute,
    # 16 ulps of c0^2 rho / 7
    q = p.copy()
    q.method = 3
    nstar = c.get("NSTAR")
    ratio = np.random.default_rng(11).uniform(0.85, 1.05, case.tp)   # both sides of 0.92
    for sim in (g, c):
        sim.set("NSTAR", ratio * c.n0()[0])
        sim.set_params(q)
        sim.PRESSURECALC()
    pg, pc = g.get("PNEW"), c.get("PNEW")
    coef = max(q.c01, q.c02) ** 2 * max(q.rho1, q.rho2) / 7.0
    assert np.max(np.abs(pg - pc)) <= 16 * np.finfo(float).eps * coef
    assert np.count_nonzero(pc[case.ptype >= 0]) > 0
    for sim in (g, c):
        sim.set("NSTAR", nstar)
        sim.set_params(p)
    # downstream stages from identical pressure
    _sync_fields(c, g, ["PNEW"])
    for st in ("PHATCALC", "PRESSGRAD", "EULERINTEGRATION"):
        getattr(g, st)()
        getattr(c, st)()
        _assert_same(g, c, FLOAT_OUT[st], f"{name} {st}")
    g.COLLISION2()
    c.COLLISION2()
    _assert_same(g, c, FLOAT_OUT["COLLISION2"], f"{name} COLLISION2")
    assert g.status().collisions == c.status().collisions
    g.PREPDATA()
    c.PREPDATA()
    _assert_same(g, c, FLOAT_OUT["PREPDATA"], f"{name} PREPDATA")
    sg, sc = g.status(), c.status()
    assert (sg.nan_count, sg.out_of_box) == (sc.nan_count, sc.out_of_box)
    g.close()
    c.close()


def _crowded_case(seed, n=600, dim=2):
    """Fluid particles packed at ~0.55 l0 spacing with random velocities:
    many pairs closer than the collision distance."""
    rng = np.random.default_rng(seed)
    l0 = 0.01
    side = int(np.ceil(n ** (1 / dim)))
    g = np.stack(np.meshgrid(*[np.arange(side)] * dim, indexing="ij"), -1).reshape(-1, dim)[:n]
    pos = 0.1 + g * 0.55 * l0 + rng.uniform(-0.1, 0.1, g.shape) * l0
    x, y = pos[:, 0], pos[:, 1]
    z = pos[:, 2] if dim == 3 else np.zeros(n)
    zero = np.zeros(n)
    case = vx.Case(dim, l0, x, y, z, zero.copy(), zero.copy(), zero.copy(), zero.copy(),
                   np.ones(n, np.int32))
    p = vx.set_box(vx.default_params(dim, l0), case)
    return case, p, rng


@pytest.mark.parametrize("dim,vscale", [(2, 0.5), (3, 0.5), (2, 400.0)])
def test_collision_replay_bit_exact(hip, oracle, dim, vscale):
    """Order-dependent in-place collision: the GPU replay equals the
    sequential reference loop bit for bit. vscale 400 moves particles far
    enough during the pass to force the widened-screen retry."""
    case, p, rng = _crowded_case(11 + dim, dim=dim)
    p.dt = 1e-4
    g, c = _pair(case, p, hip, oracle)
    g.init()
    c.init()
    vel = {n: rng.normal(0, vscale, case.tp) for n in ("UNEW", "VNEW", "WNEW")}
    for sim in (g, c):
        for n, v in vel.items():
            sim.set(n, v)
        sim.COLLISION2()
    _assert_same(g, c, FLOAT_OUT["COLLISION2"], f"collision d{dim}")
    assert g.status().collisions == c.status().collisions > 10
    g.close()
    c.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_collision_stale_list_widened_screen(hip, oracle, dim):
    """COLLISION2 runs on list S built at the step's earlier positions
    (main.cu:1050 vs 1306). Here the list is built, the particles are moved by
    up to 0.5 l0 and given velocities large enough that the screen widens
    past the list radius: the replay then takes every listed pair as a
    candidate and still equals the sequential reference loop bit for bit."""
    case, p, rng = _crowded_case(31 + dim, n=500, dim=dim)
    p = vx.two_radius(p)
    p.dt = 1e-4
    g, c = _pair(case, p, hip, oracle)
    g.init()
    c.init()
    l0 = case.l0
    moved = {n: getattr(case, n.lower()) + rng.uniform(-0.5, 0.5, case.tp) * l0
             for n in (("X", "Y", "Z") if dim == 3 else ("X", "Y"))}
    vel = {n: rng.normal(0, 400.0, case.tp) for n in ("UNEW", "VNEW", "WNEW")}
    for sim in (g, c):
        for n, v in {**moved, **vel}.items():
            sim.set(n, v)
        sim.COLLISION2()
    _assert_same(g, c, FLOAT_OUT["COLLISION2"], f"collision stale list d{dim}")
    cnt, ids = c.neighbors(vx.LIST_S)
    listed = int(cnt.sum()) - (case.tp if dim == 3 else 0)      # 3D rows list the particle itself
    sg = g.status()
    print(f"d{dim}: collisions {sg.collisions}, candidates {sg.collision_cand} of {listed} listed")
    assert sg.collisions == c.status().collisions > 10
    assert sg.collision_cand == listed                           # the widened screen took every pair
    g.close()
    c.close()


def _sparse_pairs_case(seed, dim, side):
    """A lattice at l0 with U(-0.36 l0, 0.36 l0) jitter: scattered close pairs,
    so the collision candidates form many small components."""
    rng = np.random.default_rng(seed)
    l0 = 0.01
    g = np.stack(np.meshgrid(*[np.arange(side)] * dim, indexing="ij"), -1).reshape(-1, dim)
    pos = 0.1 + g * l0 + rng.uniform(-0.36, 0.36, g.shape) * l0
    n = pos.shape[0]
    x, y = pos[:, 0], pos[:, 1]
    z = pos[:, 2] if dim == 3 else np.zeros(n)
    zero = np.zeros(n)
    case = vx.Case(dim, l0, x, y, z, zero.copy(), zero.copy(), zero.copy(), zero.copy(),
                   np.ones(n, np.int32))
    p = vx.set_box(vx.default_params(dim, l0), case)
    return case, p, rng


@pytest.mark.parametrize("dim,side,vscale", [(2, 60, 2.0), (3, 16, 2.0), (3, 16, 60.0)])
def test_collision_components_bit_exact(hip, oracle, dim, side, vscale):
    """Scattered close pairs: the replay runs one thread per candidate
    component and must equal the sequential reference loop bit for bit; the
    fast case (vscale 60) also forces the widened-screen retry."""
    case, p, rng = _sparse_pairs_case(21 + dim, dim, side)
    p.dt = 1e-4
    g, c = _pair(case, p, hip, oracle)
    g.init()
    c.init()
    vel = {n: rng.normal(0, vscale, case.tp) for n in ("UNEW", "VNEW", "WNEW")}
    for sim in (g, c):
        for n, v in vel.items():
            sim.set(n, v)
        sim.COLLISION2()
    _assert_same(g, c, FLOAT_OUT["COLLISION2"], f"collision components d{dim}")
    sg, sc = g.status(), c.status()
    assert sg.collisions == sc.collisions > 20
    assert sg.collision_cand > 0
    g.close()
    c.close()


@pytest.mark.parametrize("nsteps", [40])
def test_trajectory_2d(hip, oracle, nsteps):
    case = vx.jitter(vx.dam_break_2d_c0(), 0.05 * 0.01, seed=4)
    p = vx.dam2d_params(case, tol=1e-10)
    g, c = _pair(case, p, hip, oracle)
    g.init()
    c.init()
    g.step(nsteps)
    c.step(nsteps)
    for n in ("X", "Y", "U", "V", "P"):
        a, b = g.get(n), c.get(n)
        scale = max(np.max(np.abs(b)), 1.0)
        assert np.max(np.abs(a - b)) <= 1e-6 * scale, n
    _assert_lists(g, c, "trajectory end")
    g.close()
    c.close()


def test_dtcalculation_steps(hip, oracle):
    """dt_mode = 1: every step starts with DTcalculation (main.cu:1030,
    functions.cu:1684-1698). The run starts from a DT above the Courant
    limit, so the first step lowers it to COURANT DL / max(c01, c02); the
    GPU and the oracle hold the same DT and simulated time, and the fields
    agree as in the fixed-step trajectory."""
    case = vx.jitter(vx.dam_break_2d_c0(), 0.05 * 0.01, seed=6)
    p = vx.dam2d_params(case, tol=1e-10)
    p.dt_mode, p.courant, p.c01, p.c02, p.dt, p.dt_max = 1, 0.04, 15.0, 25.0, 1e-3, 1.0
    courant = 0.04 * p.dl / 25.0
    g, c = _pair(case, p, hip, oracle)
    g.init()
    c.init()
    g.step(12)
    c.step(12)
    pg, pc = g.get_params(), c.get_params()
    assert pg.dt == pc.dt == courant
    assert g.status().time == c.status().time
    assert abs(g.status().time - 12 * courant) <= 1e-15
    for n in ("X", "Y", "U", "V", "P"):
        a, b = g.get(n), c.get(n)
        scale = max(np.max(np.abs(b)), 1.0)
        assert np.max(np.abs(a - b)) <= 1e-6 * scale, n
    _assert_lists(g, c, "dt_mode 1")
    g.close()
    c.close()


def test_trajectory_2d_ref_gpu(hip, oracle):
    """The reference's CUDA path (ref_gpu): 40 steps of the jittered 2D dam
    break against the oracle's restatement of that path."""
    case = vx.jitter(vx.dam_break_2d_c0(), 0.05 * 0.01, seed=4)
    p = vx.dam2d_params(case, tol=1e-10)
    p.ref_gpu = 1
    g, c = _pair(case, p, hip, oracle)
    g.init()
    c.init()
    g.step(40)
    c.step(40)
    for n in ("X", "Y", "U", "V", "P"):
        a, b = g.get(n), c.get(n)
        scale = max(np.max(np.abs(b)), 1.0)
        assert np.max(np.abs(a - b)) <= 1e-6 * scale, n
    _assert_lists(g, c, "trajectory end")
    cnt, ids = g.neighbors(vx.LIST_S)
    assert all(i in ids[i, :cnt[i]] for i in range(0, case.tp, 97))   # the particle itself
    g.close()
    c.close()


def test_trajectory_weakly_compressible(hip, oracle):
    """PRESSURECALC method 3 (equation of state, no PPE) with its parameter set
    (cubic kernel, relaxp 0.5, collision coefficient 0.5): 20 steps of the 2D
    dam break. Only pow() can differ (last bits), so the trajectories agree to
    1e-9 of each field's scale."""
    case = vx.dam_break_2d_c0()
    p = vx.wc_params(case)
    g, c = _pair(case, p, hip, oracle)
    g.init()
    c.init()
    g.step(20)
    c.step(20)
    for n in ("X", "Y", "U", "V", "PNEW"):
        a, b = g.get(n), c.get(n)
        scale = max(np.max(np.abs(b)), 1.0)
        assert np.max(np.abs(a - b)) <= 1e-9 * scale, n
    assert np.max(c.get("PNEW")) > 0
    _assert_lists(g, c, "trajectory end")
    g.close()
    c.close()


def test_trajectory_3d(hip, oracle):
    case = vx.dam_break(3, 0.005, (16, 16, 12), (48, 20, 13))
    p = vx.dam3d_params(case, tol=1e-10)
    g, c = _pair(case, p, hip, oracle)
    g.init()
    c.init()
    g.step(8)
    c.step(8)
    for n in ("X", "Y", "Z", "U", "V", "W", "P"):
        a, b = g.get(n), c.get(n)
        scale = max(np.max(np.abs(b)), 1.0)
        assert np.max(np.abs(a - b)) <= 1e-6 * scale, n
    g.close()
    c.close()


def test_gpu_runs_are_deterministic(hip):
    case = vx.jitter(vx.dam_break(3, 0.005, (16, 16, 12), (48, 20, 13)), 0.0005, seed=9)
    p = vx.dam3d_params(case)
    out = []
    for _ in range(2):
        with vx.Simulation(p, case, backend=hip) as g:
            g.init()
            g.step(5)
            out.append([g.get(n) for n in ("X", "Y", "Z", "P", "U")])
    for a, b in zip(*out):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("caps", [(4, 8), (0, 8), (4, 0)], ids=["both", "list-L", "list-S"])
@pytest.mark.parametrize("dim,mode", [(2, 1), (2, 2), (3, 2)])
def test_row_capacity_grows(hip, oracle, dim, mode, caps):
    """Rows that overflow a too-small capacity are detected, storage grows and
    the search repeats, in both search kernels, for either list alone or both
    (the two lists are built on two streams)."""
    if dim == 2:
        case = vx.jitter(vx.dam_break_2d_c0(), 0.001, seed=2)
        p = vx.dam2d_params(case)
    else:
        case = vx.jitter(vx.dam_break(3, 0.005, (16, 16, 12), (48, 20, 13)), 0.0005, seed=3)
        p = vx.dam3d_params(case)
    p.cap_s, p.cap_l = caps
    p.list_mode = mode
    g, c = _pair(case, p, hip, oracle)
    g.NEIGHBOR()
    c.NEIGHBOR()
    _assert_lists(g, c, "grown rows")
    g.close()
    c.close()


def test_ppe_columns_past_16_bit_steps(hip, oracle):
    """A 2D water column 17,000 particles tall: one cell column of list L's grid
    holds more than 65,535 unknowns, so the step from a row's last column in
    one x-column of cells to its first in the next does not fit 16 bits. The
    assembly then stores the 32-bit columns (a second pass), the product
    reads them, and the system is still the oracle's, bit for bit."""
    case = vx.dam_break(2, 0.01, (8, 17000), (9, 17000))
    p = vx.dam2d_params(case)
    p.cg_maxit = 20
    g, c = _pair(case, p, hip, oracle)
    for sim in (g, c):
        sim.init()
        random_state(sim, seed=11)
    _sync_fields(c, g, ["U", "V", "W", "UNEW", "VNEW", "WNEW", "PNEW", "PHAT", "P"])
    for sim in (g, c):
        for st in ("PNUM", "NEIGHBOR", "V_FRACTION", "VISCOSITY", "PREDICTION", "PNUMSTAR",
                   "BCON", "MATRIX"):
            getattr(sim, st)()
    Ag, Ac = g.ppe(), c.ppe()
    assert Ag["rhs"].shape[0] > 2 * 65536 and np.count_nonzero(Ag["rhs"]) > 0
    for k in ("rowptr", "cols", "vals", "rhs", "ids"):
        assert np.array_equal(Ag[k], Ac[k]), f"tall column MATRIX {k}"
    try:
        g.CGM()
    except vx.VoxError:
        pass                                             # 20 iterations need not converge
    st = g.status()
    assert st.product_col_bytes == 4 and st.product_window_cols == 0
    g.close()
    c.close()


def test_domain_error_is_reported(hip):
    case = vx.dam_break_2d_c0()
    p = vx.dam2d_params(case)
    with vx.Simulation(p, case, backend=hip) as g:
        x = case.x.copy()
        x[5] = p.xmax + 1.0
        g.set("X", x)
        with pytest.raises(vx.VoxError) as e:
            g.NEIGHBOR()
        assert e.value.code == vx._abi.VOX_ERR_DOMAIN


def test_step_before_init_is_refused(hip):
    case = vx.dam_break_2d_c0()
    with vx.Simulation(vx.dam2d_params(case), case, backend=hip) as g:
        with pytest.raises(vx.VoxError) as e:
            g.step(1)
        assert e.value.code == vx._abi.VOX_ERR_STATE


def _bulk(sim, case):
    x, y, P = sim.get("X"), sim.get("Y"), sim.get("P")
    f = case.ptype > 0
    bottom = f & (y < 0.1) & (x < 0.25)
    return x[f].max(), y[f].max(), P[bottom].mean()


def test_long_run_bulk_metrics_2d(hip, oracle):
    """The 2D dam break (l0 = 0.01, water 25 x 50, dt 1e-4, PPE to 1e-9) over
    its 1000 steps on the GPU and on the oracle. Collisions and the CG
    tolerance let single particles drift apart, and the MPS pressure of a
    region swings by a factor of 3 from step to step, so bulk metrics are
    compared: the surge front (largest fluid x) and the column height to one
    spacing every 100 steps, and the bottom-left probe pressure (mean over the
    fluid in 0.25 x 0.1 m) averaged over the run, sampled every 10 steps, to 5%."""
    case = vx.dam_break_2d_c0()
    p = vx.dam2d_params(case)
    g, c = _pair(case, p, hip, oracle)
    g.init()
    c.init()
    pg, pc = [], []
    for k in range(1, 101):
        g.step(10)
        c.step(10)
        mg, mc = _bulk(g, case), _bulk(c, case)
        pg.append(mg[2])
        pc.append(mc[2])
        if k % 10 == 0:
            print(f"step {10 * k}: front {mg[0]:.4f} / {mc[0]:.4f} m, height {mg[1]:.4f} / {mc[1]:.4f} m, "
                  f"bottom p {mg[2]:.0f} / {mc[2]:.0f} Pa", flush=True)
            assert abs(mg[0] - mc[0]) <= 0.01 and abs(mg[1] - mc[1]) <= 0.01
    a, b = np.mean(pg), np.mean(pc)
    print(f"probe pressure averaged over the run: {a:.1f} / {b:.1f} Pa")
    assert abs(a - b) <= 0.05 * abs(b)
    assert g.status().nan_count == 0 and g.status().out_of_box == 0
    g.close()
    c.close()
