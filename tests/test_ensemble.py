"""Batched ensemble of independent 2D two-phase dam breaks in one context.

The batch is pinned to its own scenarios run alone: each scenario's
particles, taken from the batched run, must match the scenario's solo run.
Neighbour sets are identical in both (the tanks are further apart than any
radius); the order of equal-cell entries and the global PPE solve can differ,
so agreement is to the solver's tolerance, stated per test.
"""
import numpy as np
import pytest

import voxarmps_amd as vx


def test_ensemble_layout_invariants():
    case, region, boxes, scen = vx.ensemble_c4()
    assert len(scen) == 64 and boxes.shape == (64, 6)
    w, h = np.array(case.meta["widths"]), np.array(case.meta["heights"])
    assert w.min() >= 20 and w.max() <= 200 and h.min() >= 40 and h.max() <= 200
    assert sum(s.tp for s in scen) == case.tp
    # reference particle order over the whole batch: dummy -> wall -> fluid
    t = case.ptype
    assert np.all(np.diff(np.where(t < 0, 0, np.where(t == 0, 1, 2))) >= 0)
    # every scenario keeps its own particles, in its own relative order
    for k, s in enumerate(scen):
        m = region == k
        for t_sel in (lambda a: a < 0, lambda a: a == 0, lambda a: a > 0):
            a = t_sel(case.ptype) & m
            b = t_sel(s.ptype)
            assert np.array_equal(case.x[a], s.x[b]) and np.array_equal(case.y[a], s.y[b])
        assert np.all(case.x[m] > boxes[k, 0]) and np.all(case.x[m] < boxes[k, 1])
        # two-phase: both fluids present, fluid 2 above fluid 1
        f1, f2 = s.ptype == 1, s.ptype == 2
        assert f1.any() and f2.any() and s.y[f2].min() > s.y[f1].min()
    # tanks are further apart than twice the Laplacian radius (4 l0)
    for k in range(63):
        assert boxes[k + 1, 0] - boxes[k, 1] > 2 * 4.0 * case.l0


def _small_ensemble():
    return vx.ensemble_2d([20, 26, 31], [40, 47, 52])


@pytest.mark.gpu
def test_ensemble_matches_solo_runs(hip):
    """Three scenarios batched against each run alone, 30 steps of the
    two-phase ensemble configuration (SPS on, Bingham fluid 2, mixed precision,
    PPE to 1e-8): fluid positions within 1e-9 m and velocities within 1e-6 of
    their scale. Both runs meet the same relative residual; the batch's is
    relative to the whole batch's right-hand side."""
    case, region, boxes, scen = _small_ensemble()
    pb = vx.ensemble_params(case)
    b = vx.Simulation(pb, case, backend=hip)
    b.set_regions(region, boxes)
    b.init()
    b.step(30)
    st = b.status()
    assert st.nan_count == 0 and st.out_of_box == 0 and st.last_relres <= 1e-8
    for k, s in enumerate(scen):
        ps = vx.ensemble_params(s, box=boxes[k])
        a = vx.Simulation(ps, s, backend=hip)
        a.init()
        # n0 = max PND over the whole batch: the scenarios are lattices shifted
        # along x, so their own maxima differ in the last bits; the solo run
        # takes the batch's value
        assert np.allclose(a.n0(), b.n0(), rtol=1e-12, atol=0)
        a.set_n0(*b.n0())
        a.step(30)
        m = region == k
        f_b = m & (case.ptype > 0)
        f_a = s.ptype > 0
        dx = max(np.max(np.abs(b.get(n)[f_b] - a.get(n)[f_a])) for n in ("X", "Y"))
        vs = max(np.max(np.abs(a.get(n)[f_a])) for n in ("U", "V"))
        dv = max(np.max(np.abs(b.get(n)[f_b] - a.get(n)[f_a])) for n in ("U", "V"))
        print(f"scenario {k}: {s.tp} particles, max |dx| {dx:.2e} m, max |du| {dv:.2e} "
              f"(scale {vs:.2e})")
        assert dx <= 1e-9
        assert dv <= 1e-6 * max(vs, 1e-3)
        a.close()
    b.close()


@pytest.mark.gpu
def test_ensemble_scenario_fp64_against_oracle(hip, oracle):
    """One scenario of the ensemble configuration run alone in FP64 on the GPU
    and on the oracle: two fluids, SPS-LES on, Herschel-Bulkley (Bingham)
    fluid 2, two radii, PPE to 1e-10, 30 steps. Lists stay identical, and
    positions, velocities and pressure agree to 1e-6 of their scale (the
    solves agree to their tolerance; the Herschel-Bulkley viscosity's libm pow
    may differ in the last ulp)."""
    from voxarmps_amd._abi import PREC_FP64
    case, region, boxes, scen = _small_ensemble()
    s = scen[1]
    p = vx.ensemble_params(s, tol=1e-10, box=boxes[1])
    p.precision = PREC_FP64
    g = vx.Simulation(p, s, backend=hip)
    c = vx.Simulation(p, s, backend=oracle)
    g.init()
    c.init()
    assert g.n0() == c.n0()
    g.step(30)
    c.step(30)
    f = s.ptype > 0
    assert (s.ptype == 2).any()
    for n in ("X", "Y", "U", "V", "P", "MEU"):
        a, b = g.get(n)[f], c.get(n)[f]
        scale = max(np.max(np.abs(b)), 1.0)
        err = np.max(np.abs(a - b))
        print(f"{n}: max |gpu - oracle| {err:.2e} (scale {scale:.2e})")
        assert err <= 1e-6 * scale, n
    for w in (vx.LIST_S, vx.LIST_L):
        cg, ig = g.neighbors(w)
        cc, ic = c.neighbors(w)
        assert np.array_equal(cg, cc) and np.array_equal(ig, ic), f"list {w}"
    sg, sc = g.status(), c.status()
    assert sg.nan_count == sc.nan_count == 0
    g.close()
    c.close()


@pytest.mark.gpu
def test_autotune_keeps_lists(hip):
    """The autotuner only picks the faster search walk: the lists before and
    after it are identical."""
    case, region, boxes, scen = _small_ensemble()
    p = vx.ensemble_params(case)
    s = vx.Simulation(p, case, backend=hip)
    s.init()
    before = [s.neighbors(w) for w in (vx.LIST_S, vx.LIST_L)]
    modes = s.autotune(reps=2)
    assert all(m in (1, 2) for m in modes)
    s.NEIGHBOR()
    after = [s.neighbors(w) for w in (vx.LIST_S, vx.LIST_L)]
    for (c0, i0), (c1, i1) in zip(before, after):
        assert np.array_equal(c0, c1) and np.array_equal(i0, i1)
    s.close()


@pytest.mark.gpu
def test_ensemble_long_run_learned(hip):
    """The 64-scenario two-phase ensemble (mixed precision, SPS, Bingham fluid
    2) for its 2000 steps (dt 1e-4 s) with the learned flexible CG and the
    committed weights: every solve meets the tolerance (a failed learned solve
    falls back to Jacobi-CG and is counted), and no particle goes non-finite
    or leaves its box."""
    from voxarmps_amd.train_lpc import load_weights
    case, region, boxes, scen = vx.ensemble_c4()
    p = vx.ensemble_params(case)
    p.solver = vx._abi.SOLVER_LEARNED_FCG
    net = vx.LearnedPreconditioner(backend=hip)
    net.set_weights(load_weights())
    s = vx.Simulation(p, case, backend=hip)
    s.set_regions(region, boxes)
    s.attach_lpc(net)
    s.init()
    s.autotune(reps=1)
    iters = []
    for k in range(20):
        s.step(100)
        st = s.status()
        iters.append(st.last_iters)
        print(f"step {100 * (k + 1)}: iterations {st.last_iters}, relres {st.last_relres:.2e}, "
              f"fallbacks {st.lpc_fallbacks}, collisions {st.collisions}", flush=True)
        assert st.nan_count == 0 and st.out_of_box == 0
        assert st.last_relres <= p.cg_tol
    s.close()
    net.close()
