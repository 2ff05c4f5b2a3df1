"""Learned preconditioner on the GPU against its float64 CPU restatement
(oracle/lpc.py), and flexible CG against Jacobi-CG on the same systems.

Tolerances: features are sums over about 30 (2D) to 250 (3D) entries,
accumulated in float32 here and in float64 by the oracle (absolute 1e-5 on
features of order one); the network runs bf16 operands with float32 accumulation
where the oracle accumulates in float64 before the same bf16 roundings, so an
accumulation that lands on the other side of a bf16 rounding boundary moves an
activation by one bf16 ulp (2^-8): outputs g, c agree to 1e-3 in the median
and 3e-2 at worst, the loss to 1e-3 and the gradient to 3e-2 in L2."""
import numpy as np
import pytest

import voxarmps_amd as vx
from oracle import lpc as olpc

pytestmark = pytest.mark.gpu

STAGES = ("PNUM", "NEIGHBOR", "SPS", "V_FRACTION", "VISCOSITY", "PREDICTION", "PNUMSTAR", "BCON",
          "MATRIX")


def assembled_2d(hip, steps=20, seed=None, solver=vx._abi.SOLVER_JACOBI_CG,
                 precision=vx._abi.PREC_FP64):
    """2D dam break (l0 = 0.01, water 25 x 50) after `steps` steps, with the
    next step's PPE assembled; optionally jittered by 0.1 l0 (with the
    unjittered lattice's n0: the jittered maximum PND would put most
    particles below the free-surface threshold)."""
    case = vx.dam_break_2d_c0()
    n0 = None
    if seed is not None:
        ref = vx.Simulation(vx.dam2d_params(case), case, backend=hip)
        ref.init()
        n0 = ref.n0()
        ref.close()
        case = vx.jitter(case, 0.1 * case.l0, seed=seed)
    p = vx.dam2d_params(case, tol=1e-8)
    p.solver = solver
    p.precision = precision
    s = vx.Simulation(p, case, backend=hip)
    s.init()
    if n0 is not None:
        s.set_n0(*n0)
    if steps:
        s.step(steps)
    for st in STAGES:
        getattr(s, st)()
    return s, case, p


def oracle_system(s, case, p):
    P = s.ppe()
    thr = olpc.strong_threshold(case.dim, p.lambda_, p.re_l, p.dl, s.n0()[1], p.rho1)
    S = olpc.System(P["rowptr"], P["cols"], P["vals"], P["rhs"], P["ids"], s.get("X"), s.get("Y"),
                    s.get("Z"), s.get("N"), s.n0()[1], case.ptype, p.re_l, case.dim, thr)
    return S, P


def random_weights(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    w = rng.uniform(-scale, scale, olpc.NPARAM)
    W = olpc.unpack(w)
    for k in ("b0", "c1", "c2", "c3"):
        W[k][:] = np.abs(W[k]) + 0.1
    return w.astype(np.float32)


@pytest.fixture(scope="module")
def sys2d(hip):
    s, case, p = assembled_2d(hip)
    S, P = oracle_system(s, case, p)
    net = vx.LearnedPreconditioner(seed=1, backend=hip)
    idx = net.dump(s)
    yield net, idx, S, P, s
    net.close()
    s.close()


def test_features_match_oracle(sys2d):
    net, idx, S, _, _ = sys2d
    f, _, _ = net.eval(idx, S.nu)
    assert np.all(np.isfinite(f))
    assert np.max(np.abs(f - S.f)) <= 1e-5 * max(1.0, np.max(np.abs(S.f)))
    # the wall flag and the structural zeros are exact
    assert np.array_equal(f[:, 1], S.f[:, 1]) and np.all(f[:, 5:] == 0)


def test_forward_matches_oracle(sys2d):
    net, idx, S, _, _ = sys2d
    w = random_weights(3)
    net.set_weights(w)
    assert np.array_equal(net.weights(), w)
    _, g, c = net.eval(idx, S.nu)
    st = olpc.forward(w.astype(np.float64), S)
    for got, want in ((g, st["g"]), (c, st["c"])):
        rel = np.abs(got - want) / np.abs(want)
        print(f"median {np.median(rel):.2e}, max {rel.max():.2e}")
        assert np.median(rel) <= 1e-3 and rel.max() <= 3e-2
    # the outputs really vary over the rows (a constant would hide layout errors)
    assert np.std(st["g"]) > 1e-3 and np.std(st["c"]) > 1e-3


def test_zero_network_outputs_one(sys2d):
    net, idx, S, _, _ = sys2d
    net.set_weights(np.zeros(vx.lpc.NPARAM, np.float32))
    _, g, c = net.eval(idx, S.nu)
    assert np.all(g == 1.0) and np.allclose(c, 1.0, rtol=1e-6)


def test_gradient_matches_oracle(sys2d):
    net, idx, S, _, _ = sys2d
    w = random_weights(5)
    net.set_weights(w)
    seed = 12345
    L, g = net.loss_grad(idx, seed)
    Lo, go = olpc.loss_grad(w.astype(np.float64), S, S.probes(seed), rounding=True)
    print(f"loss {L:.6e} oracle {Lo:.6e}")
    assert L == pytest.approx(Lo, rel=1e-3)
    err = np.linalg.norm(g - go) / np.linalg.norm(go)
    print(f"gradient relative L2 error {err:.2e}")
    assert err <= 3e-2
    # every parameter block carries gradient and agrees on its own
    k = 0
    for name, shape in olpc.LAYOUT:
        n = int(np.prod(shape))
        a, b = g[k:k + n], go[k:k + n]
        k += n
        assert np.linalg.norm(b) > 0, name
        assert np.linalg.norm(a - b) <= 0.1 * np.linalg.norm(b), name


def _residual(P, p_unknown):
    nu = len(P["rhs"])
    A = np.zeros(nu)
    for r in range(nu):
        a, b = P["rowptr"][r], P["rowptr"][r + 1]
        A[r] = P["vals"][a:b] @ p_unknown[P["cols"][a:b]]
    return np.linalg.norm(P["rhs"] - A) / np.linalg.norm(P["rhs"])


@pytest.mark.parametrize("precision", [vx._abi.PREC_FP64, vx._abi.PREC_MIXED])
def test_fcg_matches_jacobi(hip, precision):
    """The same assembled system solved by Jacobi-CG and by flexible CG with a
    (random) network: both reach rel-res 1e-8 on the FP64 system, recomputed
    here from the exported matrix, and agree to 1e-6."""
    s, case, p = assembled_2d(hip, steps=30, precision=precision)
    P = s.ppe()
    net = vx.LearnedPreconditioner(seed=2, backend=hip)
    net.set_weights(random_weights(7, scale=0.1))
    s.attach_lpc(net)
    info_j = s.CGM()
    pj = s.get("PNEW")[P["ids"]]
    p.solver = vx._abi.SOLVER_LEARNED_FCG
    s.set_params(p)
    s.CGM()                               # first learned solve: Jacobi fixes the baseline
    st = s.status()
    assert st.lpc_baseline == info_j.iterations
    info_l = s.CGM()
    pl = s.get("PNEW")[P["ids"]]
    st = s.status()
    print(f"jacobi {info_j.iterations} iterations, learned FCG {info_l.iterations}, "
          f"fallbacks {st.lpc_fallbacks}")
    assert st.lpc_fallbacks == 0
    assert info_l.converged and info_l.relres <= 1e-8
    assert _residual(P, pl) <= 1.5e-8 and _residual(P, pj) <= 1.5e-8
    assert np.linalg.norm(pl - pj) <= 1e-6 * np.linalg.norm(pj)
    s.close()
    net.close()


def test_fcg_fallback_to_jacobi(hip):
    """Flexible CG held to 10% of the Jacobi baseline (the default cap is
    200%) cannot converge; the solve falls back to Jacobi-CG, still meets the
    tolerance and resets the baseline."""
    s, case, p = assembled_2d(hip, steps=30)
    P = s.ppe()
    net = vx.LearnedPreconditioner(seed=3, backend=hip)
    w = np.zeros(vx.lpc.NPARAM, np.float32)
    W = olpc.unpack(w)
    W["bo"][:] = (-30.0, 30.0)
    net.set_weights(w)
    s.attach_lpc(net)
    p.solver = vx._abi.SOLVER_LEARNED_FCG
    s.set_params(p)
    ij = s.CGM().iterations
    info = s.CGM()
    st = s.status()
    assert st.lpc_fallbacks == 0 and info.converged
    p.lpc_cap_pct = 10
    s.set_params(p)
    info = s.CGM()
    st = s.status()
    print(f"iterations {info.iterations}, fallbacks {st.lpc_fallbacks}, baseline {st.lpc_baseline}")
    assert st.lpc_fallbacks == 1 and st.lpc_baseline == ij == info.iterations
    assert info.converged and _residual(P, s.get("PNEW")[P["ids"]]) <= 1.5e-8
    s.close()
    net.close()


def test_trained_network_gains(hip):
    """Train on systems dumped from jittered 2D dam breaks (seeds 1-3), then
    solve held-out systems (seed 9): the loss falls, and the trained network's
    flexible CG needs fewer iterations than Jacobi-CG, than the same network
    with its initial weights, and no more than the all-zero network
    (g = c = 1, the bare structure)."""
    net = vx.LearnedPreconditioner(seed=4, backend=hip)
    for seed in (1, 2, 3):
        s, case, p = assembled_2d(hip, steps=0, seed=seed)
        for k in range(8):
            s.step(5)
            for st in STAGES:
                getattr(s, st)()
            net.dump(s)
        s.close()
    n, rows = net.count()
    assert n == 24 and rows > 24 * 1000      # the fluid rows are unknowns, not free surface
    initial = net.weights()
    l0, l1 = net.train(epochs=20, lr=2e-3, seed=1)
    print(f"training loss {l0:.4e} -> {l1:.4e} over {n} systems ({rows} rows)")
    assert l1 < l0
    trained = net.weights()
    results = {}
    for name, w in (("initial", initial), ("zero", np.zeros(vx.lpc.NPARAM, np.float32)),
                    ("trained", trained)):
        net.set_weights(w)
        s, case, p = assembled_2d(hip, steps=40, seed=9)
        s.attach_lpc(net)
        ij = s.CGM().iterations
        p.solver = vx._abi.SOLVER_LEARNED_FCG
        s.set_params(p)
        s.CGM()
        info = s.CGM()
        assert info.converged and s.status().lpc_fallbacks == 0
        results[name] = (ij, info.iterations)
        s.close()
    print("held-out system: Jacobi-CG %d iterations, FCG initial weights %d, all-zero network %d, "
          "trained %d" % (results["zero"][0], results["initial"][1], results["zero"][1],
                          results["trained"][1]))
    assert results["trained"][1] < results["trained"][0]          # beats Jacobi-CG
    assert results["trained"][1] < results["initial"][1]          # training gains
    assert results["trained"][1] <= results["zero"][1]            # and does not lose to g = c = 1
    net.close()
