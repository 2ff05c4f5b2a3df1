"""The learned preconditioner's CPU restatement (oracle/lpc.py), pinned on its
own: analytic gradient against central finite differences in float64,
zero-output preconditioner against the textbook Neumann/Jacobi forms, the loss
against a dense evaluation, and the probe generator against SplitMix64."""
import numpy as np
import pytest

import voxarmps_amd as vx
from oracle import lpc
from voxarmps_amd.cases import splitmix64


def small_system(oracle, steps=0, scale=None):
    """A PPE system of the 2D dam break (l0 = 0.01, water 25 x 50), from the
    sequential oracle, prepared as the library stores a training system."""
    case = vx.dam_break_2d_c0()
    p = vx.dam2d_params(case)
    s = vx.Simulation(p, case, backend=oracle)
    s.init()
    if steps:
        s.step(steps)
    for st in ("PNUM", "NEIGHBOR", "SPS", "V_FRACTION", "VISCOSITY", "PREDICTION", "PNUMSTAR",
               "BCON", "MATRIX"):
        getattr(s, st)()
    P = s.ppe()
    thr = lpc.strong_threshold(2, p.lambda_, p.re_l, p.dl, s.n0()[1], p.rho1)
    S = lpc.System(P["rowptr"], P["cols"], P["vals"], P["rhs"], P["ids"], s.get("X"),
                   s.get("Y"), s.get("Z"), s.get("N"), s.n0()[1], case.ptype, p.re_l, 2, thr)
    return S, P, s


@pytest.fixture(scope="module")
def sys2d(oracle):
    return small_system(oracle)


def random_weights(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    w = rng.uniform(-scale, scale, lpc.NPARAM)
    # biases positive so that most ReLUs are active
    W = lpc.unpack(w)
    for k in ("b0", "c1", "c2", "c3"):
        W[k][:] = np.abs(W[k]) + 0.1
    return w


def test_param_count():
    assert lpc.NPARAM == 8 * 64 + 64 + 3 * (16 * 64 + 64 * 64 + 64 * 16 + 64) + 2 * 64 + 2


def test_rademacher_is_splitmix64():
    # SplitMix64 from state seed + probe * 2^32; the first outputs for seed 0
    # are the published test vector of the generator
    z = splitmix64(0, 3)
    assert int(z[0]) == 0xE220A8397B1DCDAF
    s = lpc.rademacher(0, 0, 3)
    assert np.array_equal(s, np.where((z >> np.uint64(63)) == 1, -1.0, 1.0))
    a = lpc.rademacher(7, 2, 1000)
    assert set(np.unique(a)) == {-1.0, 1.0} and abs(a.mean()) < 0.15


def test_bf16_rounding():
    # exactly representable, ties to even, and the largest error of 2^-9 relative
    assert lpc.bf16(1.0) == 1.0 and lpc.bf16(1.0 + 2 ** -8) == 1.0          # tie -> even
    assert lpc.bf16(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6                       # tie -> even (up)
    x = np.random.default_rng(0).normal(size=10000)
    assert np.max(np.abs(lpc.bf16(x) - x) / np.abs(x)) <= 2 ** -8


def test_zero_network_is_neumann_on_strong_graph(sys2d):
    """All-zero weights give g = c = 1: M = D^-1 + D^-1 O_s D^-1, built here
    densely from the exported CSR with the coupling rule applied entrywise."""
    S, P, _ = sys2d
    st = lpc.forward(np.zeros(lpc.NPARAM), S)
    assert np.all(st["g"] == 1.0) and np.allclose(st["c"], 1.0, rtol=1e-15)
    nu = S.nu
    A = np.zeros((nu, nu))
    for r in range(nu):
        a, b = P["rowptr"][r], P["rowptr"][r + 1]
        A[r, P["cols"][a:b]] = P["vals"][a:b]
    D = np.diag(A).copy()
    O = np.diag(D) - A
    O32 = O.astype(np.float32).astype(np.float64)          # the selection reads float32 values
    # strong: a quarter of the nearest-lattice-neighbour coupling (W(l0) = 3 for re = 4 l0)
    s = sys2d[2]
    p = s.params
    oref = 2 * 2 / p.lambda_ * 3.0 / s.n0()[1] / p.rho1
    keep = O32 >= 0.25 * oref
    assert 0 < keep.sum() < (O > 0).sum()
    Os = np.where(keep & (O > 0), O, 0.0)
    M = np.diag(1 / D) + Os / np.outer(D, D)
    w = np.random.default_rng(1).normal(size=nu)
    z = lpc.apply_m(S, st["g"], st["c"], w)
    # float32 storage of Oh and 1/D: 1e-6 of the result's scale
    ref = M @ w
    assert np.max(np.abs(z - ref)) <= 1e-6 * np.max(np.abs(ref))
    # M is symmetric positive definite and the coupling term off is Jacobi
    assert np.allclose(M, M.T) and np.linalg.eigvalsh(M).min() > 0
    z0 = lpc.apply_m(S, st["g"], np.zeros(nu) + 1e-300, w)
    assert np.max(np.abs(z0 - w / D)) <= 1e-6 * np.max(np.abs(w / D))


def test_loss_dense_brute_force(sys2d):
    S, P, _ = sys2d
    w = random_weights(3)
    st = lpc.forward(w, S)
    nu = S.nu
    A = S.A32.toarray()
    Mz = np.diag(st["g"] * S.inv_d) + np.diag(np.sqrt(st["c"])) @ S.Oh.toarray() @ np.diag(
        np.sqrt(st["c"]))
    probes = S.probes(11)
    E = A @ Mz - np.eye(nu)
    want = np.mean([(E @ v) @ (E @ v) / (v @ v) for v in probes])
    got = lpc.loss_outputs(S, st["g"], st["c"], probes)[0]
    assert got == pytest.approx(want, rel=1e-10)


def test_gradient_finite_differences(sys2d):
    """Analytic gradient against central differences along 12 random
    directions and 40 single coordinates (rounding off, float64):
    relative difference <= 1e-6."""
    S, _, _ = sys2d
    probes = S.probes(5)
    w = random_weights(4)
    L, g = lpc.loss_grad(w, S, probes, rounding=False)
    assert L > 0 and np.all(np.isfinite(g))
    rng = np.random.default_rng(9)
    h = 1e-6
    for _ in range(12):
        d = rng.normal(size=lpc.NPARAM)
        fd = (lpc.loss(w + h * d, S, probes, False) - lpc.loss(w - h * d, S, probes, False)) / (2 * h)
        assert fd == pytest.approx(g @ d, rel=1e-6, abs=1e-12 * abs(L))
    # single coordinates in every parameter block
    offs, k = [], 0
    for name, shape in lpc.LAYOUT:
        n = int(np.prod(shape))
        offs += list(k + rng.choice(n, min(n, 3), replace=False))
        k += n
    for i in offs:
        e = np.zeros(lpc.NPARAM); e[i] = 1.0
        fd = (lpc.loss(w + h * e, S, probes, False) - lpc.loss(w - h * e, S, probes, False)) / (2 * h)
        assert fd == pytest.approx(g[i], rel=1e-5, abs=1e-9 * max(abs(L), np.abs(g).max()))


def test_rounded_gradient_close_to_exact(sys2d):
    """The bf16 network's straight-through gradient stays close to the exact
    gradient of the unrounded network (a sanity bound, 10% in L2)."""
    S, _, _ = sys2d
    probes = S.probes(5)
    w = random_weights(4)
    _, g0 = lpc.loss_grad(w, S, probes, rounding=False)
    _, g1 = lpc.loss_grad(w, S, probes, rounding=True)
    assert np.linalg.norm(g1 - g0) <= 0.1 * np.linalg.norm(g0)
