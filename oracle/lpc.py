"""Learned PPE preconditioner: float64 CPU restatement of the network, the
preconditioner it parameterises, the training loss and its gradient.

Test infrastructure only (tests/ load it as the checker of the HIP kernels in
voxarmps_amd/csrc/vox_lpc.hip). The reference has no learned component, so
nothing here follows a reference file. It is pinned instead by:
  * central finite differences of the loss against the analytic gradient
    (rounding switched off), in float64;
  * the preconditioner with zero network output reducing to the textbook
    form D^-1 + D^-1 O_s D^-1 (Neumann series truncated after one term on the
    strong-coupling graph) and, with the coupling term off, to Jacobi;
  * a brute-force dense evaluation of the loss on a tiny system.

System (one PPE, unknowns r = 0..nu-1; A = D - O, O >= 0 off the diagonal;
the network reads the float32 roundings O32 of O and p of the positions):
  features f_r (8): n_r/n0, wall flag, sum_k O32_rk / A_rr, the three first
    moments sum_k a_rk (p_k - p_r)/re with a_rk = O32_rk / sum_k O32_rk, 0, 0;
  strong couplings S: (r, k) with O32_rk >= theta * O_ref, O_ref the coupling
    of two particles one spacing apart (strong_threshold), a symmetric
    pattern; Oh_rk = O32_rk / (A_rr A_kk) evaluated as float32 products with
    the float32 1/A_rr.
Network (hidden 64, messages 16, bf16 operands with float accumulation):
  h0 = relu(W0 f + b0)
  for l = 1..3: s = P_l h_{l-1};  m_r = sum_k a_rk s_k;
                h_l = relu(U_l h_{l-1} + V_l m + c_l)
  o = Wo h3 + bo;  g = exp(0.5 tanh o0);  c = CMAX * sigmoid(o1 + BETA)
  (zero output gives g = c = 1).
Preconditioner: z_r = g_r w_r / A_rr + sqrt(c_r) sum_{k in S(r)} Oh_rk sqrt(c_k) w_k.
Loss over probes w_p (p = 0: the right-hand side; p >= 1: Rademacher signs):
  L = (1/P) sum_p ||A M w_p - w_p||^2 / ||w_p||^2.
"""
import numpy as np
import scipy.sparse as sp

NF, NH, NM, NO = 8, 64, 16, 2
NLAYER = 3
CMAX = 2.5
BETA = -np.log(CMAX - 1.0)
THETA = 0.25
NPROBE = 4

LAYOUT = [("W0", (NH, NF)), ("b0", (NH,))]
for _l in range(1, NLAYER + 1):
    LAYOUT += [(f"P{_l}", (NM, NH)), (f"U{_l}", (NH, NH)), (f"V{_l}", (NH, NM)), (f"c{_l}", (NH,))]
LAYOUT += [("Wo", (NO, NH)), ("bo", (NO,))]
NPARAM = sum(int(np.prod(s)) for _, s in LAYOUT)
SPLITMIX_GAMMA = np.uint64(0x9E3779B97F4A7C15)


def unpack(w):
    """Views of the flat parameter vector, in the packed order."""
    out, k = {}, 0
    for name, shape in LAYOUT:
        n = int(np.prod(shape))
        out[name] = w[k:k + n].reshape(shape)
        k += n
    return out


def bf16(x):
    """Round to float32, then to bfloat16 (nearest, ties to even), as float64."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    b = a.view(np.uint32).astype(np.uint64)
    b = (b + np.uint64(0x7FFF) + ((b >> np.uint64(16)) & np.uint64(1))) & np.uint64(0xFFFF0000)
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def f32(x):
    return np.asarray(x, np.float32).astype(np.float64)


def strong_threshold(dim, lam, re, dl, n0, rho1):
    """theta * ((2 dim / lambda) W(dl) / n0) / rho1 for the rational weight
    W(r) = re/r - 1, in the library's order of operations."""
    q = dl / re
    w = (1 / q) - 1 if q < 1.0 else 0.0
    return THETA * ((2.0 * dim / lam * w / n0) / rho1)


def rademacher(seed, probe, n):
    """Signs of probe `probe` over n rows: SplitMix64 started at
    seed + probe * 2^32, output r+1 for row r, top bit set -> -1."""
    k = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.uint64(probe) << np.uint64(32)) + k * SPLITMIX_GAMMA
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return np.where((z >> np.uint64(63)) == 1, -1.0, 1.0)


class System:
    """A PPE system prepared as the library stores a training system: float32
    matrix values; features and couplings from the float32 values and
    positions, accumulated in float64, then rounded to float32."""

    def __init__(self, rowptr, cols, vals, rhs, ids, x, y, z, pnd, n0, ptype, re, dim, thr):
        nu = len(rhs)
        self.nu = nu
        rowptr = np.asarray(rowptr); cols = np.asarray(cols); vals = np.asarray(vals, np.float64)
        rows = np.repeat(np.arange(nu), np.diff(rowptr))
        dmask = cols == rows
        diag = np.zeros(nu)
        diag[rows[dmask]] = vals[dmask]
        orow, ocol, oval = rows[~dmask], cols[~dmask], -vals[~dmask]
        o32 = f32(oval)
        osum = np.zeros(nu)
        np.add.at(osum, orow, o32)
        inv_osum = np.where(osum > 0, 1.0 / np.where(osum > 0, osum, 1.0), 0.0)
        self.inv_d = f32(1.0 / diag)
        pid = np.asarray(ids)
        px, py = f32(np.asarray(x)[pid]), f32(np.asarray(y)[pid])
        pz = f32(np.asarray(z)[pid]) if dim == 3 else np.zeros(nu)
        inv_re = float(np.float32(1.0 / re))
        f = np.zeros((nu, NF))
        f[:, 0] = np.asarray(pnd)[pid] / n0
        f[:, 1] = (np.asarray(ptype)[pid] == 0).astype(np.float64)
        f[:, 2] = osum * self.inv_d
        for a, pc in enumerate((px, py, pz)):
            mom = np.zeros(nu)
            np.add.at(mom, orow, o32 * (pc[ocol] - pc[orow]))
            f[:, 3 + a] = mom * inv_osum * inv_re
        if dim != 3:
            f[:, 5] = 0.0
        self.f = f32(f)
        self.inv_osum = f32(inv_osum)
        # float32 matrix of the loss products and the aggregation weights
        self.diag32 = f32(diag)
        self.O32 = sp.csr_matrix((o32, (orow, ocol)), shape=(nu, nu))
        self.A32 = sp.diags(self.diag32) - self.O32
        keep = o32 >= thr
        i32 = self.inv_d.astype(np.float32)
        oh = (o32[keep].astype(np.float32) * i32[orow[keep]]) * i32[ocol[keep]]
        self.Oh = sp.csr_matrix((oh.astype(np.float64), (orow[keep], ocol[keep])), shape=(nu, nu))
        self.rhs32 = f32(rhs)
        self.agg = sp.diags(self.inv_osum) @ self.O32     # a_rk as the kernels form it

    def probes(self, seed, nprobe=NPROBE):
        return [self.rhs32] + [rademacher(seed, p, self.nu) for p in range(1, nprobe)]


def forward(w, S, rounding=True):
    """Network outputs and the intermediates the backward pass needs."""
    W = unpack(np.asarray(w, np.float64))
    R = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    st = {"h": [], "m": [], "s": []}
    h = R(np.maximum(R(S.f) @ R(W["W0"]).T + W["b0"], 0.0))
    st["h"].append(h)
    for l in range(1, NLAYER + 1):
        s = R(h @ R(W[f"P{l}"]).T)
        m = R(S.agg @ s)
        h = R(np.maximum(h @ R(W[f"U{l}"]).T + m @ R(W[f"V{l}"]).T + W[f"c{l}"], 0.0))
        st["s"].append(s); st["m"].append(m); st["h"].append(h)
    o = h @ W["Wo"].T + W["bo"]
    st["o"] = o
    st["g"] = np.exp(0.5 * np.tanh(o[:, 0]))
    st["c"] = CMAX / (1.0 + np.exp(-(o[:, 1] + BETA)))
    return st


def apply_m(S, g, c, w):
    sc = np.sqrt(c)
    return g * S.inv_d * w + sc * (S.Oh @ (sc * w))


def loss_outputs(S, g, c, probes):
    """Loss and its gradient with respect to g and c."""
    sc = np.sqrt(c)
    P = len(probes)
    L = 0.0
    dg = np.zeros(S.nu); dc = np.zeros(S.nu)
    for wv in probes:
        nn = wv @ wv
        if nn == 0:
            continue
        t = S.Oh @ (sc * wv)
        z = g * S.inv_d * wv + sc * t
        e = S.A32 @ z - wv
        L += (e @ e) / (P * nn)
        y = (2.0 / (P * nn)) * (S.A32 @ e)
        u = S.Oh @ (sc * y)
        dg += y * wv * S.inv_d
        dc += 0.5 / sc * (y * t + wv * u)
    return L, dg, dc


def loss(w, S, probes, rounding=True):
    st = forward(w, S, rounding)
    return loss_outputs(S, st["g"], st["c"], probes)[0]


def loss_grad(w, S, probes, rounding=True):
    """Loss and d loss / d w (flat, packed order). Rounding to bf16 is taken
    as the identity in the backward pass."""
    W = unpack(np.asarray(w, np.float64))
    R = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    st = forward(w, S, rounding)
    L, dg, dc = loss_outputs(S, st["g"], st["c"], probes)
    o = st["o"]
    th = np.tanh(o[:, 0])
    sg = 1.0 / (1.0 + np.exp(-(o[:, 1] + BETA)))
    do = np.stack([dg * st["g"] * 0.5 * (1 - th * th), dc * CMAX * sg * (1 - sg)], 1)
    G = {k: np.zeros_like(v) for k, v in W.items()}
    h = st["h"]
    G["Wo"] = do.T @ h[NLAYER]
    G["bo"] = do.sum(0)
    dh = do @ W["Wo"]
    for l in range(NLAYER, 0, -1):
        dpre = dh * (h[l] > 0)
        m, s = st["m"][l - 1], st["s"][l - 1]
        G[f"U{l}"] = dpre.T @ h[l - 1]
        G[f"V{l}"] = dpre.T @ m
        G[f"c{l}"] = dpre.sum(0)
        dm = dpre @ R(W[f"V{l}"])
        ds = S.agg.T @ dm
        G[f"P{l}"] = ds.T @ h[l - 1]
        dh = dpre @ R(W[f"U{l}"]) + ds @ R(W[f"P{l}"])
    dpre = dh * (h[0] > 0)
    G["W0"] = dpre.T @ R(S.f)
    G["b0"] = dpre.sum(0)
    return L, np.concatenate([G[k].ravel() for k, _ in LAYOUT])
