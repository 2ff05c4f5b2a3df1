// vox_lpc.hip — learned PPE preconditioner on gfx950: a message-passing
// network over the PPE matrix graph, its training on the GPU, and the
// coupling operator it parameterises for flexible CG (vox_cg.hip).
//
// No reference code: the reference solves the PPE with IC(0)-CG only
// (functions.cu:1139-1200, 506-716). The preconditioner is
//   z_r = g_r w_r / A_rr + sqrt(c_r) sum_{k in S(r)} O_rk / (A_rr A_kk) sqrt(c_k) w_k
// with O = -offdiag(A) >= 0 and S the strong couplings O_rk >= theta * O_ref
// (O_ref the coupling of two particles one spacing apart, theta = 0.25: a
// symmetric pattern, about the neighbours within 2.3 spacings). g = c = 1 is
// the Neumann series D^-1 + D^-1 O_S D^-1 truncated after one term on S; the
// network sets g and c per unknown from local features.
//
// Network (per unknown r, hidden 64, messages 16; bf16 operands on the matrix
// cores, FP32 accumulation; weights packed as in voxarmps.h):
//   h0 = relu(W0 f + b0)
//   for l = 1..3: s = P_l h_{l-1};  m_r = (sum_k O_rk s_k) / sum_k O_rk
//                 h_l = relu(U_l h_{l-1} + V_l m + c_l)
//   o = Wo h3 + bo;  g = exp(0.5 tanh o0);  c = 2.5 sigmoid(o1 - ln 1.5)
// Features f: n_r/n0, wall flag, sum_k O_rk / A_rr, the first moments
// sum_k O_rk (x_k - x_r) / re / sum_k O_rk (three axes), 0, 0.
//
// Layout: rows of activations are contiguous (64 bf16 = 128 bytes of hidden
// state, 16 bf16 = 32 bytes of message per row), so the aggregation gathers
// one 32-byte message per matrix entry; matrix and coupling operator are
// SELL-64 (mx_index). The dense layers run one wavefront per 16-row tile:
// v_mfma_f32_16x16x32_bf16 with the weights held in registers for the whole
// launch, the 16 x 64 output tile staged in LDS for the coalesced store and
// for the message product that follows in the same kernel.
//
// Training (mean over probes w of ||A M w - w||^2 / ||w||^2, Adam): the
// backward pass runs in FP32 on the vector units, one thread per row, with
// weight gradients reduced per block in a fixed order (deterministic).
#include <algorithm>
#include <cstring>
#include <numeric>

#include "vox_internal.hpp"

namespace vox {

constexpr int LF = 8, LH = 64, LM = 16, LO = 2, LNL = 3;
constexpr int OFF_W0 = 0, OFF_B0 = OFF_W0 + LH * LF;
constexpr int LAYER_SZ = LM * LH + LH * LH + LH * LM + LH;
__host__ __device__ constexpr int off_P(int l) { return OFF_B0 + LH + (l - 1) * LAYER_SZ; }
__host__ __device__ constexpr int off_U(int l) { return off_P(l) + LM * LH; }
__host__ __device__ constexpr int off_V(int l) { return off_U(l) + LH * LH; }
__host__ __device__ constexpr int off_c(int l) { return off_V(l) + LH * LM; }
constexpr int OFF_WO = off_P(LNL + 1), OFF_BO = OFF_WO + LO * LH, NPARAM = OFF_BO + LO;
static_assert(NPARAM == VOX_LPC_NPARAM, "parameter count");
constexpr float kCMax = 2.5f;
constexpr float kBeta = -0.40546510810816438f;   // -ln(1.5): c = 1 at zero output
constexpr int kProbes = 4;
constexpr int kOuterBlocks = 256;
constexpr int kOuterMax = LH * LH;

// ---- bf16 (round to nearest even; inputs are finite) -----------------------
__device__ __forceinline__ uint16_t f2bf(float x)
{
    uint32_t b = __float_as_uint(x);
    b += 0x7FFFu + ((b >> 16) & 1u);
    return (uint16_t)(b >> 16);
}
__device__ __forceinline__ float bf2f(uint32_t h) { return __uint_as_float(h << 16); }
__device__ __forceinline__ float bf_lo(int32_t x) { return __uint_as_float((uint32_t)x << 16); }
__device__ __forceinline__ float bf_hi(int32_t x) { return __uint_as_float((uint32_t)x & 0xFFFF0000u); }

__device__ __forceinline__ double rademacher(uint64_t seed, int probe, int r)
{
    uint64_t z = seed + ((uint64_t)probe << 32) + (uint64_t)(r + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (z >> 63) ? -1.0 : 1.0;
}

// ---- features and strong couplings ------------------------------------------
// Per-row preparation: FP32 position of the unknown's particle, 1/A_rr, and
// the two features that need no neighbours (n/n0, wall flag); padding rows
// get zero features.
struct PrepArgs {
    int nu, rows, dim;
    double n0;
    const double* __restrict__ adiag;
    const int32_t* __restrict__ urow;
    const int32_t* __restrict__ ptype;
    const double *x, *y, *z, *pnd;
    v4f* upos;
    float* invd;
    float* feat;
};

__global__ __launch_bounds__(kBlock) void k_lpc_rowprep(PrepArgs a)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.rows) return;
    float* f = a.feat + 8L * r;
    if (r >= a.nu) {
        for (int q = 0; q < 8; ++q) f[q] = 0.f;
        return;
    }
    const int i = a.urow[r];
    a.upos[r] = v4f{(float)a.x[i], (float)a.y[i], a.dim == 3 ? (float)a.z[i] : 0.f, 0.f};
    a.invd[r] = (float)(1.0 / a.adiag[r]);
    f[0] = (float)(a.pnd[i] / a.n0);
    f[1] = (a.ptype[i] == 0) ? 1.f : 0.f;
}

// Row sums over the FP32 off-diagonal magnitudes O_rk = -a_rk (sum and the
// first moments of the FP32 neighbour offsets: features 2..7), and the strong
// couplings O_rk >= thr of the row (column and O_rk, in matrix order).
struct FeatArgs {
    int nu, cap;
    float inv_re;
    double thr;
    const int32_t* __restrict__ rowlen;
    const int32_t* __restrict__ acol;
    const float* __restrict__ aval;
    const v4f* __restrict__ upos;
    const float* __restrict__ invd;
    float* feat;
    float* inv_osum;
    int32_t* slen;
    int32_t* scol;
    float* sval;
};

__global__ __launch_bounds__(kBlock) void k_lpc_features(FeatArgs a)
{
    const int r = xcd_tid();
    if (r >= a.nu) return;
    const v4f pr = a.upos[r];
    float osum = 0.f, mx = 0.f, my = 0.f, mz = 0.f;
    const long b = mx_index(r, 0, a.cap);
    const int len = a.rowlen[r];
    int n = 0;
    constexpr int U = 8;
    for (int k = 0; k < len; k += U) {
        int32_t cc[U];
        float oo[U];
        v4f pp[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const bool ok = (k + u) < len;
            cc[u] = ok ? __builtin_nontemporal_load(&a.acol[b + 64L * (k + u)]) : r;
            oo[u] = ok ? -__builtin_nontemporal_load(&a.aval[b + 64L * (k + u)]) : 0.f;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) pp[u] = a.upos[cc[u]];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            osum += oo[u];
            mx = fmaf(oo[u], pp[u].x - pr.x, mx);
            my = fmaf(oo[u], pp[u].y - pr.y, my);
            mz = fmaf(oo[u], pp[u].z - pr.z, mz);
            if ((double)oo[u] >= a.thr) {   // padding has oo = 0 < thr
                a.scol[b + 64L * n] = cc[u];
                a.sval[b + 64L * n] = oo[u];
                ++n;
            }
        }
    }
    a.slen[r] = n;
    const float inv = (osum > 0.f) ? 1.f / osum : 0.f;
    float* f = a.feat + 8L * r;
    f[2] = osum * a.invd[r];
    f[3] = mx * inv * a.inv_re;
    f[4] = my * inv * a.inv_re;
    f[5] = mz * inv * a.inv_re;
    f[6] = 0.f;
    f[7] = 0.f;
    a.inv_osum[r] = inv;
}

// Training-set coupling values: O_rk / (A_rr A_kk) as (O_rk * (1/A_rr)) * (1/A_kk).
__global__ __launch_bounds__(kBlock) void k_lpc_oh(int nu, int cap, const int32_t* __restrict__ slen,
                                                   const int32_t* __restrict__ scol,
                                                   const float* __restrict__ invd, float* sval)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nu) return;
    const long b = mx_index(r, 0, cap);
    const float ir = invd[r];
    for (int k = 0; k < slen[r]; ++k) sval[b + 64L * k] = (sval[b + 64L * k] * ir) * invd[scol[b + 64L * k]];
}

// FP64 -> FP32 right-hand side and diagonal of a dumped system.
__global__ void k_lpc_dump_rows(int nu, const double* __restrict__ adiag, const double* __restrict__ rhs,
                                float* diag32, float* rhs32)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nu) return;
    diag32[r] = (float)adiag[r];
    rhs32[r] = (float)rhs[r];
}

// ---- network forward --------------------------------------------------------
struct AggArgs {
    int nu, cap;
    const int32_t* __restrict__ rowlen;
    const int32_t* __restrict__ acol;
    const float* __restrict__ aval;
    const float* __restrict__ inv_osum;
    const uint16_t* __restrict__ s;
    uint16_t* m;
};

// m_r = (sum_k O_rk s_k) / sum_k O_rk over the matrix row; one thread per row,
// 8 entries' 32-byte message gathers in flight.
__global__ __launch_bounds__(kBlock) void k_lpc_agg(AggArgs a)
{
    const int r = xcd_tid();
    if (r >= a.nu) return;
    float acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0.f;
    const long b = mx_index(r, 0, a.cap);
    const int len = a.rowlen[r];
    constexpr int U = 8;
    for (int k = 0; k < len; k += U) {
        int32_t cc[U];
        float oo[U];
        v4i x[U][2];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const bool ok = (k + u) < len;
            cc[u] = ok ? __builtin_nontemporal_load(&a.acol[b + 64L * (k + u)]) : r;
            oo[u] = ok ? -__builtin_nontemporal_load(&a.aval[b + 64L * (k + u)]) : 0.f;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const v4i* p = reinterpret_cast<const v4i*>(a.s + 16L * cc[u]);
            x[u][0] = p[0];
            x[u][1] = p[1];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const v4i v = x[u][h];
                const int32_t e[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    acc[8 * h + 2 * t] = fmaf(oo[u], bf_lo(e[t]), acc[8 * h + 2 * t]);
                    acc[8 * h + 2 * t + 1] = fmaf(oo[u], bf_hi(e[t]), acc[8 * h + 2 * t + 1]);
                }
            }
        }
    }
    const float io = a.inv_osum[r];
    int32_t w[8];
#pragma unroll
    for (int t = 0; t < 8; ++t)
        w[t] = (int32_t)((uint32_t)f2bf(acc[2 * t] * io) | ((uint32_t)f2bf(acc[2 * t + 1] * io) << 16));
    v4i* o = reinterpret_cast<v4i*>(a.m + 16L * r);
    o[0] = v4i{w[0], w[1], w[2], w[3]};
    o[1] = v4i{w[4], w[5], w[6], w[7]};
}

// Inference: t_r = sqrt(c_r) / A_rr, gd_r = g_r / A_rr from the network
// output, then coupling values (O_rk t_r) t_k.
__global__ void k_lpc_gc_inf(int nu, const float* __restrict__ o, const float* __restrict__ invd, float* gd,
                             float* t)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nu) return;
    const float gv = expf(0.5f * tanhf(o[2 * r]));
    const float cv = kCMax / (1.f + expf(-(o[2 * r + 1] + kBeta)));
    gd[r] = gv * invd[r];
    t[r] = sqrtf(cv) * invd[r];
}

__global__ __launch_bounds__(kBlock) void k_lpc_scale(int nu, int cap, const int32_t* __restrict__ slen,
                                                      const int32_t* __restrict__ scol,
                                                      const float* __restrict__ t, float* sval)
{
    const int r = xcd_tid();
    if (r >= nu) return;
    const long b = mx_index(r, 0, cap);
    const float tr = t[r];
    for (int k = 0; k < slen[r]; ++k) sval[b + 64L * k] = (sval[b + 64L * k] * tr) * t[scol[b + 64L * k]];
}

struct DenseArgs {
    int rows, layer;            // rows: multiple of 16; layer 0 = encoder
    const uint16_t* __restrict__ w16;
    const float* __restrict__ w32;
    const float* __restrict__ feat;
    const uint16_t* __restrict__ hin;
    const uint16_t* __restrict__ min;
    uint16_t* hout;
    uint16_t* sout;             // next layer's messages (layer < 3)
    float* o;                   // network output (layer 3)
};

typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

__device__ __forceinline__ v8s ld8(const uint16_t* p) { return *reinterpret_cast<const v8s*>(p); }

// D += A B on gfx950's 16 x 16 x 32 bf16 matrix instruction.
__device__ __forceinline__ v4f mfma32(v8s x, v8s y, v4f acc)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, x),
                                                   __builtin_bit_cast(v8bf, y), acc, 0, 0, 0);
}

// One 16-row tile per wavefront. Operand layouts of v_mfma_f32_16x16x32_bf16:
// lane l supplies A[row l%16][k 8*(l/16)..+7] and B[k 8*(l/16)..+7][col l%16]
// and receives D[row 4*(l/16)+i][col l%16], i = 0..3. Rows are graph rows,
// columns are output features, k runs over input features: the encoder's 8
// features are one instruction (k 8..31 zero), a message layer's 64 hidden
// inputs two and its 16 messages one (k 16..31 zero), the message
// projection's 64 inputs two.
template <int MODE>   // 0: encoder (features in), 1: message layer
__global__ __launch_bounds__(kBlock) void k_lpc_dense(DenseArgs a)
{
    __shared__ __align__(16) uint16_t th[4][16][LH + 8];
    __shared__ __align__(16) uint16_t ts[4][16][LM];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int cl = lane & 15, gp = lane >> 4;
    constexpr int KB = (MODE == 0) ? 1 : 3;
    const v8s zero8 = {0, 0, 0, 0, 0, 0, 0, 0};
    const int l = a.layer;
    v8s bw[KB][4];
    float bias[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        const int o = nt * 16 + cl;
        if (MODE == 0) {
            bw[0][nt] = (gp == 0) ? ld8(a.w16 + OFF_W0 + o * LF) : zero8;
            bias[nt] = a.w32[OFF_B0 + o];
        } else {
#pragma unroll
            for (int kb = 0; kb < 2; ++kb) bw[kb][nt] = ld8(a.w16 + off_U(l) + o * LH + kb * 32 + 8 * gp);
            bw[KB - 1][nt] = (gp < 2) ? ld8(a.w16 + off_V(l) + o * LM + 8 * gp) : zero8;
            bias[nt] = a.w32[off_c(l) + o];
        }
    }
    const bool next = (l < LNL);
    v8s bp[2];
#pragma unroll
    for (int kb = 0; kb < 2; ++kb)
        bp[kb] = next ? ld8(a.w16 + off_P(l + 1) + cl * LH + kb * 32 + 8 * gp) : zero8;
    float wo[2][LH / 2];   // head weights of this lane's output (lanes < 32)
    if (!next) {
        const int q = lane & 1;
#pragma unroll
        for (int k = 0; k < LH / 2; ++k) {
            wo[0][k] = a.w32[OFF_WO + q * LH + k];
            wo[1][k] = a.w32[OFF_WO + q * LH + LH / 2 + k];
        }
    }
    const int ntile = a.rows / 16;
    for (int base = blockIdx.x * 4; base < ntile; base += gridDim.x * 4) {
        const int tile = base + wv;
        const bool act = tile < ntile;
        const long row0 = (long)tile * 16;
        if (act) {
            v8s af[KB];
            if (MODE == 0) {
                if (gp == 0) {
                    const float* f = a.feat + (row0 + cl) * LF;
                    af[0] = v8s{(short)f2bf(f[0]), (short)f2bf(f[1]), (short)f2bf(f[2]),
                                (short)f2bf(f[3]), (short)f2bf(f[4]), (short)f2bf(f[5]),
                                (short)f2bf(f[6]), (short)f2bf(f[7])};
                } else {
                    af[0] = zero8;
                }
            } else {
#pragma unroll
                for (int kb = 0; kb < 2; ++kb) af[kb] = ld8(a.hin + (row0 + cl) * LH + kb * 32 + 8 * gp);
                if (gp < 2) af[KB - 1] = ld8(a.min + (row0 + cl) * LM + 8 * gp);
                else af[KB - 1] = zero8;
            }
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int kb = 0; kb < KB; ++kb) acc = mfma32(af[kb], bw[kb][nt], acc);
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    th[wv][4 * gp + i][nt * 16 + cl] = f2bf(fmaxf(acc[i] + bias[nt], 0.f));
            }
        }
        __syncthreads();
        if (act) {
            const int nd = lane >> 2, pt = lane & 3;
            const v4i* src = reinterpret_cast<const v4i*>(&th[wv][nd][pt * 16]);
            v4i* dst = reinterpret_cast<v4i*>(a.hout + (row0 + nd) * LH + pt * 16);
            dst[0] = src[0];
            dst[1] = src[1];
            if (next) {
                v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int kb = 0; kb < 2; ++kb) {
                    const v8s x = *reinterpret_cast<const v8s*>(&th[wv][cl][kb * 32 + 8 * gp]);
                    acc = mfma32(x, bp[kb], acc);
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) ts[wv][4 * gp + i][cl] = f2bf(acc[i]);
            } else if (lane < 32) {
                // head: lane = 2 * row + output
                const int nd2 = lane >> 1, q = lane & 1;
                float acc = a.w32[OFF_BO + q];
#pragma unroll
                for (int k = 0; k < LH / 2; ++k) acc = fmaf(wo[0][k], bf2f(th[wv][nd2][k]), acc);
#pragma unroll
                for (int k = 0; k < LH / 2; ++k) acc = fmaf(wo[1][k], bf2f(th[wv][nd2][LH / 2 + k]), acc);
                a.o[(row0 + nd2) * LO + q] = acc;
            }
        }
        __syncthreads();
        if (act && next && lane < 32) {
            const int nd = lane >> 1, hf = lane & 1;
            *reinterpret_cast<v4i*>(a.sout + (row0 + nd) * LM + hf * 8) =
                *reinterpret_cast<const v4i*>(&ts[wv][nd][hf * 8]);
        }
        __syncthreads();
    }
}

// g, sqrt(c) (and c) from the network output.
__global__ void k_lpc_gc(int nu, const float* __restrict__ o, float* g, float* c, float* sc)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nu) return;
    const float gv = expf(0.5f * tanhf(o[2 * r]));
    const float cv = kCMax / (1.f + expf(-(o[2 * r + 1] + kBeta)));
    g[r] = gv;
    if (c) c[r] = cv;
    sc[r] = sqrtf(cv);
}

// ---- training: probes and the loss gradient with respect to g and c --------
enum { LS_NN = 0, LS_COEF, LS_LOSS, LS_EPOCH, LS_COUNT };

__global__ void k_lpc_probe(int nu, int probe, uint64_t seed, const float* __restrict__ rhs, float* w)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nu) return;
    w[r] = probe == 0 ? rhs[r] : (float)rademacher(seed, probe, r);
}

__global__ __launch_bounds__(kBlock) void k_lpc_dot(int nu, const float* __restrict__ x, double* part)
{
    __shared__ double lds[kBlock];
    double s = 0.0;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nu; r += gridDim.x * blockDim.x)
        s += (double)x[r] * (double)x[r];
    s = block_sum(s, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// mode 0: scal[LS_NN] = sum, scal[LS_COEF] = 2 / (probes * sum)
// mode 1: scal[LS_LOSS] += sum / (probes * scal[LS_NN])
// mode 2: scal[LS_EPOCH] += scal[LS_LOSS]
__global__ __launch_bounds__(kBlock) void k_lpc_fin(const double* part, int n, double* scal, int mode)
{
    __shared__ double lds[kBlock];
    double s = 0.0;
    for (int k = threadIdx.x; k < n; k += kBlock) s = s + part[k];
    s = block_sum(s, lds);
    if (threadIdx.x != 0) return;
    if (mode == 0) {
        scal[LS_NN] = s;
        scal[LS_COEF] = s > 0 ? 2.0 / (kProbes * s) : 0.0;
    } else if (mode == 1) {
        scal[LS_LOSS] += scal[LS_NN] > 0 ? s / (kProbes * scal[LS_NN]) : 0.0;
    } else {
        scal[LS_EPOCH] += scal[LS_LOSS];
    }
}

__global__ void k_lpc_set(double* scal, int slot, double v) { scal[slot] = v; }

struct OpView {   // a SELL-64 operator: rows, cap, lengths, columns, values
    int nu, cap;
    const int32_t* __restrict__ len;
    const int32_t* __restrict__ col;
    const float* __restrict__ val;
};

// t = Oh (sc w); z = g w / A_rr + sc t
__global__ __launch_bounds__(kBlock) void k_lpc_mprobe(OpView S, const float* __restrict__ sc,
                                                       const float* __restrict__ g,
                                                       const float* __restrict__ inv_d,
                                                       const float* __restrict__ w, float* t, float* z)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= S.nu) return;
    const long b = mx_index(r, 0, S.cap);
    float acc = 0.f;
    for (int k = 0; k < S.len[r]; ++k) {
        const int cc = S.col[b + 64L * k];
        acc = fmaf(S.val[b + 64L * k], sc[cc] * w[cc], acc);
    }
    t[r] = acc;
    z[r] = fmaf(sc[r], acc, (g[r] * inv_d[r]) * w[r]);
}

// out = A x - sub (partial ||out||^2 per block), or out = coef * A x (sub null)
__global__ __launch_bounds__(kBlock) void k_lpc_spmv(OpView A, const float* __restrict__ diag,
                                                     const float* __restrict__ x,
                                                     const float* __restrict__ sub,
                                                     const double* __restrict__ scal, float* out,
                                                     double* part)
{
    __shared__ double lds[kBlock];
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    double ee = 0.0;
    if (r < A.nu) {
        const long b = mx_index(r, 0, A.cap);
        float acc = diag[r] * x[r];
        for (int k = 0; k < A.len[r]; ++k) acc = fmaf(A.val[b + 64L * k], x[A.col[b + 64L * k]], acc);
        if (sub) {
            const float e = acc - sub[r];
            out[r] = e;
            ee = (double)e * (double)e;
        } else {
            out[r] = (float)(scal[LS_COEF] * (double)acc);
        }
    }
    if (sub) {
        ee = block_sum(ee, lds);
        if (threadIdx.x == 0) part[blockIdx.x] = ee;
    }
}

// u = Oh (sc y); dL/dg += y w / A_rr; dL/dc += (y t + w u) / (2 sc)
__global__ __launch_bounds__(kBlock) void k_lpc_dgc(OpView S, const float* __restrict__ sc,
                                                    const float* __restrict__ inv_d,
                                                    const float* __restrict__ w,
                                                    const float* __restrict__ t,
                                                    const float* __restrict__ y, float* dgdc, int first)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= S.nu) return;
    const long b = mx_index(r, 0, S.cap);
    float u = 0.f;
    for (int k = 0; k < S.len[r]; ++k) {
        const int cc = S.col[b + 64L * k];
        u = fmaf(S.val[b + 64L * k], sc[cc] * y[cc], u);
    }
    const float dg = y[r] * w[r] * inv_d[r];
    const float dc = (0.5f / sc[r]) * (y[r] * t[r] + w[r] * u);
    dgdc[2 * r] = first ? dg : dgdc[2 * r] + dg;
    dgdc[2 * r + 1] = first ? dc : dgdc[2 * r + 1] + dc;
}

// ---- training: backward through the network (FP32) -------------------------
__global__ __launch_bounds__(kBlock) void k_lpc_head_bwd(int nu, const float* __restrict__ w32,
                                                         const float* __restrict__ o,
                                                         const float* __restrict__ g,
                                                         const float* __restrict__ dgdc, float* dO,
                                                         float* dh)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nu) return;
    const float th = tanhf(o[2 * r]);
    const float s = 1.f / (1.f + expf(-(o[2 * r + 1] + kBeta)));
    const float d0 = dgdc[2 * r] * g[r] * 0.5f * (1.f - th * th);
    const float d1 = dgdc[2 * r + 1] * kCMax * s * (1.f - s);
    dO[2 * r] = d0;
    dO[2 * r + 1] = d1;
    for (int k = 0; k < LH; ++k) dh[LH * (long)r + k] = fmaf(d0, w32[OFF_WO + k], d1 * w32[OFF_WO + LH + k]);
}

// dpre = dh * [h > 0]; for layers >= 1 also dh_prev = dpre U (direct part)
// and dmh = (dpre V) / sum_k O_rk (the aggregation's input gradient, scaled).
__global__ __launch_bounds__(kBlock) void k_lpc_dense_bwd(int nu, int l, const uint16_t* __restrict__ w16,
                                                          const uint16_t* __restrict__ h,
                                                          const float* __restrict__ dh,
                                                          const float* __restrict__ inv_osum,
                                                          float* dpre, float* dmh, float* dh_prev)
{
    __shared__ float sU[LH * LH];
    __shared__ float sV[LH * LM];
    if (l > 0) {
        for (int k = threadIdx.x; k < LH * LH; k += kBlock) sU[k] = bf2f(w16[off_U(l) + k]);
        for (int k = threadIdx.x; k < LH * LM; k += kBlock) sV[k] = bf2f(w16[off_V(l) + k]);
    }
    __syncthreads();
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nu) return;
    float dp[LH];
#pragma unroll
    for (int o = 0; o < LH; ++o) {
        const float hv = bf2f(h[LH * (long)r + o]);
        dp[o] = hv > 0.f ? dh[LH * (long)r + o] : 0.f;
        dpre[LH * (long)r + o] = dp[o];
    }
    if (l == 0) return;
    for (int k = 0; k < LH; ++k) {
        float acc = 0.f;
#pragma unroll
        for (int o = 0; o < LH; ++o) acc = fmaf(dp[o], sU[o * LH + k], acc);
        dh_prev[LH * (long)r + k] = acc;
    }
    const float io = inv_osum[r];
    for (int j = 0; j < LM; ++j) {
        float acc = 0.f;
#pragma unroll
        for (int o = 0; o < LH; ++o) acc = fmaf(dp[o], sV[o * LM + j], acc);
        dmh[LM * (long)r + j] = acc * io;
    }
}

// ds_k = sum_r O_kr dmh_r (transpose of the aggregation; O is symmetric) and
// dh_prev_k += ds_k P_l.
__global__ __launch_bounds__(kBlock) void k_lpc_agg_bwd(int nu, int cap, const int32_t* __restrict__ rowlen,
                                                        const int32_t* __restrict__ acol,
                                                        const float* __restrict__ aval, int l,
                                                        const uint16_t* __restrict__ w16,
                                                        const float* __restrict__ dmh, float* ds,
                                                        float* dh_prev)
{
    __shared__ float sP[LM * LH];
    for (int k = threadIdx.x; k < LM * LH; k += kBlock) sP[k] = bf2f(w16[off_P(l) + k]);
    __syncthreads();
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nu) return;
    float acc[LM];
#pragma unroll
    for (int q = 0; q < LM; ++q) acc[q] = 0.f;
    const long b = mx_index(r, 0, cap);
    for (int k = 0; k < rowlen[r]; ++k) {
        const float o = -aval[b + 64L * k];
        const v4f* p = reinterpret_cast<const v4f*>(dmh + LM * (long)acol[b + 64L * k]);
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            const v4f v = p[h];
            acc[4 * h + 0] = fmaf(o, v.x, acc[4 * h + 0]);
            acc[4 * h + 1] = fmaf(o, v.y, acc[4 * h + 1]);
            acc[4 * h + 2] = fmaf(o, v.z, acc[4 * h + 2]);
            acc[4 * h + 3] = fmaf(o, v.w, acc[4 * h + 3]);
        }
    }
#pragma unroll
    for (int q = 0; q < LM; ++q) ds[LM * (long)r + q] = acc[q];
    for (int k = 0; k < LH; ++k) {
        float v = dh_prev[LH * (long)r + k];
#pragma unroll
        for (int q = 0; q < LM; ++q) v = fmaf(acc[q], sP[q * LH + k], v);
        dh_prev[LH * (long)r + k] = v;
    }
}

// Partial sums over a block's rows of X_r (na) outer Y_r (nb): Y is bf16
// (YK 0), FP32 rounded to bf16 (YK 1) or all ones (YK 2, nb = 1).
template <int YK>
__global__ __launch_bounds__(kBlock) void k_lpc_outer(int nu, int na, int nb, const float* __restrict__ X,
                                                      const void* __restrict__ Yv, float* part)
{
    __shared__ float sX[32][LH];
    __shared__ float sY[32][LH];
    const int E = na * nb;
    constexpr int Q = kOuterMax / kBlock;
    float acc[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) acc[q] = 0.f;
    const int chunk = (nu + gridDim.x - 1) / gridDim.x;
    const int r0 = blockIdx.x * chunk, r1 = min(nu, r0 + chunk);
    for (int t0 = r0; t0 < r1; t0 += 32) {
        const int nt = min(32, r1 - t0);
        for (int k = threadIdx.x; k < 32 * na; k += kBlock) {
            const int rr = k / na, i = k % na;
            sX[rr][i] = rr < nt ? X[(long)(t0 + rr) * na + i] : 0.f;
        }
        if (YK != 2) {
            for (int k = threadIdx.x; k < 32 * nb; k += kBlock) {
                const int rr = k / nb, j = k % nb;
                float v = 0.f;
                if (rr < nt) {
                    if (YK == 0) v = bf2f(reinterpret_cast<const uint16_t*>(Yv)[(long)(t0 + rr) * nb + j]);
                    else v = bf2f(f2bf(reinterpret_cast<const float*>(Yv)[(long)(t0 + rr) * nb + j]));
                }
                sY[rr][j] = v;
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int e = threadIdx.x + q * kBlock;
            if (e < E) {
                const int i = e / nb, j = e % nb;
                float s = acc[q];
                for (int rr = 0; rr < nt; ++rr) s = fmaf(sX[rr][i], YK == 2 ? 1.f : sY[rr][j], s);
                acc[q] = s;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int e = threadIdx.x + q * kBlock;
        if (e < E) part[(long)blockIdx.x * E + e] = acc[q];
    }
}

__global__ void k_lpc_reduce(int E, int nblk, const float* __restrict__ part, float* out)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += (double)part[(long)b * E + e];
    out[e] = (float)s;
}

__global__ void k_lpc_adam(int n, float* w, const float* __restrict__ g, float* m, float* v, float lr,
                           float b1, float b2, float eps, float bc1, float bc2, uint16_t* w16)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.f - b1) * gk;
    const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    const float wk = w[k] - lr * (mk / bc1) / (sqrtf(vk / bc2) + eps);
    w[k] = wk;
    w16[k] = f2bf(wk);
}

__global__ void k_lpc_cast(int n, const float* __restrict__ w, uint16_t* w16)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) w16[k] = f2bf(w[k]);
}

}  // namespace vox

// ---- the network object -------------------------------------------------------
namespace {

struct LpcSys {
    int nu = 0, rows = 0, cap = 0;
    int32_t *rowlen = nullptr, *acol = nullptr, *slen = nullptr, *scol = nullptr;
    float *aval = nullptr, *diag = nullptr, *inv_d = nullptr, *inv_osum = nullptr, *feat = nullptr,
          *rhs = nullptr, *sval = nullptr;
};

}  // namespace

struct vox_lpc {
    int device = 0;
    hipStream_t stream = nullptr;
    float *w32 = nullptr, *grad = nullptr, *am = nullptr, *av = nullptr, *gpart = nullptr;
    uint16_t* w16 = nullptr;
    double *dpart = nullptr, *dscal = nullptr, *hscal = nullptr;
    std::vector<LpcSys> sys;
    int rows_cap = 0;
    uint16_t *h[4] = {}, *s[3] = {}, *m[3] = {};
    float *o = nullptr, *g = nullptr, *c = nullptr, *sc = nullptr;
    float *wv = nullptr, *t = nullptr, *z = nullptr, *e = nullptr, *y = nullptr, *dgdc = nullptr,
          *dO = nullptr, *dh[2] = {}, *dpre = nullptr, *dmh = nullptr, *ds = nullptr;
    long adam_t = 0;
};

namespace vox {

template <class T>
static int lalloc(T** p, size_t n)
{
    if (n == 0) n = 1;
    VOX_HIP(hipMalloc((void**)p, sizeof(T) * n));
    VOX_HIP(hipMemset(*p, 0, sizeof(T) * n));
    return VOX_OK;
}
template <class T>
static void lfree(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}
static int round64(int n) { return (n + 63) & ~63; }

// He-uniform weights from SplitMix64(seed) in the packed order; biases zero,
// the output layer scaled by 0.1 (outputs start near g = c = 1).
static void init_weights(uint64_t seed, std::vector<float>& w)
{
    w.assign(NPARAM, 0.f);
    uint64_t st = seed;
    auto next = [&]() {
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z = z ^ (z >> 31);
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    auto fill = [&](int off, int rows, int cols, double gain) {
        const double lim = gain * std::sqrt(6.0 / cols);
        for (int k = 0; k < rows * cols; ++k) w[off + k] = (float)((2.0 * next() - 1.0) * lim);
    };
    fill(OFF_W0, LH, LF, 1.0);
    for (int l = 1; l <= LNL; ++l) {
        fill(off_P(l), LM, LH, 1.0);
        fill(off_U(l), LH, LH, 1.0);
        fill(off_V(l), LH, LM, 1.0);
    }
    fill(OFF_WO, LO, LH, 0.1);
}

static int lpc_training_buffers(vox_lpc* L, int rows)
{
    if (rows <= L->rows_cap) return VOX_OK;
    for (auto*& p : L->h) lfree(p);
    for (auto*& p : L->s) lfree(p);
    for (auto*& p : L->m) lfree(p);
    for (float** p : {&L->o, &L->g, &L->c, &L->sc, &L->wv, &L->t, &L->z, &L->e, &L->y, &L->dgdc, &L->dO,
                      &L->dh[0], &L->dh[1], &L->dpre, &L->dmh, &L->ds})
        lfree(*p);
    const size_t n = (size_t)rows;
    for (auto*& p : L->h) VOX_TRY(lalloc(&p, n * LH));
    for (auto*& p : L->s) VOX_TRY(lalloc(&p, n * LM));
    for (auto*& p : L->m) VOX_TRY(lalloc(&p, n * LM));
    VOX_TRY(lalloc(&L->o, n * LO));
    for (float** p : {&L->g, &L->c, &L->sc, &L->wv, &L->t, &L->z, &L->e, &L->y}) VOX_TRY(lalloc(p, n));
    VOX_TRY(lalloc(&L->dgdc, n * 2));
    VOX_TRY(lalloc(&L->dO, n * 2));
    VOX_TRY(lalloc(&L->dh[0], n * LH));
    VOX_TRY(lalloc(&L->dh[1], n * LH));
    VOX_TRY(lalloc(&L->dpre, n * LH));
    VOX_TRY(lalloc(&L->dmh, n * LM));
    VOX_TRY(lalloc(&L->ds, n * LM));
    L->rows_cap = rows;
    return VOX_OK;
}

// Network forward over `rows` rows (multiple of 64) of a graph with nu rows;
// h[0..3], s[0..2], m[0..2] may alias as long as layer l's inputs differ from
// its outputs.
// Network forward over `rows` rows (multiple of 64) of a graph with nu rows;
// h[0..3], s[0..2], m[0..2] may alias as long as layer l's inputs differ from
// its outputs.
static int lpc_forward(hipStream_t st, const vox_lpc* L, int rows, AggArgs ag, const float* feat,
                       uint16_t* const h[4], uint16_t* const s[3], uint16_t* const m[3], float* o)
{
    const int ntile = rows / 16;
    const int nblk = std::max(1, std::min(2048, (ntile + 3) / 4));
    DenseArgs d{};
    d.rows = rows; d.w16 = L->w16; d.w32 = L->w32; d.feat = feat;
    d.layer = 0; d.hout = h[0]; d.sout = s[0]; d.o = o;
    k_lpc_dense<0><<<nblk, kBlock, 0, st>>>(d);
    for (int l = 1; l <= LNL; ++l) {
        ag.s = s[l - 1];
        ag.m = m[l - 1];
        k_lpc_agg<<<grid_for(ag.nu), kBlock, 0, st>>>(ag);
        d.layer = l; d.hin = h[l - 1]; d.min = m[l - 1]; d.hout = h[l];
        d.sout = (l < LNL) ? s[l] : nullptr;
        k_lpc_dense<1><<<nblk, kBlock, 0, st>>>(d);
    }
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

static int lpc_forward_sys(vox_lpc* L, const LpcSys& S)
{
    AggArgs ag{};
    ag.nu = S.nu; ag.cap = S.cap; ag.rowlen = S.rowlen; ag.acol = S.acol; ag.aval = S.aval;
    ag.inv_osum = S.inv_osum;
    VOX_TRY(lpc_forward(L->stream, L, S.rows, ag, S.feat, L->h, L->s, L->m, L->o));
    k_lpc_gc<<<grid_for(S.nu), kBlock, 0, L->stream>>>(S.nu, L->o, L->g, L->c, L->sc);
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

template <int YK>
static int outer(vox_lpc* L, int nu, int na, int nb, const float* X, const void* Y, int off)
{
    const int nblk = std::max(1, std::min(kOuterBlocks, (nu + 63) / 64));
    k_lpc_outer<YK><<<nblk, kBlock, 0, L->stream>>>(nu, na, nb, X, Y, L->gpart);
    const int E = na * nb;
    k_lpc_reduce<<<grid_for(E), kBlock, 0, L->stream>>>(E, nblk, L->gpart, L->grad + off);
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

// Loss (into dscal[LS_LOSS]) and gradient (into L->grad) on one system.
static int lpc_loss_grad_sys(vox_lpc* L, const LpcSys& S, uint64_t seed)
{
    hipStream_t st = L->stream;
    const int nu = S.nu;
    if (grid_for(nu) > kMaxPartials) return fail(VOX_ERR_ARG, "training system too large");
    VOX_TRY(lpc_forward_sys(L, S));
    k_lpc_set<<<1, 1, 0, st>>>(L->dscal, LS_LOSS, 0.0);
    const OpView SO{nu, S.cap, S.slen, S.scol, S.sval};
    const OpView AO{nu, S.cap, S.rowlen, S.acol, S.aval};
    const int nb = std::min(kMaxPartials, grid_for(nu));
    for (int p = 0; p < kProbes; ++p) {
        k_lpc_probe<<<grid_for(nu), kBlock, 0, st>>>(nu, p, seed, S.rhs, L->wv);
        k_lpc_dot<<<nb, kBlock, 0, st>>>(nu, L->wv, L->dpart);
        k_lpc_fin<<<1, kBlock, 0, st>>>(L->dpart, nb, L->dscal, 0);
        k_lpc_mprobe<<<grid_for(nu), kBlock, 0, st>>>(SO, L->sc, L->g, S.inv_d, L->wv, L->t, L->z);
        k_lpc_spmv<<<grid_for(nu), kBlock, 0, st>>>(AO, S.diag, L->z, L->wv, L->dscal, L->e, L->dpart);
        k_lpc_fin<<<1, kBlock, 0, st>>>(L->dpart, grid_for(nu), L->dscal, 1);
        k_lpc_spmv<<<grid_for(nu), kBlock, 0, st>>>(AO, S.diag, L->e, nullptr, L->dscal, L->y, nullptr);
        k_lpc_dgc<<<grid_for(nu), kBlock, 0, st>>>(SO, L->sc, S.inv_d, L->wv, L->t, L->y, L->dgdc, p == 0);
    }
    k_lpc_head_bwd<<<grid_for(nu), kBlock, 0, st>>>(nu, L->w32, L->o, L->g, L->dgdc, L->dO, L->dh[0]);
    VOX_HIP(hipGetLastError());
    VOX_TRY(outer<0>(L, nu, LO, LH, L->dO, L->h[LNL], OFF_WO));
    VOX_TRY(outer<2>(L, nu, LO, 1, L->dO, nullptr, OFF_BO));
    int cur = 0;
    for (int l = LNL; l >= 1; --l) {
        float* dh = L->dh[cur];
        float* dhp = L->dh[cur ^ 1];
        k_lpc_dense_bwd<<<grid_for(nu), kBlock, 0, st>>>(nu, l, L->w16, L->h[l], dh, S.inv_osum, L->dpre,
                                                         L->dmh, dhp);
        k_lpc_agg_bwd<<<grid_for(nu), kBlock, 0, st>>>(nu, S.cap, S.rowlen, S.acol, S.aval, l, L->w16,
                                                       L->dmh, L->ds, dhp);
        VOX_HIP(hipGetLastError());
        VOX_TRY(outer<0>(L, nu, LH, LH, L->dpre, L->h[l - 1], off_U(l)));
        VOX_TRY(outer<0>(L, nu, LH, LM, L->dpre, L->m[l - 1], off_V(l)));
        VOX_TRY(outer<2>(L, nu, LH, 1, L->dpre, nullptr, off_c(l)));
        VOX_TRY(outer<0>(L, nu, LM, LH, L->ds, L->h[l - 1], off_P(l)));
        cur ^= 1;
    }
    k_lpc_dense_bwd<<<grid_for(nu), kBlock, 0, st>>>(nu, 0, L->w16, L->h[0], L->dh[cur], S.inv_osum,
                                                     L->dpre, nullptr, nullptr);
    VOX_HIP(hipGetLastError());
    VOX_TRY(outer<1>(L, nu, LH, LF, L->dpre, S.feat, OFF_W0));
    VOX_TRY(outer<2>(L, nu, LH, 1, L->dpre, nullptr, OFF_B0));
    return VOX_OK;
}

static int check_lpc(vox_lpc* L)
{
    if (!L) return fail(VOX_ERR_ARG, "null learned preconditioner");
    VOX_HIP(hipSetDevice(L->device));
    return VOX_OK;
}

static int check_index(vox_lpc* L, int index)
{
    if (index < 0 || index >= (int)L->sys.size()) return fail(VOX_ERR_ARG, "no stored system of that index");
    return VOX_OK;
}

// ---- inference for a context ------------------------------------------------
void lpc_free_ctx(vox_ctx* c)
{
    lfree(c->lpc_feat); lfree(c->lpc_inv_osum); lfree(c->lpc_t); lfree(c->lpc_invd);
    lfree(c->lpc_upos);
    lfree(c->lpc_h[0]); lfree(c->lpc_h[1]); lfree(c->lpc_s); lfree(c->lpc_m); lfree(c->lpc_o);
    lfree(c->lpc_gd); lfree(c->lpc_slen); lfree(c->lpc_scol);
    lfree(c->lpc_sval); lfree(c->cg_z); lfree(c->cg_z32);
    c->lpc_rows = 0;
    c->lpc_scap = 0;
}

// Strong-coupling threshold: theta times the coupling of two particles one
// spacing dl apart, ((2 dim / lambda) W(dl) / n0) / rho1 as the assembly forms
// it (W of functions.cu:32-70).
static double lpc_threshold(const vox_ctx* c)
{
    const vox_params& p = c->p;
    const double q = p.dl / p.re_l;
    double w = 0.0;
    if (p.ktype == 2) w = (q < 1.0) ? (1 / q) - 1 : 0.0;
    else if (p.ktype == 6) w = (q <= 1.0) ? (1 - q) * (1 - q) * (1 - q) : 0.0;
    else if (p.ktype == 1) w = (q < 0.5) ? 2.0 - 4.0 * (q * q) : (q <= 1.0 ? (2 * q - 2) * (2 * q - 2) : 0.0);
    else if (p.ktype == 5) w = (q <= 1.0) ? 1.5 * std::log(1 / (q + 0.000001)) : 0.0;
    return 0.25 * ((2.0 * p.dim / p.lambda * w / c->n0l) / p.rho1);
}

// Features and raw strong couplings of the context's assembled PPE (FP32
// values): feat (rows x 8), inv_osum, invd, slen/scol/sval; upos is scratch.
static void ctx_features(vox_ctx* c, int rows, v4f* upos, float* invd, float* feat, float* inv_osum,
                         int32_t* slen, int32_t* scol, float* sval)
{
    PrepArgs pa;
    pa.nu = c->nu; pa.rows = rows; pa.dim = c->p.dim; pa.n0 = c->n0l;
    pa.adiag = c->adiag; pa.urow = c->urow; pa.ptype = c->ptype;
    pa.x = c->f[VOX_F_X]; pa.y = c->f[VOX_F_Y]; pa.z = c->f[VOX_F_Z]; pa.pnd = c->f[VOX_F_N];
    pa.upos = upos; pa.invd = invd; pa.feat = feat;
    k_lpc_rowprep<<<grid_for(rows), kBlock, 0, c->stream>>>(pa);
    FeatArgs fa;
    fa.nu = c->nu; fa.cap = c->a_cap; fa.inv_re = (float)(1.0 / c->p.re_l); fa.thr = lpc_threshold(c);
    fa.rowlen = c->rowlen; fa.acol = c->acol; fa.aval = c->aval32; fa.upos = upos; fa.invd = invd;
    fa.feat = feat; fa.inv_osum = inv_osum; fa.slen = slen; fa.scol = scol; fa.sval = sval;
    k_lpc_features<<<grid_for(c->nu), kBlock, 0, c->stream>>>(fa);
}

int lpc_prepare(vox_ctx* c)
{
    VOX_TRY(ensure_acol(c));
    if (!c->lpc) return fail(VOX_ERR_STATE, "learned solver without an attached vox_lpc");
    if (!c->aval32_ok) VOX_TRY(ppe_make_aval32(c));
    ProfScope ps(c, KC_LPC_NET);
    const int nu = c->nu;
    const int rows = round64(c->tp);
    if (c->lpc_rows < rows) {
        lpc_free_ctx(c);
        const size_t n = (size_t)rows;
        VOX_TRY(lalloc(&c->lpc_feat, n * LF));
        VOX_TRY(lalloc(&c->lpc_inv_osum, n));
        VOX_TRY(lalloc(&c->lpc_t, n));
        VOX_TRY(lalloc(&c->lpc_invd, n));
        VOX_TRY(lalloc(&c->lpc_upos, n));
        VOX_TRY(lalloc(&c->lpc_h[0], n * LH));
        VOX_TRY(lalloc(&c->lpc_h[1], n * LH));
        VOX_TRY(lalloc(&c->lpc_s, n * LM));
        VOX_TRY(lalloc(&c->lpc_m, n * LM));
        VOX_TRY(lalloc(&c->lpc_o, n * LO));
        VOX_TRY(lalloc(&c->lpc_gd, n));
        VOX_TRY(lalloc(&c->lpc_slen, n));
        VOX_TRY(lalloc(&c->cg_z, n));
        VOX_TRY(lalloc(&c->cg_z32, n));
        c->lpc_rows = rows;
    }
    if (c->lpc_scap != c->a_cap) {
        lfree(c->lpc_scol);
        lfree(c->lpc_sval);
        VOX_TRY(lalloc(&c->lpc_scol, (size_t)sl_size(c->tp, c->a_cap)));
        VOX_TRY(lalloc(&c->lpc_sval, (size_t)sl_size(c->tp, c->a_cap)));
        c->lpc_scap = c->a_cap;
    }
    if (nu == 0) return VOX_OK;
    const int frows = round64(nu);
    ctx_features(c, frows, c->lpc_upos, c->lpc_invd, c->lpc_feat, c->lpc_inv_osum, c->lpc_slen,
                 c->lpc_scol, c->lpc_sval);
    AggArgs ag{};
    ag.nu = nu; ag.cap = c->a_cap; ag.rowlen = c->rowlen; ag.acol = c->acol; ag.aval = c->aval32;
    ag.inv_osum = c->lpc_inv_osum;
    uint16_t* const h[4] = {c->lpc_h[0], c->lpc_h[1], c->lpc_h[0], c->lpc_h[1]};
    uint16_t* const s[3] = {c->lpc_s, c->lpc_s, c->lpc_s};
    uint16_t* const m[3] = {c->lpc_m, c->lpc_m, c->lpc_m};
    VOX_TRY(lpc_forward(c->stream, c->lpc, frows, ag, c->lpc_feat, h, s, m, c->lpc_o));
    k_lpc_gc_inf<<<grid_for(nu), kBlock, 0, c->stream>>>(nu, c->lpc_o, c->lpc_invd, c->lpc_gd, c->lpc_t);
    k_lpc_scale<<<grid_for(nu), kBlock, 0, c->stream>>>(nu, c->a_cap, c->lpc_slen, c->lpc_scol, c->lpc_t,
                                                        c->lpc_sval);
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

}  // namespace vox

using namespace vox;

extern "C" {

int vox_lpc_create(vox_lpc** out, uint64_t seed)
{
    if (!out) return fail(VOX_ERR_ARG, "null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(VOX_ERR_HIP, "no HIP device available");
    vox_lpc* L = new vox_lpc();
    auto bail = [&](int rc) { vox_lpc_destroy(L); return rc; };
    if (hipGetDevice(&L->device) != hipSuccess) return bail(fail(VOX_ERR_HIP, "hipGetDevice"));
    if (hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(VOX_ERR_HIP, "hipStreamCreate"));
    int rc;
    if ((rc = lalloc(&L->w32, NPARAM)) || (rc = lalloc(&L->grad, NPARAM)) || (rc = lalloc(&L->am, NPARAM)) ||
        (rc = lalloc(&L->av, NPARAM)) || (rc = lalloc(&L->w16, NPARAM)) ||
        (rc = lalloc(&L->gpart, (size_t)kOuterBlocks * kOuterMax)) ||
        (rc = lalloc(&L->dpart, (size_t)kMaxPartials)) || (rc = lalloc(&L->dscal, LS_COUNT)))
        return bail(rc);
    if (hipHostMalloc((void**)&L->hscal, sizeof(double) * LS_COUNT, hipHostMallocDefault) != hipSuccess)
        return bail(fail(VOX_ERR_HIP, "hipHostMalloc"));
    std::vector<float> w;
    init_weights(seed, w);
    if ((rc = vox_lpc_set_weights(L, w.data()))) return bail(rc);
    *out = L;
    return VOX_OK;
}

int vox_lpc_destroy(vox_lpc* L)
{
    if (!L) return VOX_OK;
    (void)hipSetDevice(L->device);
    if (L->stream) (void)hipStreamSynchronize(L->stream);
    for (auto& S : L->sys) {
        lfree(S.rowlen); lfree(S.acol); lfree(S.slen); lfree(S.scol); lfree(S.aval); lfree(S.diag);
        lfree(S.inv_d); lfree(S.inv_osum); lfree(S.feat); lfree(S.rhs); lfree(S.sval);
    }
    for (auto*& p : L->h) lfree(p);
    for (auto*& p : L->s) lfree(p);
    for (auto*& p : L->m) lfree(p);
    for (float** p : {&L->o, &L->g, &L->c, &L->sc, &L->wv, &L->t, &L->z, &L->e, &L->y, &L->dgdc, &L->dO,
                      &L->dh[0], &L->dh[1], &L->dpre, &L->dmh, &L->ds, &L->w32, &L->grad, &L->am, &L->av,
                      &L->gpart})
        lfree(*p);
    lfree(L->w16); lfree(L->dpart); lfree(L->dscal);
    if (L->hscal) (void)hipHostFree(L->hscal);
    if (L->stream) (void)hipStreamDestroy(L->stream);
    delete L;
    return VOX_OK;
}

int vox_lpc_get_weights(vox_lpc* L, float* w)
{
    VOX_TRY(check_lpc(L));
    if (!w) return fail(VOX_ERR_ARG, "null weights");
    VOX_HIP(hipMemcpyAsync(w, L->w32, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, L->stream));
    VOX_HIP(hipStreamSynchronize(L->stream));
    return VOX_OK;
}

int vox_lpc_set_weights(vox_lpc* L, const float* w)
{
    VOX_TRY(check_lpc(L));
    if (!w) return fail(VOX_ERR_ARG, "null weights");
    for (int k = 0; k < NPARAM; ++k)
        if (!std::isfinite(w[k])) return fail(VOX_ERR_ARG, "non-finite weight");
    VOX_HIP(hipMemcpyAsync(L->w32, w, sizeof(float) * NPARAM, hipMemcpyHostToDevice, L->stream));
    k_lpc_cast<<<grid_for(NPARAM), kBlock, 0, L->stream>>>(NPARAM, L->w32, L->w16);
    VOX_HIP(hipGetLastError());
    VOX_HIP(hipStreamSynchronize(L->stream));
    // a fresh set of weights restarts the optimiser
    VOX_HIP(hipMemset(L->am, 0, sizeof(float) * NPARAM));
    VOX_HIP(hipMemset(L->av, 0, sizeof(float) * NPARAM));
    L->adam_t = 0;
    return VOX_OK;
}

int vox_lpc_dump(vox_lpc* L, vox_ctx* c, int32_t* index)
{
    VOX_TRY(check_lpc(L));
    if (!c) return fail(VOX_ERR_ARG, "null context");
    if (c->device != L->device) return fail(VOX_ERR_ARG, "context and network on different devices");
    if (!c->acol || c->nu <= 0) return fail(VOX_ERR_STATE, "no assembled PPE in the context");
    if (c->uorder_cells) {
        // training systems are dumped with the unknowns in particle order, the
        // order of vox_ppe_export: the system is assembled again in that order
        c->uorder_particles = true;
        const int rc0 = ppe_assemble(c);
        c->uorder_particles = false;
        VOX_TRY(rc0);
    }
    if (!c->aval32_ok) VOX_TRY(ppe_make_aval32(c));
    VOX_TRY(ensure_acol(c));
    LpcSys S;
    S.nu = c->nu; S.rows = round64(c->nu); S.cap = c->a_cap;
    const size_t nn = (size_t)S.rows, mat = (size_t)sl_size(S.nu, S.cap);
    int rc;
    v4f* upos = nullptr;
    auto cleanup = [&]() {
        lfree(S.rowlen); lfree(S.acol); lfree(S.slen); lfree(S.scol); lfree(S.aval); lfree(S.diag);
        lfree(S.inv_d); lfree(S.inv_osum); lfree(S.feat); lfree(S.rhs); lfree(S.sval); lfree(upos);
    };
    if ((rc = lalloc(&S.rowlen, nn)) || (rc = lalloc(&S.acol, mat)) || (rc = lalloc(&S.aval, mat)) ||
        (rc = lalloc(&S.slen, nn)) || (rc = lalloc(&S.scol, mat)) || (rc = lalloc(&S.sval, mat)) ||
        (rc = lalloc(&S.diag, nn)) || (rc = lalloc(&S.inv_d, nn)) || (rc = lalloc(&S.inv_osum, nn)) ||
        (rc = lalloc(&S.feat, nn * LF)) || (rc = lalloc(&S.rhs, nn)) || (rc = lalloc(&upos, nn))) {
        cleanup();
        return rc;
    }
    hipStream_t st = c->stream;
    if (hipMemcpyAsync(S.rowlen, c->rowlen, sizeof(int32_t) * S.nu, hipMemcpyDeviceToDevice, st) != hipSuccess ||
        hipMemcpyAsync(S.acol, c->acol, sizeof(int32_t) * mat, hipMemcpyDeviceToDevice, st) != hipSuccess ||
        hipMemcpyAsync(S.aval, c->aval32, sizeof(float) * mat, hipMemcpyDeviceToDevice, st) != hipSuccess) {
        cleanup();
        return fail(VOX_ERR_HIP, "dump copy");
    }
    k_lpc_dump_rows<<<grid_for(S.nu), kBlock, 0, st>>>(S.nu, c->adiag, c->rhs, S.diag, S.rhs);
    ctx_features(c, S.rows, upos, S.inv_d, S.feat, S.inv_osum, S.slen, S.scol, S.sval);
    k_lpc_oh<<<grid_for(S.nu), kBlock, 0, st>>>(S.nu, S.cap, S.slen, S.scol, S.inv_d, S.sval);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
        cleanup();
        return fail(VOX_ERR_HIP, "dump kernels");
    }
    lfree(upos);
    if ((rc = lpc_training_buffers(L, S.rows))) {
        cleanup();
        return rc;
    }
    L->sys.push_back(S);
    if (index) *index = (int32_t)L->sys.size() - 1;
    return VOX_OK;
}

int vox_lpc_count(vox_lpc* L, int32_t* systems, int64_t* rows)
{
    VOX_TRY(check_lpc(L));
    if (systems) *systems = (int32_t)L->sys.size();
    if (rows) {
        int64_t t = 0;
        for (auto& S : L->sys) t += S.nu;
        *rows = t;
    }
    return VOX_OK;
}

int vox_lpc_eval(vox_lpc* L, int32_t index, float* feat, float* g, float* c)
{
    VOX_TRY(check_lpc(L));
    VOX_TRY(check_index(L, index));
    const LpcSys& S = L->sys[index];
    VOX_TRY(lpc_forward_sys(L, S));
    if (feat) VOX_HIP(hipMemcpyAsync(feat, S.feat, sizeof(float) * LF * S.nu, hipMemcpyDeviceToHost, L->stream));
    if (g) VOX_HIP(hipMemcpyAsync(g, L->g, sizeof(float) * S.nu, hipMemcpyDeviceToHost, L->stream));
    if (c) VOX_HIP(hipMemcpyAsync(c, L->c, sizeof(float) * S.nu, hipMemcpyDeviceToHost, L->stream));
    VOX_HIP(hipStreamSynchronize(L->stream));
    return VOX_OK;
}

int vox_lpc_loss_grad(vox_lpc* L, int32_t index, uint64_t seed, double* loss, float* grad)
{
    VOX_TRY(check_lpc(L));
    VOX_TRY(check_index(L, index));
    VOX_TRY(lpc_loss_grad_sys(L, L->sys[index], seed));
    VOX_HIP(hipMemcpyAsync(L->hscal, L->dscal, sizeof(double) * LS_COUNT, hipMemcpyDeviceToHost, L->stream));
    if (grad) VOX_HIP(hipMemcpyAsync(grad, L->grad, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, L->stream));
    VOX_HIP(hipStreamSynchronize(L->stream));
    if (loss) *loss = L->hscal[LS_LOSS];
    return VOX_OK;
}

int vox_lpc_train(vox_lpc* L, int32_t epochs, double lr, uint64_t seed, double* loss_first, double* loss_last)
{
    VOX_TRY(check_lpc(L));
    if (epochs < 1 || !(lr > 0)) return fail(VOX_ERR_ARG, "epochs >= 1 and lr > 0 required");
    const int ns = (int)L->sys.size();
    if (ns == 0) return fail(VOX_ERR_STATE, "no stored systems to train on");
    std::vector<int> order(ns);
    std::iota(order.begin(), order.end(), 0);
    uint64_t st = seed ^ 0xD1B54A32D192ED03ull;
    auto next = [&]() {
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int ep = 0; ep < epochs; ++ep) {
        for (int k = ns - 1; k > 0; --k) std::swap(order[k], order[(int)(next() % (uint64_t)(k + 1))]);
        k_lpc_set<<<1, 1, 0, L->stream>>>(L->dscal, LS_EPOCH, 0.0);
        for (int k = 0; k < ns; ++k) {
            const int i = order[k];
            VOX_TRY(lpc_loss_grad_sys(L, L->sys[i], seed + (uint64_t)ep * 1000003ull + (uint64_t)i));
            k_lpc_fin<<<1, kBlock, 0, L->stream>>>(L->dpart, 0, L->dscal, 2);
            ++L->adam_t;
            const float bc1 = 1.f - std::pow(b1, (float)L->adam_t);
            const float bc2 = 1.f - std::pow(b2, (float)L->adam_t);
            k_lpc_adam<<<grid_for(NPARAM), kBlock, 0, L->stream>>>(NPARAM, L->w32, L->grad, L->am, L->av,
                                                                   (float)lr, b1, b2, eps, bc1, bc2, L->w16);
            VOX_HIP(hipGetLastError());
        }
        VOX_HIP(hipMemcpyAsync(L->hscal, L->dscal, sizeof(double) * LS_COUNT, hipMemcpyDeviceToHost, L->stream));
        VOX_HIP(hipStreamSynchronize(L->stream));
        const double mean = L->hscal[LS_EPOCH] / ns;
        if (!std::isfinite(mean)) return fail(VOX_ERR_SOLVER, "training loss is not finite");
        if (ep == 0 && loss_first) *loss_first = mean;
        if (loss_last) *loss_last = mean;
    }
    return VOX_OK;
}

int vox_attach_lpc(vox_ctx* c, vox_lpc* L)
{
    if (!c) return fail(VOX_ERR_ARG, "null context");
    if (L && L->device != c->device) return fail(VOX_ERR_ARG, "context and network on different devices");
    c->lpc = L;
    c->lpc_baseline = 0;
    return VOX_OK;
}

}  // extern "C"
