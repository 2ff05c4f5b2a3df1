// vox_cg.hip — Jacobi-preconditioned CG for the assembled PPE on gfx950, in
// FP64 or as FP64 iterative refinement around an FP32 inner solve.
//
// Replaces INCDECOM + CGM (functions.cu:1139-1200, 506-716) and the GPU loop of
// main.cu:1159-1261 (which synchronised with the host and used FP32 atomics for
// every dot product). One CG iteration is three kernels on a fixed grid:
//   K1 q = A p                         partial p.q per block
//   K2 alpha = rz/p.q; x += alpha p; r -= alpha q
//                                      partials r.r and r.(r/d) per block
//   K3 beta = rz'/rz; convergence test; p = r/d + beta p
// There are no atomics: each kernel writes one partial per block and every
// block of the next kernel sums the partials in the same fixed order, so all
// blocks derive bit-identical scalars and results are run-to-run identical.
// rz is double-buffered by iteration parity (K3 of iteration k writes slot
// k&1 while its blocks read slot (k-1)&1). The stop flag stays on the device;
// kernels launched after convergence return at entry.
//
// The kernels are templated on the vector type VT and the matrix value type
// MT; dot products always accumulate in FP64. VT = MT = double is the FP64
// solver. In mixed precision the outer loop keeps x, the residual b - A x and
// its norm in FP64 on the FP64 matrix, and each correction solves A d = r with
// FP32 matrix values and FP32 vectors (FP64 dots) to a relative eta.
//
// The matrix is SELL-64 (mx_index); the product runs one thread per row.
// Its default form (k_cg_spmv_win) stages p at each 256-row block's distinct
// columns in LDS (the windows k_win_build lists once per assembly) and reads
// each entry's 16-bit rank in that window instead of its column; the other
// forms read the columns as 16-bit steps (dcol) or 32-bit ids (acol). Every
// form sums each row in the same order: the same q, bit for bit.
//
// Flexible CG (solver VOX_SOLVER_LEARNED_FCG) takes z = M r from the learned
// coupling operator (vox_lpc.hip) instead of r / d, and one more kernel per
// iteration:
//   K1 q = A p, partial p.q           (the product above)
//   K2 x += alpha p; r -= alpha q     partial r.r
//   K3 convergence test; z = M r      partials z.r and z.q
//   K4 beta = -alpha z.q / rz (Polak-Ribiere form, since r_new - r_old =
//      -alpha q); p = z + beta p
// which stays a conjugate-direction method when M varies between solves or
// is not exactly symmetric positive definite.
#include <type_traits>

#include "vox_internal.hpp"

namespace vox {

// The vector kernels (everything but the products and the preconditioner):
// 512 blocks of 1024 threads, 8192 waves (8 per SIMD on 256 CUs). Against
// 2048 blocks of 256 threads, each kernel's blocks read the previous
// kernel's partials a quarter as often, and there are a quarter as many.
constexpr int kVecThreads = 1024;
constexpr int kVecBlocks = 512;

// Fixed-order sum over a block of NT threads (LDS tree of NT doubles).
template <int NT>
__device__ __forceinline__ double block_sum_n(double v, double* lds)
{
    const int t = threadIdx.x;
    lds[t] = v;
    __syncthreads();
#pragma unroll
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (t < s) lds[t] = lds[t] + lds[t + s];
        __syncthreads();
    }
    const double r = lds[0];
    __syncthreads();
    return r;
}

template <int NT>
__device__ __forceinline__ double sum_partials_n(const double* part, int n, double* lds)
{
    double s = 0.0;
    for (int k = threadIdx.x; k < n; k += NT) s = s + part[k];
    return block_sum_n<NT>(s, lds);
}

// partial slots (kMaxPartials doubles each)
enum { P_BB = 0, P_RZ0, P_NNZ, P_PQ, P_RR, P_RZ, P_RES, P_RHS, P_ZQ, P_COUNT };
// scalar slots
enum { S_BNORM = 0, S_RZ0, S_RZ1, S_RELRES, S_NNZ, S_RESN, S_RHSN, S_COUNT_ };
static_assert(S_COUNT_ <= SC_COUNT, "scalar slots");
static_assert(P_COUNT <= kPartialSlots, "partial slots");

template <class VT, class MT>
struct CgArgs {
    int nu, cap, maxit, it, npq;   // npq: blocks of the product kernel
    double tol;
    const int32_t* __restrict__ rowlen;
    const int32_t* __restrict__ acol;
    const v4u* __restrict__ dcol;     // the columns as 16-bit steps (null: not valid)
    const int32_t* __restrict__ cbase;
    const v4u* __restrict__ wcol;     // ranks in the chunk windows (null: not built)
    const int32_t* __restrict__ wuniq;
    const int32_t* __restrict__ wn;
    int win_n;
    const MT* __restrict__ aval;
    const double* __restrict__ diag;
    const double* __restrict__ rhs;   // FP64 right-hand side of this solve
    VT* x;
    VT* r;
    VT* p;
    VT* q;
    VT* z;                            // flexible CG: z = M r
    const int32_t* __restrict__ slen; // coupling operator (SELL-64, cap scap)
    const int32_t* __restrict__ scol;
    const float* __restrict__ sval;
    const float* __restrict__ gd;     // g_r / A_rr
    int scap;
    double* part;
    double* scal;
    int32_t* flags;
};

#define PSLOT(a, s) ((a).part + (long)(s) * kMaxPartials)

// Sum of the first n partials of one slot, identical in every block.
__device__ __forceinline__ double sum_partials(const double* part, int n, double* lds)
{
    double s = 0.0;
    for (int k = threadIdx.x; k < n; k += kBlock) s = s + part[k];
    return block_sum(s, lds);
}

template <class VT, class MT>
__global__ __launch_bounds__(kVecThreads) void k_cg_init(CgArgs<VT, MT> a)
{
    __shared__ double lds[kVecThreads];
    double bb = 0.0, rz = 0.0, nz = 0.0;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < a.nu; r += gridDim.x * blockDim.x) {
        const double b = a.rhs[r];
        const VT rv = (VT)b;
        const VT z = rv / (VT)a.diag[r];
        a.x[r] = (VT)0;
        a.r[r] = rv;
        a.p[r] = z;
        bb += b * b;
        rz += (double)rv * (double)z;
        nz += (double)(a.rowlen[r] + 1);
    }
    bb = block_sum_n<kVecThreads>(bb, lds);
    rz = block_sum_n<kVecThreads>(rz, lds);
    nz = block_sum_n<kVecThreads>(nz, lds);
    if (threadIdx.x == 0) {
        PSLOT(a, P_BB)[blockIdx.x] = bb;
        PSLOT(a, P_RZ0)[blockIdx.x] = rz;
        PSLOT(a, P_NNZ)[blockIdx.x] = nz;
    }
}

template <class VT, class MT>
__global__ __launch_bounds__(kBlock) void k_cg_init_fin(CgArgs<VT, MT> a)
{
    __shared__ double lds[kBlock];
    double bb = sum_partials(PSLOT(a, P_BB), kVecBlocks, lds);
    double rz = sum_partials(PSLOT(a, P_RZ0), kVecBlocks, lds);
    double nz = sum_partials(PSLOT(a, P_NNZ), kVecBlocks, lds);
    if (threadIdx.x == 0) {
        double bn = sqrt(bb);
        a.scal[S_BNORM] = bn;
        a.scal[S_RZ0] = rz;
        a.scal[S_NNZ] = nz;
        a.scal[S_RELRES] = 0.0;
        a.flags[FLAG_CG_ITERS] = 0;
        a.flags[FLAG_CG_BREAK] = 0;
        a.flags[FLAG_CG_CONV] = (bn == 0.0) ? 1 : 0;
        a.flags[FLAG_CG_DONE] = (bn == 0.0) ? 1 : 0;
    }
}

// Row products over the SELL-64 matrix, entries in row order after the
// diagonal: acc = d_r p_r + sum_k a_rk p_col(rk). A wave's entry k of its 64
// rows is one contiguous block; 8 entries' loads are in flight per lane. The
// matrix streams from HBM once per product (non-temporal loads keep it from
// evicting the gathered vector from L2).

// q = A p and the partial p.q of each block; one thread per row. U: entries
// whose loads are issued together (the sum is in row order for every U, so
// the autotuner's choice does not change any bit). kD: the columns come from
// their 16-bit steps (U = 8: one 16-byte chunk per lane per batch, 10 instead
// of 12 bytes per FP64 entry and one column load per 8 entries, not 8).
template <class VT, class MT, int U = 8, bool kD = false>
__global__ __launch_bounds__(kBlock) void k_cg_spmv(CgArgs<VT, MT> a)
{
    static_assert(!kD || U == 8, "column steps come in chunks of 8");
    __shared__ double lds[kBlock];
    if (a.flags[FLAG_CG_DONE]) return;
    double pq = 0.0;
    const int r = xcd_tid();
    if (kD && r < a.nu) {
        const VT pr = a.p[r];
        VT acc = (VT)a.diag[r] * pr;
        const int len = a.rowlen[r];
        const v4u* __restrict__ dc = a.dcol + dc_index(r, 0, a.cap);   // chunk g at dc[64 g]
        const MT* __restrict__ val = a.aval + mx_index(r, 0, a.cap);
        int cj = a.cbase[r];
        int k = 0;
        for (; k + 8 <= len; k += 8) {
            const v4u w = __builtin_nontemporal_load(&dc[8L * k]);
            MT vv[8];
            VT pp[8];
            int cc[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) vv[u] = __builtin_nontemporal_load(&val[64L * (k + u)]);
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                cj += (int)(w[h] & 0xffffu);
                cc[2 * h] = cj;
                cj += (int)(w[h] >> 16);
                cc[2 * h + 1] = cj;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) pp[u] = a.p[cc[u]];
#pragma unroll
            for (int u = 0; u < 8; ++u) acc = acc + (VT)vv[u] * pp[u];
        }
        if (k < len) {
            const v4u w = __builtin_nontemporal_load(&dc[8L * k]);
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                if (k + u < len) {
                    cj += (int)((w[u >> 1] >> (16 * (u & 1))) & 0xffffu);
                    acc = acc + (VT)val[64L * (k + u)] * a.p[cj];
                }
            }
        }
        a.q[r] = acc;
        pq = (double)pr * (double)acc;
    } else if (!kD && r < a.nu) {
        // (all 16 streaming loads of a step are issued before the first
        // gather: keep this loop's shape, other forms scheduled the gathers
        // earlier and measured about 18% slower; 8 entries per step measured
        // 11% faster than 16 and 4)
        const VT pr = a.p[r];
        VT acc = (VT)a.diag[r] * pr;
        const int len = a.rowlen[r];
        const int32_t* __restrict__ col = a.acol + mx_index(r, 0, a.cap);
        const MT* __restrict__ val = a.aval + mx_index(r, 0, a.cap);
        int k = 0;
        for (; k + U <= len; k += U) {
            int32_t cc[U];
            MT vv[U];
            VT pp[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                cc[u] = __builtin_nontemporal_load(&col[64L * (k + u)]);
                vv[u] = __builtin_nontemporal_load(&val[64L * (k + u)]);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) pp[u] = a.p[cc[u]];
#pragma unroll
            for (int u = 0; u < U; ++u) acc = acc + (VT)vv[u] * pp[u];
        }
        for (; k < len; ++k) acc = acc + (VT)val[64L * k] * a.p[col[64L * k]];
        a.q[r] = acc;
        pq = (double)pr * (double)acc;
    }
    pq = block_sum(pq, lds);
    if (threadIdx.x == 0) PSLOT(a, P_PQ)[blockIdx.x] = pq;
}

// Step or rank u (0..7) of a 16-byte chunk of 16-bit values.
__device__ __forceinline__ int u16_of(const v4u& w, int u)
{
    return (int)((w[u >> 1] >> (16 * (u & 1))) & 0xffffu);
}

// The same product with p gathered from LDS: each block (one chunk of 256
// rows, as in k_cg_spmv) first stages p at the chunk's distinct columns
// (wuniq, ascending, so mostly whole lines) into a window in LDS, then reads
// each entry's value and its 16-bit rank in that window. The gathers from
// global memory, one cache-line lookup per distinct line touched by each
// wave-wide load, become one coalesced staging pass per block; the row sums
// are the same operations in the same order, so q and p.q are the same bits.
template <class VT, class MT>
__global__ __launch_bounds__(kBlock) void k_cg_spmv_win(CgArgs<VT, MT> a)
{
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    VT* win = reinterpret_cast<VT*>(dyn_lds);
    __shared__ double lds[kBlock];
    if (a.flags[FLAG_CG_DONE]) return;
    const int ch = xcd_block();
    const int n = a.wn[ch];
    const int32_t* __restrict__ uq = a.wuniq + (long)ch * kWinMax;
    for (int t = threadIdx.x; t < n; t += kBlock) win[t] = a.p[uq[t]];
    __syncthreads();
    double pq = 0.0;
    const int r = ch * kBlock + threadIdx.x;
    if (r < a.nu) {
        // (issuing the first batch before the staging, and each batch before
        // the previous one is consumed, measured the same)
        const VT pr = a.p[r];
        VT acc = (VT)a.diag[r] * pr;
        const int len = a.rowlen[r];
        const v4u* __restrict__ wc = a.wcol + dc_index(r, 0, a.cap);
        const MT* __restrict__ val = a.aval + mx_index(r, 0, a.cap);
        int k = 0;
        for (; k + 8 <= len; k += 8) {
            const v4u w = __builtin_nontemporal_load(&wc[8L * k]);
            MT vv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) vv[u] = __builtin_nontemporal_load(&val[64L * (k + u)]);
#pragma unroll
            for (int u = 0; u < 8; ++u) acc = acc + (VT)vv[u] * win[u16_of(w, u)];
        }
        if (k < len) {
            const v4u w = __builtin_nontemporal_load(&wc[8L * k]);
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (k + u < len) acc = acc + (VT)val[64L * (k + u)] * win[u16_of(w, u)];
        }
        a.q[r] = acc;
        pq = (double)pr * (double)acc;
    }
    pq = block_sum(pq, lds);
    if (threadIdx.x == 0) PSLOT(a, P_PQ)[blockIdx.x] = pq;
}

struct WinArgs {
    int nu, cap, words;        // words: bitmap words in the dynamic LDS
    const int32_t* __restrict__ rowlen;
    const v4u* __restrict__ dcol;
    const int32_t* __restrict__ cbase;
    v4u* wcol;
    int32_t* wuniq;
    int32_t* wn;
    int32_t* flags;
};

// The windows of the LDS-gather product, one block per chunk of 256 rows
// (rows ascending): a bitmap of the columns the rows use, based at the
// chunk's smallest first column (in dynamic LDS, `words` words, sized from
// the assembly's bound on how far the rows reach below and above their own
// index), the prefix of its popcounts per word, the distinct columns in
// ascending order and each entry's rank among them. A chunk with a column
// past the bitmap or more distinct columns than kWinMax raises FLAG_WIN_N
// past kWinMax, and the product then reads the column steps. Each lane walks
// its row 4 chunks (32 entries) per round trip.
__global__ __launch_bounds__(kBlock) void k_win_build(WinArgs a)
{
    extern __shared__ __align__(16) uint32_t wlds[];
    uint32_t* bits = wlds;
    uint16_t* pre = reinterpret_cast<uint16_t*>(wlds + a.words);
    __shared__ int red[kBlock];
    __shared__ int outside;
    const int ch = blockIdx.x, t = threadIdx.x;
    const int r = ch * kBlock + t;
    const int len = (r < a.nu) ? a.rowlen[r] : 0;
    const int first = (len > 0) ? a.cbase[r] : 0x7fffffff;
    red[t] = first;
    if (t == 0) outside = 0;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (t < s) red[t] = min(red[t], red[t + s]);
        __syncthreads();
    }
    const int lo = red[0];
    __syncthreads();
    if (lo == 0x7fffffff) {                            // no entries (uniform)
        if (t == 0) a.wn[ch] = 0;
        return;
    }
    const int nw = a.words;
    const long nbits = 32L * nw;
    for (int w = t; w < nw; w += kBlock) bits[w] = 0u;
    __syncthreads();
    const v4u* __restrict__ dc = a.dcol + dc_index(r, 0, a.cap);
    int cj = first;
    // a row's columns ascend: the bits of consecutive entries in one word are
    // gathered in a register and set with one atomic when the word changes
    int cw = -1;
    uint32_t cm = 0u;
    for (int k = 0; k < len; k += 32) {
        v4u w[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            w[q] = (k + 8 * q < len) ? dc[8L * (k + 8 * q)] : v4u{0u, 0u, 0u, 0u};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                if (k + 8 * q + u < len) {
                    cj += u16_of(w[q], u);
                    const long o = (long)cj - lo;
                    if (o >= nbits) {
                        outside = 1;
                        continue;
                    }
                    if ((int)(o >> 5) != cw) {
                        if (cw >= 0) atomicOr(&bits[cw], cm);
                        cw = (int)(o >> 5);
                        cm = 0u;
                    }
                    cm |= 1u << (o & 31);
                }
            }
        }
    }
    if (cw >= 0) atomicOr(&bits[cw], cm);
    __syncthreads();
    if (outside) {                                     // uniform
        if (t == 0) { a.wn[ch] = -1; atomicMax(&a.flags[FLAG_WIN_N], kWinMax + 1); }
        return;
    }
    // exclusive prefix of the words' popcounts: thread t owns words [w0, w1)
    const int per = (nw + kBlock - 1) / kBlock;
    const int w0 = min(nw, t * per), w1 = min(nw, w0 + per);
    int sum = 0;
    for (int w = w0; w < w1; ++w) sum += __popc(bits[w]);
    red[t] = sum;
    __syncthreads();
    for (int s = 1; s < kBlock; s <<= 1) {
        const int v = (t >= s) ? red[t - s] : 0;
        __syncthreads();
        red[t] += v;
        __syncthreads();
    }
    const int total = red[kBlock - 1];
    if (total > kWinMax) {                             // uniform
        if (t == 0) { a.wn[ch] = -1; atomicMax(&a.flags[FLAG_WIN_N], kWinMax + 1); }
        return;
    }
    int run = red[t] - sum;
    int32_t* __restrict__ uq = a.wuniq + (long)ch * kWinMax;
    for (int w = w0; w < w1; ++w) {
        pre[w] = (uint16_t)run;
        uint32_t b = bits[w];
        while (b) {
            uq[run++] = lo + 32 * w + (__ffs(b) - 1);
            b &= b - 1u;
        }
    }
    __syncthreads();
    v4u* __restrict__ wc = a.wcol + dc_index(r, 0, a.cap);
    cj = first;
    cw = -1;
    uint32_t cb = 0u;                                  // bits of word cw
    int cp = 0;                                        // and its prefix
    for (int k = 0; k < len; k += 32) {
        v4u w[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            w[q] = (k + 8 * q < len) ? dc[8L * (k + 8 * q)] : v4u{0u, 0u, 0u, 0u};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (k + 8 * q < len) {
                uint32_t o4[4] = {0u, 0u, 0u, 0u};
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    if (k + 8 * q + u < len) {
                        cj += u16_of(w[q], u);
                        const int o = cj - lo;
                        if ((o >> 5) != cw) {
                            cw = o >> 5;
                            cb = bits[cw];
                            cp = pre[cw];
                        }
                        const int rank = cp + __popc(cb & ((1u << (o & 31)) - 1u));
                        o4[u >> 1] |= (uint32_t)rank << (16 * (u & 1));
                    }
                }
                wc[8L * (k + 8 * q)] = v4u{o4[0], o4[1], o4[2], o4[3]};
            }
        }
    }
    if (t == 0) {
        a.wn[ch] = total;
        atomicMax(&a.flags[FLAG_WIN_N], total);
        atomicAdd(&a.flags[FLAG_WIN_SUM], total);
    }
}

template <class VT, class MT>
__global__ __launch_bounds__(kVecThreads) void k_cg_update(CgArgs<VT, MT> a)
{
    __shared__ double lds[kVecThreads];
    if (a.flags[FLAG_CG_DONE]) return;
    const double pq = sum_partials_n<kVecThreads>(PSLOT(a, P_PQ), a.npq, lds);
    const double rz_old = a.scal[S_RZ0 + ((a.it - 1) & 1)];
    if (!(pq > 0.0) || !isfinite(pq)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            a.flags[FLAG_CG_BREAK] = 1;
            a.flags[FLAG_CG_DONE] = 1;
            a.flags[FLAG_CG_ITERS] = a.it - 1;
        }
        return;
    }
    const VT alpha = (VT)(rz_old / pq);
    double rr = 0.0, rz = 0.0;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < a.nu; r += gridDim.x * blockDim.x) {
        const VT pr = a.p[r];
        a.x[r] = a.x[r] + alpha * pr;
        const VT res = a.r[r] - alpha * a.q[r];
        a.r[r] = res;
        const VT z = res / (VT)a.diag[r];
        rr += (double)res * (double)res;
        rz += (double)res * (double)z;
    }
    rr = block_sum_n<kVecThreads>(rr, lds);
    rz = block_sum_n<kVecThreads>(rz, lds);
    if (threadIdx.x == 0) {
        PSLOT(a, P_RR)[blockIdx.x] = rr;
        PSLOT(a, P_RZ)[blockIdx.x] = rz;
    }
}

template <class VT, class MT>
__global__ __launch_bounds__(kVecThreads) void k_cg_p(CgArgs<VT, MT> a)
{
    __shared__ double lds[kVecThreads];
    if (a.flags[FLAG_CG_DONE]) return;
    const double rr = sum_partials_n<kVecThreads>(PSLOT(a, P_RR), kVecBlocks, lds);
    const double rz = sum_partials_n<kVecThreads>(PSLOT(a, P_RZ), kVecBlocks, lds);
    const double rel = sqrt(rr) / a.scal[S_BNORM];
    const bool lead = (blockIdx.x == 0 && threadIdx.x == 0);
    if (lead) {
        a.scal[S_RELRES] = rel;
        a.flags[FLAG_CG_ITERS] = a.it;
    }
    if (rel <= a.tol) {
        if (lead) { a.flags[FLAG_CG_CONV] = 1; a.flags[FLAG_CG_DONE] = 1; }
        return;
    }
    if (a.it >= a.maxit || !isfinite(rel)) {
        if (lead) a.flags[FLAG_CG_DONE] = 1;
        return;
    }
    const VT beta = (VT)(rz / a.scal[S_RZ0 + ((a.it - 1) & 1)]);
    if (lead) a.scal[S_RZ0 + (a.it & 1)] = rz;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < a.nu; r += gridDim.x * blockDim.x)
        a.p[r] = a.r[r] / (VT)a.diag[r] + beta * a.p[r];
}

// ---- flexible CG ---------------------------------------------------------------------------
template <class VT, class MT>
__global__ __launch_bounds__(kVecThreads) void k_fcg_init(CgArgs<VT, MT> a)
{
    __shared__ double lds[kVecThreads];
    double bb = 0.0, nz = 0.0;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < a.nu; r += gridDim.x * blockDim.x) {
        const double b = a.rhs[r];
        a.x[r] = (VT)0;
        a.r[r] = (VT)b;
        bb += b * b;
        nz += (double)(a.rowlen[r] + 1);
    }
    bb = block_sum_n<kVecThreads>(bb, lds);
    nz = block_sum_n<kVecThreads>(nz, lds);
    if (threadIdx.x == 0) {
        PSLOT(a, P_BB)[blockIdx.x] = bb;
        PSLOT(a, P_NNZ)[blockIdx.x] = nz;
    }
}

// z = M r: g_r r_r / A_rr plus the coupling row; one thread per row.
template <class VT, class MT>
__device__ __forceinline__ VT apply_m(const CgArgs<VT, MT>& a, int r)
{
    VT acc = (VT)a.gd[r] * a.r[r];
    const long b = mx_index(r, 0, a.scap);
    const int len = a.slen[r];
    constexpr int U = 8;
    int k = 0;
    for (; k + U <= len; k += U) {
        int32_t cc[U];
        float vv[U];
        VT rr[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            cc[u] = __builtin_nontemporal_load(&a.scol[b + 64L * (k + u)]);
            vv[u] = __builtin_nontemporal_load(&a.sval[b + 64L * (k + u)]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) rr[u] = a.r[cc[u]];
#pragma unroll
        for (int u = 0; u < U; ++u) acc = acc + (VT)vv[u] * rr[u];
    }
    for (; k < len; ++k) acc = acc + (VT)a.sval[b + 64L * k] * a.r[a.scol[b + 64L * k]];
    return acc;
}

// first direction: z = M b, p = z, partial z.b
template <class VT, class MT>
__global__ __launch_bounds__(kBlock) void k_fcg_prec0(CgArgs<VT, MT> a)
{
    __shared__ double lds[kBlock];
    double zr = 0.0;
    const int r = xcd_tid();
    if (r < a.nu) {
        const VT z = apply_m(a, r);
        a.z[r] = z;
        a.p[r] = z;
        zr = (double)z * (double)a.r[r];
    }
    zr = block_sum(zr, lds);
    if (threadIdx.x == 0) PSLOT(a, P_RZ0)[blockIdx.x] = zr;
}

template <class VT, class MT>
__global__ __launch_bounds__(kBlock) void k_fcg_init_fin(CgArgs<VT, MT> a)
{
    __shared__ double lds[kBlock];
    const double bb = sum_partials(PSLOT(a, P_BB), kVecBlocks, lds);
    const double nz = sum_partials(PSLOT(a, P_NNZ), kVecBlocks, lds);
    const double rz = sum_partials(PSLOT(a, P_RZ0), a.npq, lds);
    if (threadIdx.x == 0) {
        const double bn = sqrt(bb);
        a.scal[S_BNORM] = bn;
        a.scal[S_RZ0] = rz;
        a.scal[S_NNZ] = nz;
        a.scal[S_RELRES] = 0.0;
        a.flags[FLAG_CG_ITERS] = 0;
        const bool zero = (bn == 0.0);
        const bool brk = !zero && !(rz > 0.0 && isfinite(rz));
        a.flags[FLAG_CG_BREAK] = brk ? 1 : 0;
        a.flags[FLAG_CG_CONV] = zero ? 1 : 0;
        a.flags[FLAG_CG_DONE] = (zero || brk) ? 1 : 0;
    }
}

template <class VT, class MT>
__global__ __launch_bounds__(kVecThreads) void k_fcg_update(CgArgs<VT, MT> a)
{
    __shared__ double lds[kVecThreads];
    if (a.flags[FLAG_CG_DONE]) return;
    const double pq = sum_partials_n<kVecThreads>(PSLOT(a, P_PQ), a.npq, lds);
    const double rz_old = a.scal[S_RZ0 + ((a.it - 1) & 1)];
    if (!(pq > 0.0) || !isfinite(pq)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            a.flags[FLAG_CG_BREAK] = 1;
            a.flags[FLAG_CG_DONE] = 1;
            a.flags[FLAG_CG_ITERS] = a.it - 1;
        }
        return;
    }
    const VT alpha = (VT)(rz_old / pq);
    double rr = 0.0;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < a.nu; r += gridDim.x * blockDim.x) {
        a.x[r] = a.x[r] + alpha * a.p[r];
        const VT res = a.r[r] - alpha * a.q[r];
        a.r[r] = res;
        rr += (double)res * (double)res;
    }
    rr = block_sum_n<kVecThreads>(rr, lds);
    if (threadIdx.x == 0) PSLOT(a, P_RR)[blockIdx.x] = rr;
}

template <class VT, class MT>
__global__ __launch_bounds__(kBlock) void k_fcg_prec(CgArgs<VT, MT> a)
{
    __shared__ double lds[kBlock];
    if (a.flags[FLAG_CG_DONE]) return;
    const double rr = sum_partials(PSLOT(a, P_RR), kVecBlocks, lds);
    const double rel = sqrt(rr) / a.scal[S_BNORM];
    const bool lead = (blockIdx.x == 0 && threadIdx.x == 0);
    if (lead) {
        a.scal[S_RELRES] = rel;
        a.flags[FLAG_CG_ITERS] = a.it;
    }
    if (rel <= a.tol) {
        if (lead) { a.flags[FLAG_CG_CONV] = 1; a.flags[FLAG_CG_DONE] = 1; }
        return;
    }
    if (a.it >= a.maxit || !isfinite(rel)) {
        if (lead) a.flags[FLAG_CG_DONE] = 1;
        return;
    }
    double zr = 0.0, zq = 0.0;
    const int r = xcd_tid();
    if (r < a.nu) {
        const VT z = apply_m(a, r);
        a.z[r] = z;
        zr = (double)z * (double)a.r[r];
        zq = (double)z * (double)a.q[r];
    }
    zr = block_sum(zr, lds);
    zq = block_sum(zq, lds);
    if (threadIdx.x == 0) {
        PSLOT(a, P_RZ)[blockIdx.x] = zr;
        PSLOT(a, P_ZQ)[blockIdx.x] = zq;
    }
}

template <class VT, class MT>
__global__ __launch_bounds__(kVecThreads) void k_fcg_p(CgArgs<VT, MT> a)
{
    __shared__ double lds[kVecThreads];
    if (a.flags[FLAG_CG_DONE]) return;
    const double pq = sum_partials_n<kVecThreads>(PSLOT(a, P_PQ), a.npq, lds);
    const double zr = sum_partials_n<kVecThreads>(PSLOT(a, P_RZ), a.npq, lds);
    const double zq = sum_partials_n<kVecThreads>(PSLOT(a, P_ZQ), a.npq, lds);
    const double rz_old = a.scal[S_RZ0 + ((a.it - 1) & 1)];
    const bool lead = (blockIdx.x == 0 && threadIdx.x == 0);
    if (!(zr > 0.0) || !isfinite(zr)) {
        // M r is not a descent direction: the preconditioner is not positive here
        if (lead) { a.flags[FLAG_CG_BREAK] = 1; a.flags[FLAG_CG_DONE] = 1; }
        return;
    }
    const double alpha = rz_old / pq;
    const VT beta = (VT)(-alpha * zq / rz_old);
    if (lead) a.scal[S_RZ0 + (a.it & 1)] = zr;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < a.nu; r += gridDim.x * blockDim.x)
        a.p[r] = a.z[r] + beta * a.p[r];
}

// ---- outer FP64 refinement kernels ---------------------------------------------------------
template <class MT>
struct IrArgs {
    int nu, cap, npq, win_n;
    const int32_t* __restrict__ rowlen;
    const int32_t* __restrict__ acol;
    const v4u* __restrict__ wcol;     // ranks in the product's LDS windows (see k_cg_spmv_win)
    const int32_t* __restrict__ wuniq;
    const int32_t* __restrict__ wn;
    const MT* __restrict__ aval;
    const double* __restrict__ diag;
    const double* __restrict__ b;
    const double* __restrict__ x;
    double* res;
    double* part;
    double* scal;
};

// res = b - A x in FP64 arithmetic with the matrix values as stored (FP64, or
// FP32 after a mixed-precision assembly); partials ||res||^2 and ||b||^2 per
// block.
template <class MT>
__global__ __launch_bounds__(kVecThreads) void k_ir_residual(IrArgs<MT> a)
{
    __shared__ double lds[kVecThreads];
    double rr = 0.0, bb = 0.0;
    const int r = xcd_tid();
    if (r < a.nu) {
        double acc = a.diag[r] * a.x[r];
        const int len = a.rowlen[r];
        const int32_t* __restrict__ col = a.acol + mx_index(r, 0, a.cap);
        const MT* __restrict__ val = a.aval + mx_index(r, 0, a.cap);
        constexpr int U = 8;
        int k = 0;
        for (; k + U <= len; k += U) {
            int32_t cc[U];
            double vv[U], xx[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                cc[u] = __builtin_nontemporal_load(&col[64L * (k + u)]);
                vv[u] = (double)__builtin_nontemporal_load(&val[64L * (k + u)]);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) xx[u] = a.x[cc[u]];
#pragma unroll
            for (int u = 0; u < U; ++u) acc = acc + vv[u] * xx[u];
        }
        for (; k < len; ++k) acc = acc + (double)val[64L * k] * a.x[col[64L * k]];
        const double b = a.b[r];
        const double rv = b - acc;
        a.res[r] = rv;
        rr = rv * rv;
        bb = b * b;
    }
    rr = block_sum(rr, lds);
    bb = block_sum(bb, lds);
    if (threadIdx.x == 0) {
        PSLOT(a, P_RES)[blockIdx.x] = rr;
        PSLOT(a, P_RHS)[blockIdx.x] = bb;
    }
}

// The same residual with x gathered from the product's LDS windows.
template <class MT>
__global__ __launch_bounds__(kVecThreads) void k_ir_residual_win(IrArgs<MT> a)
{
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    double* win = reinterpret_cast<double*>(dyn_lds);
    __shared__ double lds[kVecThreads];
    const int ch = xcd_block();
    const int n = a.wn[ch];
    const int32_t* __restrict__ uq = a.wuniq + (long)ch * kWinMax;
    for (int t = threadIdx.x; t < n; t += kBlock) win[t] = a.x[uq[t]];
    __syncthreads();
    double rr = 0.0, bb = 0.0;
    const int r = ch * kBlock + threadIdx.x;
    if (r < a.nu) {
        double acc = a.diag[r] * a.x[r];
        const int len = a.rowlen[r];
        const v4u* __restrict__ wc = a.wcol + dc_index(r, 0, a.cap);
        const MT* __restrict__ val = a.aval + mx_index(r, 0, a.cap);
        int k = 0;
        for (; k + 8 <= len; k += 8) {
            const v4u w = __builtin_nontemporal_load(&wc[8L * k]);
            double vv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) vv[u] = (double)__builtin_nontemporal_load(&val[64L * (k + u)]);
#pragma unroll
            for (int u = 0; u < 8; ++u) acc = acc + vv[u] * win[u16_of(w, u)];
        }
        if (k < len) {
            const v4u w = __builtin_nontemporal_load(&wc[8L * k]);
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (k + u < len) acc = acc + (double)val[64L * (k + u)] * win[u16_of(w, u)];
        }
        const double b = a.b[r];
        const double rv = b - acc;
        a.res[r] = rv;
        rr = rv * rv;
        bb = b * b;
    }
    rr = block_sum(rr, lds);
    bb = block_sum(bb, lds);
    if (threadIdx.x == 0) {
        PSLOT(a, P_RES)[blockIdx.x] = rr;
        PSLOT(a, P_RHS)[blockIdx.x] = bb;
    }
}

template <class MT>
__global__ __launch_bounds__(kVecThreads) void k_ir_residual_fin(IrArgs<MT> a)
{
    __shared__ double lds[kVecThreads];
    const double rr = sum_partials(PSLOT(a, P_RES), a.npq, lds);
    const double bb = sum_partials(PSLOT(a, P_RHS), a.npq, lds);
    if (threadIdx.x == 0) {
        a.scal[S_RESN] = sqrt(rr);
        a.scal[S_RHSN] = sqrt(bb);
    }
}

template <class VT>
__global__ void k_ir_correct(int nu, double* __restrict__ x, const VT* __restrict__ d)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nu) x[r] = x[r] + (double)d[r];
}

// Start vector of a warm-started solve: each unknown's pressure of the
// previous step (the P field), by its particle index.
__global__ void k_gather_x0(int nu, const int32_t* __restrict__ urow, const double* __restrict__ p,
                            double* __restrict__ x)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nu) x[r] = p[urow[r]];
}

__global__ void k_scatter_x(int nu, const int32_t* __restrict__ urow, const double* __restrict__ x,
                            double* __restrict__ pnew)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nu) pnew[urow[r]] = x[r];
}

// ---- host ------------------------------------------------------------------------------------
// The product variant the autotuner chose (variant 0 unless tuned). Variant 0
// gathers p from the blocks' LDS windows when they are built, variant 4 reads
// the columns' 16-bit steps when the assembly stored them all; either falls
// back to the next form (window -> steps -> 32-bit columns). Variants 1-3
// read the 32-bit columns, 4, 16 and 8 entries per batch.
template <class VT, class MT>
static void launch_spmv(vox_ctx* c, const CgArgs<VT, MT>& a, int variant)
{
    if (variant == 0 && a.wcol) {
        c->st.product_col_bytes = 2;
        c->st.product_window_cols = c->win_sum;
        k_cg_spmv_win<VT, MT><<<a.npq, kBlock, sizeof(VT) * (size_t)a.win_n, c->stream>>>(a);
        return;
    }
    const bool steps = (variant == 0 || variant == 4) && a.dcol != nullptr;
    c->st.product_col_bytes = steps ? 2 : 4;
    c->st.product_window_cols = 0;
    if (steps) {
        k_cg_spmv<VT, MT, 8, true><<<a.npq, kBlock, 0, c->stream>>>(a);
        return;
    }
    switch (variant) {
    case 1: k_cg_spmv<VT, MT, 4><<<a.npq, kBlock, 0, c->stream>>>(a); break;
    case 2: k_cg_spmv<VT, MT, 16><<<a.npq, kBlock, 0, c->stream>>>(a); break;
    default: k_cg_spmv<VT, MT, 8><<<a.npq, kBlock, 0, c->stream>>>(a); break;
    }
}

// Runs CG on A x = rhs to ||r|| <= tol ||rhs||. Returns iterations; *conv, *brk.
template <class VT, class MT>
static int run_cg(vox_ctx* c, CgArgs<VT, MT>& a, int* iters_out)
{
    {
        ProfScope ps(c, KC_CG_INIT);
        k_cg_init<VT, MT><<<kVecBlocks, kVecThreads, 0, c->stream>>>(a);
        k_cg_init_fin<VT, MT><<<1, kBlock, 0, c->stream>>>(a);
        VOX_HIP(hipGetLastError());
    }
    // Check schedule: the first batch runs the last solve's count (consecutive
    // solves need similar counts), then the flag is read after every cg_check
    // iterations (default 1). The convergence test is on the device and kernels
    // queued past it return at entry, so the host reads the flags about once per
    // solve; the setup's partial sums (nnz) are read with them.
    const int check = c->p.cg_check > 0 ? c->p.cg_check : 1;
    int& expect = std::is_same<VT, double>::value ? c->cg_expect : c->cg_expect_inner;
    const int first = expect;
    int it = 0;
    bool read = false;
    c->hflags[FLAG_CG_DONE] = 0;        // host copy; the device decides
    while (!c->hflags[FLAG_CG_DONE] && it < a.maxit) {
        const int batch = (it == 0 && first > check) ? first : check;
        for (int k = 0; k < batch && it < a.maxit; ++k) {
            a.it = ++it;
            {
                ProfScope ps(c, KC_SPMV);
                launch_spmv<VT, MT>(c, a, c->tune_spmv);
            }
            {
                ProfScope ps(c, KC_CG_UPDATE);
                k_cg_update<VT, MT><<<kVecBlocks, kVecThreads, 0, c->stream>>>(a);
            }
            {
                ProfScope ps(c, KC_CG_P);
                k_cg_p<VT, MT><<<kVecBlocks, kVecThreads, 0, c->stream>>>(a);
            }
        }
        VOX_HIP(hipGetLastError());
        VOX_TRY(read_flags(c));
        read = true;
    }
    if (!read) VOX_TRY(read_flags(c));
    c->nnz_total = (long)c->hscal[S_NNZ];
    expect = c->hflags[FLAG_CG_ITERS];
    *iters_out = c->hflags[FLAG_CG_ITERS];
    return VOX_OK;
}

// Flexible CG with the learned coupling operator; same check schedule as run_cg.
template <class VT, class MT>
static int run_fcg(vox_ctx* c, CgArgs<VT, MT>& a, int* iters_out)
{
    {
        ProfScope ps(c, KC_CG_INIT);
        k_fcg_init<VT, MT><<<kVecBlocks, kVecThreads, 0, c->stream>>>(a);
    }
    {
        ProfScope ps(c, KC_LPC_APPLY);
        k_fcg_prec0<VT, MT><<<a.npq, kBlock, 0, c->stream>>>(a);
    }
    {
        ProfScope ps(c, KC_CG_INIT);
        k_fcg_init_fin<VT, MT><<<1, kBlock, 0, c->stream>>>(a);
        VOX_HIP(hipGetLastError());
    }
    const int check = c->p.cg_check > 0 ? c->p.cg_check : 1;
    int& expect = std::is_same<VT, double>::value ? c->fcg_expect : c->fcg_expect_inner;
    const int first = expect;
    int it = 0;
    bool read = false;
    c->hflags[FLAG_CG_DONE] = 0;        // host copy; the device decides
    while (!c->hflags[FLAG_CG_DONE] && it < a.maxit) {
        const int batch = (it == 0 && first > check) ? first : check;
        for (int k = 0; k < batch && it < a.maxit; ++k) {
            a.it = ++it;
            {
                ProfScope ps(c, KC_SPMV);
                launch_spmv<VT, MT>(c, a, c->tune_spmv);
            }
            {
                ProfScope ps(c, KC_CG_UPDATE);
                k_fcg_update<VT, MT><<<kVecBlocks, kVecThreads, 0, c->stream>>>(a);
            }
            {
                ProfScope ps(c, KC_LPC_APPLY);
                k_fcg_prec<VT, MT><<<a.npq, kBlock, 0, c->stream>>>(a);
            }
            {
                ProfScope ps(c, KC_CG_P);
                k_fcg_p<VT, MT><<<kVecBlocks, kVecThreads, 0, c->stream>>>(a);
            }
        }
        VOX_HIP(hipGetLastError());
        VOX_TRY(read_flags(c));
        read = true;
    }
    if (!read) VOX_TRY(read_flags(c));
    c->nnz_total = (long)c->hscal[S_NNZ];
    expect = c->hflags[FLAG_CG_ITERS];
    *iters_out = c->hflags[FLAG_CG_ITERS];
    return VOX_OK;
}

// Builds the LDS windows of the current assembly once (k_win_build); win_ok
// says whether every chunk fit.
static int ensure_window(vox_ctx* c)
{
    if (c->win_asm == c->asm_count) return VOX_OK;
    c->win_asm = c->asm_count;
    c->win_ok = false;
    if (!c->dcol_ok || c->nu <= 0) return VOX_OK;
    const long nch = grid_for(c->nu);
    if (c->wcap_alloc != c->a_cap) {
        if (c->wcol) (void)hipFree(c->wcol);
        c->wcol = nullptr;
        c->wcap_alloc = 0;
        VOX_HIP(hipMalloc(&c->wcol, sizeof(uint16_t) * (size_t)sl_size(c->tp, c->a_cap)));
        c->wcap_alloc = c->a_cap;
    }
    if (c->wchunks_alloc < nch) {
        if (c->wuniq) (void)hipFree(c->wuniq);
        if (c->wn) (void)hipFree(c->wn);
        c->wuniq = nullptr;
        c->wn = nullptr;
        c->wchunks_alloc = 0;
        const long m = grid_for(c->tp);
        VOX_HIP(hipMalloc(&c->wuniq, sizeof(int32_t) * (size_t)m * kWinMax));
        VOX_HIP(hipMalloc(&c->wn, sizeof(int32_t) * (size_t)m));
        c->wchunks_alloc = m;
    }
    // a chunk's columns lie in [r0 - below, r0 + 255 + above]
    const long span = (long)c->col_below + c->col_above + kBlock;
    const int words = (int)((span + 31) / 32);
    if (words > kWinSpanWords) return VOX_OK;
    VOX_TRY(fill_i32(c, c->dflags + FLAG_WIN_N, 0, 2));          // and FLAG_WIN_SUM
    WinArgs w;
    w.nu = c->nu; w.cap = c->a_cap; w.words = words; w.rowlen = c->rowlen;
    w.dcol = reinterpret_cast<const v4u*>(c->dcol); w.cbase = c->cbase;
    w.wcol = reinterpret_cast<v4u*>(c->wcol); w.wuniq = c->wuniq; w.wn = c->wn; w.flags = c->dflags;
    {
        ProfScope ps(c, KC_ASSEMBLE);
        const size_t lds = sizeof(uint32_t) * (size_t)words + sizeof(uint16_t) * (size_t)(words + 1);
        k_win_build<<<(unsigned)nch, kBlock, lds, c->stream>>>(w);
        VOX_HIP(hipGetLastError());
    }
    VOX_TRY(read_flags(c));
    c->win_n = c->hflags[FLAG_WIN_N];
    c->win_sum = c->hflags[FLAG_WIN_SUM];
    c->win_ok = c->win_n <= kWinMax;
    return VOX_OK;
}

// The windows when the product variant in use (or being timed) reads them,
// and the 32-bit columns when it reads those.
static int prepare_product(vox_ctx* c, bool all_variants = false)
{
    const int v = c->tune_spmv;
    if (all_variants || v == 1 || v == 2 || v == 3 || !c->dcol_ok) VOX_TRY(ensure_acol(c));
    if (all_variants || v == 0) return ensure_window(c);
    return VOX_OK;
}

template <class VT, class MT>
static void fill_args(vox_ctx* c, CgArgs<VT, MT>& a, const MT* aval, VT* x, VT* r, VT* p, VT* q,
                      const double* rhs, double tol, int maxit)
{
    const bool win = c->win_ok && c->win_asm == c->asm_count;
    a.wcol = win ? reinterpret_cast<const v4u*>(c->wcol) : nullptr;
    a.wuniq = c->wuniq; a.wn = c->wn; a.win_n = c->win_n;
    a.nu = c->nu; a.cap = c->a_cap; a.tol = tol; a.maxit = maxit; a.it = 0;
    a.npq = grid_for(c->nu > 0 ? c->nu : 1);
    a.rowlen = c->rowlen; a.acol = c->acol; a.aval = aval; a.diag = c->adiag; a.rhs = rhs;
    a.dcol = c->dcol_ok ? reinterpret_cast<const v4u*>(c->dcol) : nullptr; a.cbase = c->cbase;
    a.x = x; a.r = r; a.p = p; a.q = q;
    a.part = c->partials; a.scal = c->dscal; a.flags = c->dflags;
    a.z = nullptr; a.slen = c->lpc_slen; a.scol = c->lpc_scol; a.sval = c->lpc_sval; a.gd = c->lpc_gd;
    a.scap = c->lpc_scap;
}

static int finish(vox_ctx* c, vox_solve_info* info, int iters, int conv, double relres,
                  double bnorm, int brk)
{
    const int nu = c->nu;
    if (nu > 0) {
        k_scatter_x<<<grid_for(nu), kBlock, 0, c->stream>>>(nu, c->urow, c->cg_x, c->f[VOX_F_PNEW]);
        VOX_HIP(hipGetLastError());
    }
    c->st.last_iters = iters;
    c->st.last_relres = relres;
    if (info) {
        info->iterations = iters;
        info->converged = conv;
        info->unknowns = nu;
        info->nnz = (int32_t)c->nnz_total;
        info->relres = relres;
        info->bnorm = bnorm;
    }
    if (brk) return fail(VOX_ERR_SOLVER, "CG breakdown (p.Ap <= 0 or not finite)");
    if (!conv)
        return fail(VOX_ERR_SOLVER, "PPE solve did not reach the relative residual " +
                                        std::to_string(c->p.cg_tol) + " in " +
                                        std::to_string(iters) + " iterations");
    return VOX_OK;
}

static int solve_fp64(vox_ctx* c, vox_solve_info* info, bool fcg, int maxit)
{
    VOX_TRY(prepare_product(c));
    CgArgs<double, double> a;
    fill_args(c, a, (const double*)c->aval, c->cg_x, c->cg_r, c->cg_p, c->cg_q, c->rhs,
              c->p.cg_tol, maxit);
    a.z = c->cg_z;
    if (a.npq > kMaxPartials) return fail(VOX_ERR_ARG, "PPE larger than the solver's row limit");
    int iters = 0;
    if (fcg) VOX_TRY(run_fcg(c, a, &iters));
    else VOX_TRY(run_cg(c, a, &iters));
    return finish(c, info, iters, c->hflags[FLAG_CG_CONV], c->hscal[S_RELRES], c->hscal[S_BNORM],
                  c->hflags[FLAG_CG_BREAK]);
}

// res = b - A x (FP64) over the matrix as stored, then the norms into the
// scalar slots; FP64 values when the assembly kept them, else its FP32 values.
template <class MT>
static int launch_ir(vox_ctx* c, IrArgs<MT>& ir, const MT* aval, int npq)
{
    ir.nu = c->nu; ir.cap = c->a_cap; ir.npq = npq;
    ir.rowlen = c->rowlen; ir.acol = c->acol; ir.aval = aval; ir.diag = c->adiag;
    ir.b = c->rhs; ir.x = c->cg_x; ir.res = c->cg_r; ir.part = c->partials; ir.scal = c->dscal;
    const bool win = c->win_ok && c->win_asm == c->asm_count;
    if (!win) VOX_TRY(ensure_acol(c));
    ir.wcol = win ? reinterpret_cast<const v4u*>(c->wcol) : nullptr;
    ir.wuniq = c->wuniq; ir.wn = c->wn; ir.win_n = c->win_n;
    if (win)
        k_ir_residual_win<MT><<<npq, kBlock, sizeof(double) * (size_t)ir.win_n, c->stream>>>(ir);
    else
        k_ir_residual<MT><<<npq, kBlock, 0, c->stream>>>(ir);
    k_ir_residual_fin<MT><<<1, kBlock, 0, c->stream>>>(ir);
    return VOX_OK;
}

// (x gathered from the product's LDS windows when they are built)
static int ir_residual(vox_ctx* c, int npq)
{
    ProfScope ps(c, KC_IR);
    if (c->aval64_ok) {
        IrArgs<double> ir;
        VOX_TRY(launch_ir(c, ir, (const double*)c->aval, npq));
    } else {
        IrArgs<float> ir;
        VOX_TRY(launch_ir(c, ir, (const float*)c->aval32, npq));
    }
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

// Refinement driver: x = x0 (zero, or the previous pressure when warm);
// repeat r = b - A x in FP64 on the stored matrix, stop when ||r|| <= tol ||b||,
// else solve A d = r with CG<VT, MT> to a relative tolerance and x += d.
// Mixed precision: VT = MT = float, each correction to max(eta, the remaining
// factor). FP64 warm start: VT = MT = double, one correction normally reaches
// the target.
template <class VT, class MT>
static int solve_refined(vox_ctx* c, vox_solve_info* info, const MT* aval, VT* x, VT* r, VT* p,
                         VT* q, VT* z, double eta, bool warm, bool fcg, int maxit)
{
    const int nu = c->nu;
    const int npq = grid_for(nu > 0 ? nu : 1);
    if (npq > kMaxPartials) return fail(VOX_ERR_ARG, "PPE larger than the solver's row limit");
    VOX_TRY(prepare_product(c));
    if (warm && nu > 0) {
        k_gather_x0<<<grid_for(nu), kBlock, 0, c->stream>>>(nu, c->urow, c->f[VOX_F_P], c->cg_x);
        VOX_HIP(hipGetLastError());
    } else {
        VOX_TRY(fill_f64(c, c->cg_x, 0.0, nu));
    }
    const double tol = c->p.cg_tol;
    int total = 0, conv = 0, brk = 0, outer = 0;
    double bnorm = 0.0, rel = 1.0;
    for (;; ++outer) {
        // From a zero start the first residual is b itself: no product is
        // needed, and the first inner solve takes b directly.
        const bool from_b = (outer == 0 && !warm);
        if (!from_b) {
            VOX_TRY(ir_residual(c, npq));
            VOX_TRY(read_flags(c));
            const double rn = c->hscal[S_RESN];
            bnorm = c->hscal[S_RHSN];
            rel = (bnorm > 0) ? rn / bnorm : 0.0;
            if (rel <= tol) { conv = 1; break; }
            if (outer >= c->p.ir_maxouter || total >= maxit || !std::isfinite(rel)) break;
        }
        // relative target of this correction: no tighter than the remaining
        // factor to the goal (half of it, for the rounding of the update)
        const double need = 0.5 * tol / rel;
        CgArgs<VT, MT> a;
        fill_args(c, a, aval, x, r, p, q, from_b ? c->rhs : c->cg_r, eta > need ? eta : need,
                  maxit - total);
        a.z = z;
        int it = 0;
        if (fcg) VOX_TRY(run_fcg(c, a, &it));
        else VOX_TRY(run_cg(c, a, &it));
        total += it;
        if (c->hflags[FLAG_CG_BREAK] && it == 0) { brk = 1; break; }
        k_ir_correct<VT><<<grid_for(nu), kBlock, 0, c->stream>>>(nu, c->cg_x, x);
        VOX_HIP(hipGetLastError());
    }
    c->ir_outer_last = outer;
    return finish(c, info, total, conv, rel, bnorm, brk);
}

static int solve_mixed(vox_ctx* c, vox_solve_info* info, bool fcg, int maxit)
{
    const double eta = c->p.ir_eta > 0 ? c->p.ir_eta : 1e-4;
    if (!c->aval32_ok) VOX_TRY(ppe_make_aval32(c));
    return solve_refined<float, float>(c, info, (const float*)c->aval32, c->cg_x32, c->cg_r32,
                                       c->cg_p32, c->cg_q32, c->cg_z32, eta, c->p.warm_start != 0,
                                       fcg, maxit);
}

static int solve_plain(vox_ctx* c, vox_solve_info* info, bool fcg, int maxit)
{
    VOX_TRY(fill_f64(c, c->f[VOX_F_PNEW], 0.0, c->tp));
    if (c->p.precision == VOX_PREC_MIXED) return solve_mixed(c, info, fcg, maxit);
    // an FP64 solve of a system assembled in mixed precision solves its FP32
    // values, which FP64 holds exactly
    if (!c->aval64_ok) VOX_TRY(ppe_make_aval64(c));
    if (c->p.warm_start)
        return solve_refined<double, double>(c, info, (const double*)c->aval, c->cg_d, c->cg_e,
                                             c->cg_p, c->cg_q, c->cg_z, 0.0, true, fcg, maxit);
    return solve_fp64(c, info, fcg, maxit);
}

// Learned solve: flexible CG with the network's preconditioner, held to twice
// the iterations of the last Jacobi-CG solve (the baseline); a first solve
// without a baseline, a solve over the cap and a breakdown run Jacobi-CG,
// which sets the bar every result must meet (same tolerance on the same system).
static int solve_learned(vox_ctx* c, vox_solve_info* info)
{
    if (c->lpc_baseline <= 0) {
        const int rc = solve_plain(c, info, false, c->p.cg_maxit);
        if (rc == VOX_OK) c->lpc_baseline = c->st.last_iters > 0 ? c->st.last_iters : 1;
        c->st.lpc_baseline = c->lpc_baseline;
        return rc;
    }
    VOX_TRY(lpc_prepare(c));
    const long want = ((long)c->lpc_baseline * c->p.lpc_cap_pct + 99) / 100;
    const int cap = (int)std::max(1L, std::min((long)c->p.cg_maxit, want));
    int rc = solve_plain(c, info, true, cap);
    if (rc != VOX_ERR_SOLVER) return rc;
    c->st.lpc_fallbacks += 1;
    rc = solve_plain(c, info, false, c->p.cg_maxit);
    if (rc == VOX_OK) c->lpc_baseline = c->st.last_iters > 0 ? c->st.last_iters : 1;
    c->st.lpc_baseline = c->lpc_baseline;
    return rc;
}

// Times each product variant on the assembled system (the context's solver
// vectors hold p = D^-1 b afterwards, as after a solve's setup): ms[v] is the
// mean of `reps` launches. Used by the autotuner; needs nu > 0.
template <class VT, class MT>
static int time_spmv_t(vox_ctx* c, CgArgs<VT, MT>& a, int reps, float* ms)
{
    k_cg_init<VT, MT><<<kVecBlocks, kVecThreads, 0, c->stream>>>(a);
    k_cg_init_fin<VT, MT><<<1, kBlock, 0, c->stream>>>(a);
    VOX_HIP(hipGetLastError());
    hipEvent_t e0, e1;
    VOX_HIP(hipEventCreate(&e0));
    VOX_HIP(hipEventCreate(&e1));
    int rc = VOX_OK;
    for (int v = 0; v < kSpmvVariants && rc == VOX_OK; ++v) {
        a.it = 1;
        launch_spmv<VT, MT>(c, a, v);                          // warm
        (void)hipEventRecord(e0, c->stream);
        for (int k = 0; k < reps; ++k) launch_spmv<VT, MT>(c, a, v);
        (void)hipEventRecord(e1, c->stream);
        if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
            rc = fail(VOX_ERR_HIP, "product timing failed");
        float t = 0.f;
        (void)hipEventElapsedTime(&t, e0, e1);
        ms[v] = t / reps;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int time_spmv_variants(vox_ctx* c, int reps, float* ms)
{
    if (c->nu <= 0) return fail(VOX_ERR_STATE, "no assembled PPE to time the product on");
    if (grid_for(c->nu) > kMaxPartials) return fail(VOX_ERR_ARG, "PPE larger than the solver's row limit");
    VOX_TRY(prepare_product(c, true));
    if (c->p.precision == VOX_PREC_MIXED) {
        if (!c->aval32_ok) VOX_TRY(ppe_make_aval32(c));
        CgArgs<float, float> a;
        fill_args(c, a, (const float*)c->aval32, c->cg_x32, c->cg_r32, c->cg_p32, c->cg_q32,
                  c->rhs, c->p.cg_tol, c->p.cg_maxit);
        return time_spmv_t(c, a, reps, ms);
    }
    if (!c->aval64_ok) VOX_TRY(ppe_make_aval64(c));
    CgArgs<double, double> a;
    fill_args(c, a, (const double*)c->aval, c->cg_x, c->cg_r, c->cg_p, c->cg_q, c->rhs,
              c->p.cg_tol, c->p.cg_maxit);
    return time_spmv_t(c, a, reps, ms);
}

int ppe_solve(vox_ctx* c, vox_solve_info* info)
{
    if (c->p.solver == VOX_SOLVER_LEARNED_FCG) return solve_learned(c, info);
    return solve_plain(c, info, false, c->p.cg_maxit);
}

}  // namespace vox
