// vox_cg.hip — FP64 Jacobi-preconditioned CG for the assembled PPE on gfx950.
//
// Replaces INCDECOM + CGM (functions.cu:1139-1200, 506-716) and the GPU loop of
// main.cu:1159-1261 (which synchronised with the host and used FP32 atomics for
// every dot product). One iteration is three kernels on a fixed grid:
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
// The matrix is sliced ELL (SELL-64); the product runs one thread per row.
#include "vox_internal.hpp"

namespace vox {

constexpr int kCgBlocks = 2048;   // 8192 waves: 8 per SIMD on 256 CUs

// partial slots (kMaxPartials doubles each)
enum { P_BB = 0, P_RZ0, P_NNZ, P_PQ, P_RR, P_RZ, P_COUNT };
#define PSLOT(a, s) ((a).part + (long)(s) * kMaxPartials)
// scalar slots
enum { S_BNORM = 0, S_RZ0, S_RZ1, S_RELRES, S_NNZ, S_COUNT_ };
static_assert(S_COUNT_ <= SC_COUNT, "scalar slots");

struct CgArgs {
    int nu, cap, maxit, it, npq;   // npq: blocks of the product kernel
    double tol;
    const int32_t* __restrict__ rowlen;
    const int32_t* __restrict__ acol;
    const double* __restrict__ aval;
    const double* __restrict__ diag;
    const double* __restrict__ rhs;
    double* x;
    double* r;
    double* p;
    double* q;
    double* part;
    double* scal;
    int32_t* flags;
};

// Sum of the first n partials of one slot, identical in every block.
__device__ __forceinline__ double sum_partials(const double* part, int n, double* lds)
{
    double s = 0.0;
    for (int k = threadIdx.x; k < n; k += kBlock) s = s + part[k];
    return block_sum(s, lds);
}

__global__ __launch_bounds__(kBlock) void k_cg_init(CgArgs a)
{
    __shared__ double lds[kBlock];
    double bb = 0.0, rz = 0.0, nz = 0.0;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < a.nu; r += gridDim.x * blockDim.x) {
        double b = a.rhs[r];
        double z = b / a.diag[r];
        a.x[r] = 0.0;
        a.r[r] = b;
        a.p[r] = z;
        bb += b * b;
        rz += b * z;
        nz += (double)(a.rowlen[r] + 1);
    }
    bb = block_sum(bb, lds);
    rz = block_sum(rz, lds);
    nz = block_sum(nz, lds);
    if (threadIdx.x == 0) {
        PSLOT(a, P_BB)[blockIdx.x] = bb;
        PSLOT(a, P_RZ0)[blockIdx.x] = rz;
        PSLOT(a, P_NNZ)[blockIdx.x] = nz;
    }
}

__global__ __launch_bounds__(kBlock) void k_cg_init_fin(CgArgs a)
{
    __shared__ double lds[kBlock];
    double bb = sum_partials(PSLOT(a, P_BB), kCgBlocks, lds);
    double rz = sum_partials(PSLOT(a, P_RZ0), kCgBlocks, lds);
    double nz = sum_partials(PSLOT(a, P_NNZ), kCgBlocks, lds);
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

// One thread per row over the sliced layout: a wave's 64 rows read slot k as
// one contiguous line set; 16 entries' loads are in flight per lane and the
// row is accumulated in order. The matrix streams from HBM once per product
// (non-temporal loads keep it from evicting the gathered vector p from L2).
__global__ __launch_bounds__(kBlock) void k_cg_spmv(CgArgs a)
{
    __shared__ double lds[kBlock];
    if (a.flags[FLAG_CG_DONE]) return;
    double pq = 0.0;
    const int r = xcd_tid();
    if (r < a.nu) {
        const double pr = a.p[r];
        double acc = a.diag[r] * pr;
        const int len = a.rowlen[r];
        const int32_t* __restrict__ col = a.acol + sl_index(r, 0, a.cap);
        const double* __restrict__ val = a.aval + sl_index(r, 0, a.cap);
        constexpr int U = 16;
        int k = 0;
        for (; k + U <= len; k += U) {
            int32_t cc[U];
            double vv[U], pp[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                cc[u] = __builtin_nontemporal_load(&col[64L * (k + u)]);
                vv[u] = __builtin_nontemporal_load(&val[64L * (k + u)]);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) pp[u] = a.p[cc[u]];
#pragma unroll
            for (int u = 0; u < U; ++u) acc = acc + vv[u] * pp[u];
        }
        for (; k < len; ++k) acc = acc + val[64L * k] * a.p[col[64L * k]];
        a.q[r] = acc;
        pq = pr * acc;
    }
    pq = block_sum(pq, lds);
    if (threadIdx.x == 0) PSLOT(a, P_PQ)[blockIdx.x] = pq;
}

__global__ __launch_bounds__(kBlock) void k_cg_update(CgArgs a)
{
    __shared__ double lds[kBlock];
    if (a.flags[FLAG_CG_DONE]) return;
    const double pq = sum_partials(PSLOT(a, P_PQ), a.npq, lds);
    const double rz_old = a.scal[S_RZ0 + ((a.it - 1) & 1)];
    if (!(pq > 0.0) || !isfinite(pq)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            a.flags[FLAG_CG_BREAK] = 1;
            a.flags[FLAG_CG_DONE] = 1;
            a.flags[FLAG_CG_ITERS] = a.it - 1;
        }
        return;
    }
    const double alpha = rz_old / pq;
    double rr = 0.0, rz = 0.0;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < a.nu; r += gridDim.x * blockDim.x) {
        const double pr = a.p[r];
        a.x[r] = a.x[r] + alpha * pr;
        const double res = a.r[r] - alpha * a.q[r];
        a.r[r] = res;
        const double z = res / a.diag[r];
        rr += res * res;
        rz += res * z;
    }
    rr = block_sum(rr, lds);
    rz = block_sum(rz, lds);
    if (threadIdx.x == 0) {
        PSLOT(a, P_RR)[blockIdx.x] = rr;
        PSLOT(a, P_RZ)[blockIdx.x] = rz;
    }
}

__global__ __launch_bounds__(kBlock) void k_cg_p(CgArgs a)
{
    __shared__ double lds[kBlock];
    if (a.flags[FLAG_CG_DONE]) return;
    const double rr = sum_partials(PSLOT(a, P_RR), kCgBlocks, lds);
    const double rz = sum_partials(PSLOT(a, P_RZ), kCgBlocks, lds);
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
    const double beta = rz / a.scal[S_RZ0 + ((a.it - 1) & 1)];
    if (lead) a.scal[S_RZ0 + (a.it & 1)] = rz;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < a.nu; r += gridDim.x * blockDim.x)
        a.p[r] = a.r[r] / a.diag[r] + beta * a.p[r];
}

__global__ void k_scatter_x(int nu, const int32_t* __restrict__ urow, const double* __restrict__ x,
                            double* __restrict__ pnew)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nu) pnew[urow[r]] = x[r];
}

int ppe_solve(vox_ctx* c, vox_solve_info* info)
{
    const int nu = c->nu;
    CgArgs a;
    a.nu = nu; a.cap = c->a_cap; a.tol = c->p.cg_tol; a.maxit = c->p.cg_maxit; a.it = 0;
    a.npq = grid_for(nu > 0 ? nu : 1);
    if (a.npq > kMaxPartials) return fail(VOX_ERR_ARG, "PPE larger than the solver's row limit");
    a.rowlen = c->rowlen; a.acol = c->acol; a.aval = c->aval; a.diag = c->adiag; a.rhs = c->rhs;
    a.x = c->cg_x; a.r = c->cg_r; a.p = c->cg_p; a.q = c->cg_q;
    a.part = c->partials; a.scal = c->dscal; a.flags = c->dflags;
    VOX_TRY(fill_f64(c, c->f[VOX_F_PNEW], 0.0, c->tp));
    {
        ProfScope ps(c, KC_CG_INIT);
        k_cg_init<<<kCgBlocks, kBlock, 0, c->stream>>>(a);
        k_cg_init_fin<<<1, kBlock, 0, c->stream>>>(a);
        VOX_HIP(hipGetLastError());
    }
    // Check schedule: the first batch runs to two iterations short of the last
    // solve's count (consecutive steps need similar counts), then the flag is
    // read after every cg_check iterations (default 1), so no kernel is queued
    // after convergence.
    const int check = c->p.cg_check > 0 ? c->p.cg_check : 1;
    const int first = c->cg_expect - 2;
    int it = 0;
    VOX_TRY(read_flags(c));
    while (!c->hflags[FLAG_CG_DONE] && it < c->p.cg_maxit) {
        const int batch = (it == 0 && first > check) ? first : check;
        for (int k = 0; k < batch && it < c->p.cg_maxit; ++k) {
            a.it = ++it;
            {
                ProfScope ps(c, KC_SPMV);
                k_cg_spmv<<<a.npq, kBlock, 0, c->stream>>>(a);
            }
            {
                ProfScope ps(c, KC_CG_UPDATE);
                k_cg_update<<<kCgBlocks, kBlock, 0, c->stream>>>(a);
            }
            {
                ProfScope ps(c, KC_CG_P);
                k_cg_p<<<kCgBlocks, kBlock, 0, c->stream>>>(a);
            }
        }
        VOX_HIP(hipGetLastError());
        VOX_TRY(read_flags(c));
    }
    if (nu > 0) {
        k_scatter_x<<<grid_for(nu), kBlock, 0, c->stream>>>(nu, c->urow, c->cg_x, c->f[VOX_F_PNEW]);
        VOX_HIP(hipGetLastError());
    }
    const int iters = c->hflags[FLAG_CG_ITERS];
    const int conv = c->hflags[FLAG_CG_CONV];
    c->cg_expect = iters;
    c->st.last_iters = iters;
    c->st.last_relres = c->hscal[S_RELRES];
    if (info) {
        info->iterations = iters;
        info->converged = conv;
        info->unknowns = nu;
        info->nnz = (int32_t)c->hscal[S_NNZ];
        info->relres = c->hscal[S_RELRES];
        info->bnorm = c->hscal[S_BNORM];
    }
    if (c->hflags[FLAG_CG_BREAK]) return fail(VOX_ERR_SOLVER, "CG breakdown (p.Ap <= 0 or not finite)");
    if (!conv)
        return fail(VOX_ERR_SOLVER, "CG did not reach the relative residual " +
                                        std::to_string(c->p.cg_tol) + " in " +
                                        std::to_string(iters) + " iterations");
    return VOX_OK;
}

}  // namespace vox
