// vox_cgm.hip — Jacobi-preconditioned CG on a PPE held in the reference's
// device slot layout, for callers that keep the CUDA path's own arrays.
//
// The reference's GPU branch (main.cu:1159-1248) assembles the PPE into
// d_poiss / d_neighb, rows of `stride` slots (functions.cu:1007-1060): for a
// 1-based particle i, neighb[i*stride+1] = count+1, neighb[i*stride+l] for
// l = 2..count+1 is a 1-based neighbour id, poiss[i*stride+1] is the diagonal
// and poiss[i*stride+l] the coefficient of that neighbour. It then runs
// INCDECOM + cgm1..cgm12 with FP32 atomic dot products and one host read per
// iteration. vox_cgm_slot solves the same system in place with Jacobi-PCG:
// rows with bcon == 0 are the unknowns, a neighbour with bcon != 0 is a
// Dirichlet p = 0 (the reference's cold start leaves its pressure at zero), and
// pnew receives the solution from a zero start.
//
// Kernels (one fixed grid of kSlotBlocks blocks, partial sums per block summed
// in a fixed order by the next kernel; no atomics, run-to-run identical):
//   init     x = 0; r = b; p = r/d; partials b.b and r.p
//   spmv     q = A p, kSlotLanes lanes per row reading the row's slots as one
//            contiguous run; partial p.q
//   update   x += alpha p; r -= alpha q; partials r.r and r.(r/d)
//   dir      convergence test on the device; p = r/d + beta p
// Kernels launched after the device has set its stop flag return at entry,
// so the host reads the flags once per batch of iterations.
#include <mutex>

#include "vox_internal.hpp"

namespace vox {

constexpr int kSlotBlocks = 1024;
constexpr int kSlotLanes = 16;            // lanes per matrix row in the product
constexpr int kSlotRowsPerBlock = kBlock / kSlotLanes;

enum { SP_BB = 0, SP_RZ0, SP_PQ, SP_RR, SP_RZ, SP_COUNT };
enum { SS_BNORM = 0, SS_RZ0, SS_RZ1, SS_RELRES, SS_COUNT };
enum { SF_DONE = 0, SF_CONV, SF_BREAK, SF_ITERS, SF_BADROW, SF_COUNT };

// Workspace: three vectors of tp+1 doubles (each rounded up to 64), the
// partial slots, the scalars and the flags.
constexpr long slot_vec_len(long tp) { return ((tp + 1 + 63) / 64) * 64; }
constexpr long slot_ws_doubles(long tp)
{
    return 3 * slot_vec_len(tp) + (long)SP_COUNT * kSlotBlocks + 64;
}

struct SlotArgs {
    int tp, stride, maxit, it;
    double tol;
    const int32_t* __restrict__ neighb;
    const double* __restrict__ poiss;
    const double* __restrict__ b;
    const int32_t* __restrict__ bcon;
    double* x;
    double* r;
    double* p;
    double* q;
    double* part;     // SP_COUNT * kSlotBlocks
    double* scal;     // SS_COUNT
    int32_t* flags;   // SF_COUNT
};

static void slot_carve(SlotArgs& a, double* ws, long tp)
{
    const long v = slot_vec_len(tp);
    a.r = ws;
    a.p = ws + v;
    a.q = ws + 2 * v;
    a.part = ws + 3 * v;
    a.scal = a.part + (long)SP_COUNT * kSlotBlocks;
    a.flags = reinterpret_cast<int32_t*>(a.scal + 16);
}

#define SPART(a, s) ((a).part + (long)(s) * kSlotBlocks)

__device__ __forceinline__ double slot_sum(const double* part, double* lds)
{
    double s = 0.0;
    for (int k = threadIdx.x; k < kSlotBlocks; k += kBlock) s = s + part[k];
    return block_sum(s, lds);
}

__global__ __launch_bounds__(kBlock) void k_slot_init(SlotArgs a)
{
    __shared__ double lds[kBlock];
    double bb = 0.0, rz = 0.0;
    for (long i = 1 + blockIdx.x * (long)blockDim.x + threadIdx.x; i <= a.tp;
         i += (long)gridDim.x * blockDim.x) {
        a.x[i] = 0.0;
        double rv = 0.0, z = 0.0;
        if (a.bcon[i] == 0) {
            rv = a.b[i];
            z = rv / a.poiss[i * (long)a.stride + 1];
            bb += rv * rv;
            rz += rv * z;
        }
        a.r[i] = rv;
        a.p[i] = z;
        a.q[i] = 0.0;
    }
    bb = block_sum(bb, lds);
    rz = block_sum(rz, lds);
    if (threadIdx.x == 0) {
        SPART(a, SP_BB)[blockIdx.x] = bb;
        SPART(a, SP_RZ0)[blockIdx.x] = rz;
    }
}

__global__ __launch_bounds__(kBlock) void k_slot_init_fin(SlotArgs a)
{
    __shared__ double lds[kBlock];
    const double bb = slot_sum(SPART(a, SP_BB), lds);
    const double rz = slot_sum(SPART(a, SP_RZ0), lds);
    if (threadIdx.x == 0) {
        const double bn = sqrt(bb);
        a.scal[SS_BNORM] = bn;
        a.scal[SS_RZ0] = rz;
        a.scal[SS_RELRES] = 0.0;
        a.flags[SF_ITERS] = 0;
        a.flags[SF_BREAK] = 0;
        a.flags[SF_BADROW] = 0;
        a.flags[SF_CONV] = (bn == 0.0) ? 1 : 0;
        a.flags[SF_DONE] = (bn == 0.0) ? 1 : 0;
    }
}

// q = A p over the unknown rows. A group of kSlotLanes lanes takes one row:
// lane t reads slots 2+t, 2+t+16, ..., so the group reads the row's slots as
// one contiguous run. The group's sum is formed by a fixed shuffle tree, and
// the diagonal term is added last. Neighbour ids outside 1..tp and rows whose
// count does not fit the stride are skipped and reported.
__global__ __launch_bounds__(kBlock) void k_slot_spmv(SlotArgs a)
{
    __shared__ double lds[kBlock];
    if (a.flags[SF_DONE]) return;
    const int t = threadIdx.x & (kSlotLanes - 1);
    double pq = 0.0;
    for (long i = 1 + (long)blockIdx.x * kSlotRowsPerBlock + threadIdx.x / kSlotLanes; i <= a.tp;
         i += (long)gridDim.x * kSlotRowsPerBlock) {
        if (a.bcon[i] != 0) continue;          // uniform within the group
        const long base = i * (long)a.stride;
        int last = a.neighb[base + 1];          // count + 1
        if (last > a.stride - 1 || last < 1) {
            if (t == 0) a.flags[SF_BADROW] = 1;
            last = last < 1 ? 1 : a.stride - 1;
        }
        double acc = 0.0;
        for (int l = 2 + t; l <= last; l += kSlotLanes) {
            const int j = a.neighb[base + l];
            const bool ok = (j >= 1 && j <= a.tp);
            const int jj = ok ? j : (int)i;
            const double pj = a.p[jj];
            const bool use = ok && j != (int)i && a.bcon[jj] == 0;
            acc = acc + (use ? a.poiss[base + l] * pj : 0.0);
        }
#pragma unroll
        for (int s = kSlotLanes / 2; s > 0; s >>= 1) acc = acc + __shfl_xor(acc, s, kSlotLanes);
        if (t == 0) {
            const double pi = a.p[i];
            const double qi = acc + a.poiss[base + 1] * pi;
            a.q[i] = qi;
            pq += pi * qi;
        }
    }
    pq = block_sum(pq, lds);
    if (threadIdx.x == 0) SPART(a, SP_PQ)[blockIdx.x] = pq;
}

__global__ __launch_bounds__(kBlock) void k_slot_update(SlotArgs a)
{
    __shared__ double lds[kBlock];
    if (a.flags[SF_DONE]) return;
    const double pq = slot_sum(SPART(a, SP_PQ), lds);
    const double rz_old = a.scal[SS_RZ0 + ((a.it - 1) & 1)];
    if (!(pq > 0.0) || !isfinite(pq)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            a.flags[SF_BREAK] = 1;
            a.flags[SF_DONE] = 1;
            a.flags[SF_ITERS] = a.it - 1;
        }
        return;
    }
    const double alpha = rz_old / pq;
    double rr = 0.0, rz = 0.0;
    for (long i = 1 + blockIdx.x * (long)blockDim.x + threadIdx.x; i <= a.tp;
         i += (long)gridDim.x * blockDim.x) {
        if (a.bcon[i] != 0) continue;
        a.x[i] = a.x[i] + alpha * a.p[i];
        const double res = a.r[i] - alpha * a.q[i];
        a.r[i] = res;
        const double z = res / a.poiss[i * (long)a.stride + 1];
        rr += res * res;
        rz += res * z;
    }
    rr = block_sum(rr, lds);
    rz = block_sum(rz, lds);
    if (threadIdx.x == 0) {
        SPART(a, SP_RR)[blockIdx.x] = rr;
        SPART(a, SP_RZ)[blockIdx.x] = rz;
    }
}

__global__ __launch_bounds__(kBlock) void k_slot_dir(SlotArgs a)
{
    __shared__ double lds[kBlock];
    if (a.flags[SF_DONE]) return;
    const double rr = slot_sum(SPART(a, SP_RR), lds);
    const double rz = slot_sum(SPART(a, SP_RZ), lds);
    const double rel = sqrt(rr) / a.scal[SS_BNORM];
    const bool lead = (blockIdx.x == 0 && threadIdx.x == 0);
    if (lead) {
        a.scal[SS_RELRES] = rel;
        a.flags[SF_ITERS] = a.it;
    }
    if (rel <= a.tol) {
        if (lead) { a.flags[SF_CONV] = 1; a.flags[SF_DONE] = 1; }
        return;
    }
    if (a.it >= a.maxit || !isfinite(rel)) {
        if (lead) a.flags[SF_DONE] = 1;
        return;
    }
    const double beta = rz / a.scal[SS_RZ0 + ((a.it - 1) & 1)];
    if (lead) a.scal[SS_RZ0 + (a.it & 1)] = rz;
    for (long i = 1 + blockIdx.x * (long)blockDim.x + threadIdx.x; i <= a.tp;
         i += (long)gridDim.x * blockDim.x) {
        if (a.bcon[i] != 0) continue;
        a.p[i] = a.r[i] / a.poiss[i * (long)a.stride + 1] + beta * a.p[i];
    }
}

static int slot_check_args(int tp, int stride, const int32_t* neighb, const double* poiss,
                           const double* source, const int32_t* bcon, double* pnew, double tol,
                           int maxit)
{
    if (tp < 1) return fail(VOX_ERR_ARG, "vox_cgm_slot: tp must be >= 1");
    if (stride < 3) return fail(VOX_ERR_ARG, "vox_cgm_slot: stride must be >= 3");
    if (!neighb || !poiss || !source || !bcon || !pnew)
        return fail(VOX_ERR_ARG, "vox_cgm_slot: null device array");
    if (!(tol > 0.0)) return fail(VOX_ERR_ARG, "vox_cgm_slot: tol must be > 0");
    if (maxit < 1) return fail(VOX_ERR_ARG, "vox_cgm_slot: maxit must be >= 1");
    if ((long)(tp + 1) * stride > (1L << 40))
        return fail(VOX_ERR_ARG, "vox_cgm_slot: tp * stride too large");
    return VOX_OK;
}

// Runs the solve on the null stream with the workspace `ws`
// (slot_ws_doubles(tp) doubles of device memory) and reports through info.
static int slot_run(int tp, int stride, const int32_t* neighb, const double* poiss,
                    const double* source, const int32_t* bcon, double* pnew, double tol, int maxit,
                    vox_solve_info* info, double* ws)
{
    SlotArgs a;
    a.tp = tp; a.stride = stride; a.maxit = maxit; a.it = 0; a.tol = tol;
    a.neighb = neighb; a.poiss = poiss; a.b = source; a.bcon = bcon; a.x = pnew;
    slot_carve(a, ws, tp);
    hipStream_t s = nullptr;
    k_slot_init<<<kSlotBlocks, kBlock, 0, s>>>(a);
    k_slot_init_fin<<<1, kBlock, 0, s>>>(a);
    VOX_HIP(hipGetLastError());
    int32_t fl[SF_COUNT] = {0, 0, 0, 0, 0};
    double sc[SS_COUNT] = {0, 0, 0, 0};
    int it = 0;
    const int batch = 8;
    while (!fl[SF_DONE] && it < maxit) {
        for (int k = 0; k < batch && it < maxit; ++k) {
            a.it = ++it;
            k_slot_spmv<<<kSlotBlocks, kBlock, 0, s>>>(a);
            k_slot_update<<<kSlotBlocks, kBlock, 0, s>>>(a);
            k_slot_dir<<<kSlotBlocks, kBlock, 0, s>>>(a);
        }
        VOX_HIP(hipGetLastError());
        VOX_HIP(hipMemcpy(fl, a.flags, sizeof(fl), hipMemcpyDeviceToHost));
    }
    VOX_HIP(hipMemcpy(fl, a.flags, sizeof(fl), hipMemcpyDeviceToHost));
    VOX_HIP(hipMemcpy(sc, a.scal, sizeof(sc), hipMemcpyDeviceToHost));
    if (info) {
        info->iterations = fl[SF_ITERS];
        info->converged = fl[SF_CONV];
        info->unknowns = -1;
        info->nnz = -1;
        info->relres = sc[SS_RELRES];
        info->bnorm = sc[SS_BNORM];
    }
    if (fl[SF_BADROW])
        return fail(VOX_ERR_ARG, "vox_cgm_slot: a row count exceeds the slot stride");
    if (fl[SF_BREAK]) return fail(VOX_ERR_SOLVER, "CG breakdown (p.Ap <= 0 or not finite)");
    if (!fl[SF_CONV])
        return fail(VOX_ERR_SOLVER, "vox_cgm_slot did not reach the relative residual in " +
                                        std::to_string(fl[SF_ITERS]) + " iterations");
    return VOX_OK;
}

// ---- allocation-free variant ---------------------------------------------------------
// For callers that cannot allocate device memory (code under stream capture,
// serving loops): the workspace is this file-scope array, one copy per GPU,
// sized for tp <= kNoAllocMaxTp (24.1 MiB). Its device address is looked up
// once per device with that device current; host code uses that address,
// never the array's own name. Every call on a GPU shares the array, so calls
// that overlap on one GPU race; each call initialises everything it reads.
constexpr long kNoAllocMaxTp = (1L << 20) - 1;
static __device__ __align__(256) double scratch[slot_ws_doubles(kNoAllocMaxTp)];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};

static int scratch_for_current_device(double** out)
{
    int dev = 0;
    VOX_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return fail(VOX_ERR_ARG, "vox_cgm_slot_noalloc: device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        VOX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return fail(VOX_ERR_HIP, "hipGetSymbolAddress returned a null scratch address");
        g_scratch_ptr[dev] = static_cast<double*>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return VOX_OK;
}

}  // namespace vox

extern "C" {

int vox_cgm_slot_noalloc(int tp, int stride, const int32_t* neighb, const double* poiss,
                         const double* source, const int32_t* bcon, double* pnew, double tol,
                         int maxit, vox_solve_info* info)
{
    using namespace vox;
    VOX_TRY(slot_check_args(tp, stride, neighb, poiss, source, bcon, pnew, tol, maxit));
    if (tp > kNoAllocMaxTp)
        return fail(VOX_ERR_ARG, "vox_cgm_slot_noalloc: tp = " + std::to_string(tp) +
                                     " exceeds the static workspace limit of " +
                                     std::to_string(kNoAllocMaxTp) + " particles");
    double* ws = nullptr;
    VOX_TRY(scratch_for_current_device(&ws));
    return slot_run(tp, stride, neighb, poiss, source, bcon, pnew, tol, maxit, info, ws);
}

int vox_cgm_slot(int tp, int stride, const int32_t* neighb, const double* poiss,
                 const double* source, const int32_t* bcon, double* pnew, double tol, int maxit,
                 vox_solve_info* info)
{
    using namespace vox;
    VOX_TRY(slot_check_args(tp, stride, neighb, poiss, source, bcon, pnew, tol, maxit));
    double* ws = nullptr;
    VOX_HIP(hipMalloc((void**)&ws, sizeof(double) * (size_t)slot_ws_doubles(tp)));
    const int rc = slot_run(tp, stride, neighb, poiss, source, bcon, pnew, tol, maxit, info, ws);
    const std::string err = rc != VOX_OK ? std::string(vox_last_error()) : std::string();
    const hipError_t e = hipFree(ws);
    if (rc != VOX_OK) return fail(rc, err);
    if (e != hipSuccess) return fail(VOX_ERR_HIP, std::string("hipFree: ") + hipGetErrorString(e));
    return VOX_OK;
}

}  // extern "C"
