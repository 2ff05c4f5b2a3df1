// vox_ppe.hip — pressure Poisson equation on gfx950: device-resident sparse
// assembly and FP64 Jacobi-preconditioned CG.
//
// Assembly replaces MATRIX (functions.cu:869-1005, Method 1, no matopt/srcopt):
// rows are the bcon == 0 particles renumbered 0..nu-1 in particle order; the
// off-diagonal a_ij = -((2*dim/lambda)*W(d_ij)/n0/rho1) is kept for unknown
// columns, free-surface (bcon == 1) neighbours only add to the diagonal and
// dummy (bcon == -1) neighbours are skipped; a_ii = sum of |a_ij| in list order
// plus 1e-7/dt/dt. Storage is slot-major ELL (col/val[k*tp + r]) so a wave's
// loads of slot k are contiguous; only nonzero unknown columns are stored.
//
// The solver replaces INCDECOM + CGM (functions.cu:1139-1200, 506-716). One CG
// iteration is three kernels, each a single pass over its vectors:
//   K1 q = A p, partial p.q                       -> alpha = rz / p.q
//   K2 x += alpha p, r -= alpha q, r.r, r.(r/d)   -> beta, convergence test
//   K3 p = r/d + beta p
// Dot products reduce in a fixed order (block tree, then the last block sums
// the block partials in block order), so results are run-to-run identical.
// Scalars and the stop flag stay on the device: the host checks the flag every
// cg_check iterations only.
#include "vox_internal.hpp"

namespace vox {

constexpr int kCgGrid = 2048;   // persistent-ish grid for the CG kernels

__global__ void k_uflag(int n, const int32_t* __restrict__ bcon, int32_t* __restrict__ flag)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = (bcon[i] == 0) ? 1 : 0;
}

__global__ void k_umap(int n, const int32_t* __restrict__ bcon, int32_t* __restrict__ uidx,
                       int32_t* __restrict__ urow)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (bcon[i] == 0) {
        int r = uidx[i];
        urow[r] = i;
    } else {
        uidx[i] = -1;
    }
}

struct AsmArgs {
    int n, dim, ktype;
    double re, c0, n0l, rho1, dshift, inv_dt_dt, n0s;
    ListView L;
    const double *x, *y, *z, *pnd;
    const int32_t *bcon, *uidx;
    int32_t *acol, *rowlen;
    double *aval, *adiag, *rhs, *source;
};

__global__ __launch_bounds__(kBlock) void k_assemble(AsmArgs a)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int bi = a.bcon[i];
    // right-hand side of the default source term (functions.cu:990-1001)
    double src = 0.0;
    if (bi == 0) src = a.inv_dt_dt * (a.pnd[i] - a.n0s) / a.n0s;
    a.source[i] = src;
    if (bi != 0) return;
    const int r = a.uidx[i];
    const long stride = a.n;
    const int dim = a.dim;
    double xi = a.x[i], yi = a.y[i], zi = (dim == 3) ? a.z[i] : 0.0;
    double diag = 0.0;
    int k = 0;
    int cnt = a.L.cnt[i];
    for (int l = 0; l < cnt; ++l) {
        int j = a.L.ids[(long)l * a.L.stride + i];
        if (j == i) continue;
        int bj = a.bcon[j];
        if (bj == -1) continue;
        double dx = a.x[j] - xi, dy = a.y[j] - yi, d;
        if (dim == 3) { double dz = a.z[j] - zi; d = sqrt(dx * dx + dy * dy + dz * dz); }
        else d = sqrt(dx * dx + dy * dy);
        if (fabs(d) <= 1.0e-15) continue;
        double val = a.c0 * wfun(d, a.ktype, a.re) / a.n0l / a.rho1;
        if (bj == 0 && val != 0.0) {
            a.acol[(long)k * stride + r] = a.uidx[j];
            a.aval[(long)k * stride + r] = -val;
            ++k;
        }
        diag += val;
    }
    diag += a.dshift;
    a.adiag[r] = diag;
    a.rowlen[r] = k;
    a.rhs[r] = src;
}

int ppe_assemble(vox_ctx* c)
{
    ProfScope ps(c, KC_ASSEMBLE);
    const int n = c->tp;
    const DevList& L = list_of(c, VOX_LIST_L);
    if (c->a_cap < L.cap) {
        if (c->acol) (void)hipFree(c->acol);
        if (c->aval) (void)hipFree(c->aval);
        c->acol = nullptr;
        c->aval = nullptr;
        VOX_HIP(hipMalloc(&c->acol, sizeof(int32_t) * (size_t)L.cap * n));
        VOX_HIP(hipMalloc(&c->aval, sizeof(double) * (size_t)L.cap * n));
        c->a_cap = L.cap;
    }
    k_uflag<<<grid_for(n), kBlock, 0, c->stream>>>(n, c->bcon, c->uflag);
    VOX_TRY(scan_exclusive(c, c->uflag, c->uidx, n, c->dflags + FLAG_NU));
    k_umap<<<grid_for(n), kBlock, 0, c->stream>>>(n, c->bcon, c->uidx, c->urow);
    const vox_params& p = c->p;
    AsmArgs a;
    a.n = n; a.dim = p.dim; a.ktype = p.ktype; a.re = p.re_l;
    a.c0 = 2.0 * p.dim / p.lambda;
    a.n0l = c->n0l; a.rho1 = p.rho1;
    a.dshift = (1.00e-7) / p.dt / p.dt;
    a.inv_dt_dt = 1.0 / p.dt / p.dt;
    a.n0s = c->n0s;
    a.L = view(c, VOX_LIST_L);
    a.x = c->f[VOX_F_X]; a.y = c->f[VOX_F_Y]; a.z = c->f[VOX_F_Z]; a.pnd = c->f[VOX_F_N];
    a.bcon = c->bcon; a.uidx = c->uidx;
    a.acol = c->acol; a.rowlen = c->rowlen; a.aval = c->aval; a.adiag = c->adiag; a.rhs = c->rhs;
    a.source = c->f[VOX_F_SOURCE];
    k_assemble<<<grid_for(n), kBlock, 0, c->stream>>>(a);
    VOX_HIP(hipGetLastError());
    VOX_TRY(read_flags(c));
    c->nu = c->hflags[FLAG_NU];
    return VOX_OK;
}

// ---- CG kernels ------------------------------------------------------------------------------------
struct CgArgs {
    int nu;
    long stride;
    double tol;
    int maxit;
    const int32_t* __restrict__ rowlen;
    const int32_t* __restrict__ acol;
    const double* __restrict__ aval;
    const double* __restrict__ diag;
    const double* __restrict__ rhs;
    double* x;
    double* r;
    double* p;
    double* q;
    double* partials;
    double* scal;
    int32_t* flags;
};

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
    double v[3] = {bb, rz, nz};
    grid_reduce<3>(v, a.partials, &a.flags[FLAG_COUNTER0], lds, [&](const double (&t)[3]) {
        double bn = sqrt(t[0]);
        a.scal[SC_BB] = t[0];
        a.scal[SC_BNORM] = bn;
        a.scal[SC_RZ] = t[1];
        a.scal[SC_NNZ] = t[2];
        a.scal[SC_RELRES] = 0.0;
        a.flags[FLAG_CG_ITERS] = 0;
        a.flags[FLAG_CG_BREAK] = 0;
        a.flags[FLAG_CG_CONV] = (bn == 0.0) ? 1 : 0;
        a.flags[FLAG_CG_DONE] = (bn == 0.0) ? 1 : 0;
    });
}

__global__ __launch_bounds__(kBlock) void k_cg_spmv(CgArgs a)
{
    __shared__ double lds[kBlock];
    if (load_agent_i(&a.flags[FLAG_CG_DONE])) return;
    double pq = 0.0;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < a.nu; r += gridDim.x * blockDim.x) {
        double pr = a.p[r];
        double acc = a.diag[r] * pr;
        int len = a.rowlen[r];
        for (int k = 0; k < len; ++k) {
            long o = (long)k * a.stride + r;
            acc = acc + a.aval[o] * a.p[a.acol[o]];
        }
        a.q[r] = acc;
        pq += pr * acc;
    }
    double v[1] = {pq};
    grid_reduce<1>(v, a.partials, &a.flags[FLAG_COUNTER1], lds, [&](const double (&t)[1]) {
        double d = t[0];
        if (!(d > 0.0) || !isfinite(d)) {
            a.flags[FLAG_CG_BREAK] = 1;
            a.flags[FLAG_CG_DONE] = 1;
        } else {
            a.scal[SC_PQ] = d;
            a.scal[SC_ALPHA] = a.scal[SC_RZ] / d;
        }
    });
}

__global__ __launch_bounds__(kBlock) void k_cg_update(CgArgs a)
{
    __shared__ double lds[kBlock];
    if (load_agent_i(&a.flags[FLAG_CG_DONE])) return;
    const double alpha = load_agent(&a.scal[SC_ALPHA]);
    double rr = 0.0, rz = 0.0;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < a.nu; r += gridDim.x * blockDim.x) {
        double pr = a.p[r];
        a.x[r] = a.x[r] + alpha * pr;
        double res = a.r[r] - alpha * a.q[r];
        a.r[r] = res;
        double z = res / a.diag[r];
        rr += res * res;
        rz += res * z;
    }
    double v[2] = {rr, rz};
    grid_reduce<2>(v, a.partials, &a.flags[FLAG_COUNTER2], lds, [&](const double (&t)[2]) {
        int it = a.flags[FLAG_CG_ITERS] + 1;
        a.flags[FLAG_CG_ITERS] = it;
        double rel = sqrt(t[0]) / a.scal[SC_BNORM];
        a.scal[SC_RR] = t[0];
        a.scal[SC_RELRES] = rel;
        if (rel <= a.tol) {
            a.flags[FLAG_CG_CONV] = 1;
            a.flags[FLAG_CG_DONE] = 1;
        } else if (it >= a.maxit || !isfinite(rel)) {
            a.flags[FLAG_CG_DONE] = 1;
        } else {
            a.scal[SC_BETA] = t[1] / a.scal[SC_RZ];
            a.scal[SC_RZ] = t[1];
        }
    });
}

__global__ __launch_bounds__(kBlock) void k_cg_p(CgArgs a)
{
    if (load_agent_i(&a.flags[FLAG_CG_DONE])) return;
    const double beta = load_agent(&a.scal[SC_BETA]);
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
    a.nu = nu; a.stride = c->tp; a.tol = c->p.cg_tol; a.maxit = c->p.cg_maxit;
    a.rowlen = c->rowlen; a.acol = c->acol; a.aval = c->aval; a.diag = c->adiag; a.rhs = c->rhs;
    a.x = c->cg_x; a.r = c->cg_r; a.p = c->cg_p; a.q = c->cg_q;
    a.partials = c->partials; a.scal = c->dscal; a.flags = c->dflags;
    int g = grid_for(nu > 0 ? nu : 1);
    if (g > kCgGrid) g = kCgGrid;
    VOX_TRY(fill_f64(c, c->f[VOX_F_PNEW], 0.0, c->tp));
    {
        ProfScope ps(c, KC_CG_INIT);
        k_cg_init<<<g, kBlock, 0, c->stream>>>(a);
        VOX_HIP(hipGetLastError());
    }
    // Check schedule: the first batch runs to two iterations short of the last
    // solve's count (consecutive steps need similar counts), then the flag is
    // read after every cg_check iterations (default 1), so no kernel is queued
    // after convergence.
    const int check = c->p.cg_check > 0 ? c->p.cg_check : 1;
    int first = c->cg_expect - 2;
    int launched = 0;
    VOX_TRY(read_flags(c));
    while (!c->hflags[FLAG_CG_DONE] && launched < c->p.cg_maxit) {
        int batch = (launched == 0 && first > check) ? first : check;
        for (int k = 0; k < batch && launched < c->p.cg_maxit; ++k, ++launched) {
            {
                ProfScope ps(c, KC_SPMV);
                k_cg_spmv<<<g, kBlock, 0, c->stream>>>(a);
            }
            {
                ProfScope ps(c, KC_CG_UPDATE);
                k_cg_update<<<g, kBlock, 0, c->stream>>>(a);
            }
            {
                ProfScope ps(c, KC_CG_P);
                k_cg_p<<<g, kBlock, 0, c->stream>>>(a);
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
    c->st.last_relres = c->hscal[SC_RELRES];
    if (info) {
        info->iterations = iters;
        info->converged = conv;
        info->unknowns = nu;
        info->nnz = (int32_t)c->hscal[SC_NNZ];
        info->relres = c->hscal[SC_RELRES];
        info->bnorm = c->hscal[SC_BNORM];
    }
    if (c->hflags[FLAG_CG_BREAK]) return fail(VOX_ERR_SOLVER, "CG breakdown (p.Ap <= 0 or not finite)");
    if (!conv)
        return fail(VOX_ERR_SOLVER, "CG did not reach the relative residual " +
                                        std::to_string(c->p.cg_tol) + " in " +
                                        std::to_string(iters) + " iterations");
    return VOX_OK;
}

int ppe_export(vox_ctx* c, int32_t* nu_out, int32_t* nnz_out, int32_t* rowptr, int32_t* cols,
               double* vals, double* rhs, int32_t* ids)
{
    const int nu = c->nu;
    const long n = c->tp;
    std::vector<int32_t> len(nu > 0 ? nu : 1);
    if (nu > 0)
        VOX_HIP(hipMemcpy(len.data(), c->rowlen, sizeof(int32_t) * nu, hipMemcpyDeviceToHost));
    long nnz = 0;
    for (int r = 0; r < nu; ++r) nnz += len[r] + 1;
    if (nu_out) *nu_out = nu;
    if (nnz_out) *nnz_out = (int32_t)nnz;
    if (!rowptr && !cols && !vals && !rhs && !ids) return VOX_OK;
    int maxlen = 0;
    for (int r = 0; r < nu; ++r) maxlen = len[r] > maxlen ? len[r] : maxlen;
    std::vector<int32_t> acol((size_t)maxlen * n);
    std::vector<double> aval((size_t)maxlen * n);
    std::vector<double> diag(nu > 0 ? nu : 1), b(nu > 0 ? nu : 1);
    std::vector<int32_t> urow(nu > 0 ? nu : 1);
    if (maxlen > 0) {
        VOX_HIP(hipMemcpy(acol.data(), c->acol, sizeof(int32_t) * maxlen * n, hipMemcpyDeviceToHost));
        VOX_HIP(hipMemcpy(aval.data(), c->aval, sizeof(double) * maxlen * n, hipMemcpyDeviceToHost));
    }
    if (nu > 0) {
        VOX_HIP(hipMemcpy(diag.data(), c->adiag, sizeof(double) * nu, hipMemcpyDeviceToHost));
        VOX_HIP(hipMemcpy(b.data(), c->rhs, sizeof(double) * nu, hipMemcpyDeviceToHost));
        VOX_HIP(hipMemcpy(urow.data(), c->urow, sizeof(int32_t) * nu, hipMemcpyDeviceToHost));
    }
    long k = 0;
    for (int r = 0; r < nu; ++r) {
        if (rowptr) rowptr[r] = (int32_t)k;
        if (cols) cols[k] = r;
        if (vals) vals[k] = diag[r];
        ++k;
        for (int s = 0; s < len[r]; ++s) {
            if (cols) cols[k] = acol[(size_t)s * n + r];
            if (vals) vals[k] = aval[(size_t)s * n + r];
            ++k;
        }
        if (rhs) rhs[r] = b[r];
        if (ids) ids[r] = urow[r];
    }
    if (rowptr) rowptr[nu] = (int32_t)k;
    return VOX_OK;
}

}  // namespace vox
