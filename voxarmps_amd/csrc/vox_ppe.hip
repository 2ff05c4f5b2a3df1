// vox_ppe.hip — pressure Poisson equation on gfx950: device-resident sparse
// assembly and FP64 Jacobi-preconditioned CG.
//
// Assembly replaces MATRIX (functions.cu:869-1005, Method 1, no matopt/srcopt):
// rows are the bcon == 0 particles renumbered 0..nu-1 in particle order; the
// off-diagonal a_ij = -((2*dim/lambda)*W(d_ij)/n0/rho1) is kept for unknown
// columns, free-surface (bcon == 1) neighbours only add to the diagonal and
// dummy (bcon == -1) neighbours are skipped; a_ii = sum of |a_ij| in list order
// plus 1e-7/dt/dt. Storage is SELL-64 (mx_index): a wave's entry k of its 64
// rows is one contiguous block, which is what the product reads; only nonzero
// unknown columns are stored. The solver lives in vox_cg.hip.
//
#include <algorithm>

#include "vox_internal.hpp"

namespace vox {

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

// ---- numbering of the unknowns -------------------------------------------------------
// The unknowns are the bcon = 0 particles, numbered in the cell order of list
// L's grid with z fastest (then y, then x; inside a cell by particle id), so
// the assembly's rows and the solver's vectors run through space cell by
// cell: a wave's 64 rows are neighbours, and so are most of their columns.
// The reference numbers them in particle order (functions.cu:869-925); the
// export maps back to it. Before the first NEIGHBOR, and for a dump of the
// learned network's training systems, particle order is used.
__device__ __forceinline__ int zkey(const Grid& g, int c)
{
    const int ix = c % g.ncx, iy = (c / g.ncx) % g.ncy, iz = c / (g.ncx * g.ncy);
    return iz + g.ncz * (iy + g.ncy * ix);
}

__global__ void k_zcount(Grid g, const int32_t* __restrict__ start, int32_t* __restrict__ zcnt)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < g.tnc) zcnt[zkey(g, (int)c)] = start[c + 1] - start[c];
}

// Position t of the list's cell-sorted records goes to its cell's start in
// z-fastest order plus its rank inside the cell.
__global__ void k_uorder(int n, Grid g, const PosRec* __restrict__ rec,
                         const int32_t* __restrict__ start, const int32_t* __restrict__ zstart,
                         const int32_t* __restrict__ bcon, int32_t* __restrict__ uflag,
                         int32_t* __restrict__ uorder)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const PosRec r = rec[t];
    const int np = zstart[zkey(g, r.aux)] + (t - start[r.aux]);
    uorder[np] = r.id;
    uflag[np] = (bcon[r.id] == 0) ? 1 : 0;
}

__global__ void k_umap_order(int n, const int32_t* __restrict__ uorder,
                             const int32_t* __restrict__ uflag, const int32_t* __restrict__ uscan,
                             int32_t* __restrict__ uidx, int32_t* __restrict__ urow)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int id = uorder[t];
    if (uflag[t]) {
        const int r = uscan[t];
        uidx[id] = r;
        urow[r] = id;
    } else {
        uidx[id] = -1;
    }
}

static int number_unknowns(vox_ctx* c)
{
    const int n = c->tp;
    const DevList& L = list_of(c, VOX_LIST_L);
    if (c->uorder_particles || !L.sorted) {
        k_uflag<<<grid_for(n), kBlock, 0, c->stream>>>(n, c->bcon, c->uflag);
        VOX_TRY(scan_exclusive(c, c->uflag, c->uidx, n, c->dflags + FLAG_NU));
        k_umap<<<grid_for(n), kBlock, 0, c->stream>>>(n, c->bcon, c->uidx, c->urow);
        VOX_HIP(hipGetLastError());
        c->uorder_cells = false;
        return VOX_OK;
    }
    const long tnc = L.g.tnc;
    if (c->zcell_alloc < tnc) {
        if (c->zcell) (void)hipFree(c->zcell);
        c->zcell = nullptr;
        VOX_HIP(hipMalloc(&c->zcell, sizeof(int32_t) * 2 * (size_t)(tnc + 1)));
        c->zcell_alloc = tnc;
    }
    int32_t* zcnt = c->zcell;
    int32_t* zstart = c->zcell + (tnc + 1);
    k_zcount<<<grid_for(tnc), kBlock, 0, c->stream>>>(L.g, L.cell_start, zcnt);
    VOX_TRY(scan_exclusive(c, zcnt, zstart, tnc, zstart + tnc));
    k_uorder<<<grid_for(n), kBlock, 0, c->stream>>>(n, L.g, L.rec, L.cell_start, zstart, c->bcon,
                                                     c->uflag, c->uorder);
    VOX_TRY(scan_exclusive(c, c->uflag, c->uscan, n, c->dflags + FLAG_NU));
    k_umap_order<<<grid_for(n), kBlock, 0, c->stream>>>(n, c->uorder, c->uflag, c->uscan, c->uidx,
                                                         c->urow);
    VOX_HIP(hipGetLastError());
    c->uorder_cells = true;
    return VOX_OK;
}

struct AsmArgs {
    int n, dim, ktype, acap, matopt;
    double re, c0, n0l, rho1, dshift, inv_dt_dt, n0s;
    double hl;                             // matopt: 3 re (2D) or 2 re (3D)
    double inv_re, inv_n0l, inv_rho1;     // correctly rounded reciprocals
    ListView L;
    const double *x, *y, *z, *pnd;
    const int32_t *bcon, *uidx, *urow;
    const int32_t* nu;                     // device: the number of unknowns
    int32_t *acol, *rowlen;                // acol null: the columns are not stored
    v4u* dcol;                             // 16-bit column steps (null: not stored)
    int32_t *cbase, *dbad;                 //   first columns; set when a step does not fit
    int32_t* creach;                       //   max row - first column, max last column - row
    double *aval, *adiag, *rhs, *source;
    float* aval32;                         // FP32 values (mixed precision, learned network)
    PosRec* rec;                           // the records the row entries index
    const PosRec* order;                   // rows of positions: list L's sorted records
    const int32_t* pos_of;                 //   and each particle's position there (else null)
};

// 32-byte gather record per particle: position, bcon (id slot), unknown
// index - at the particle's id, or at its position in list L's sorted records
// when the rows hold positions; and the source term of the default MATRIX
// (functions.cu:990-1001), 0 for every particle that is not an unknown.
__global__ void k_pack_asm(AsmArgs a)
{
    // threads in particle order (fields read in lines), whole records stored
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int t = a.order ? a.pos_of[i] : i;
    PosRec r;
    r.x = a.x[i]; r.y = a.y[i]; r.z = (a.dim == 3) ? a.z[i] : 0.0;
    const int bi = a.bcon[i];
    r.id = bi;
    r.aux = a.uidx[i];
    a.rec[t] = r;
    a.source[i] = (bi == 0) ? a.inv_dt_dt * (a.pnd[i] - a.n0s) / a.n0s : 0.0;
}

// Per-thread ring of the last 16 outputs in LDS (strides of 17 keep the lanes'
// equal-slot accesses on distinct banks).
constexpr int kRing = 17;

// One thread per unknown row r (particle urow[r]), so a wave is one SELL-64
// slice: its lanes produce their rows' entries 8 at a time into LDS, and once
// every lane holds its group g (or has run out of row), the wave stores entry
// 8g + u of its 64 rows as one contiguous block for u = 0..7. Rows are written
// in their final layout in one pass. Entries are processed strictly in list
// order (the reference's), so the values and the diagonal are bit-identical
// to the sequential MATRIX.
// SV = double: FP64 values (and their FP32 copy when aval32 is set, for the
// learned preconditioner's network); SV = float: mixed precision, FP32 values
// only (computed in FP64 and rounded once; the diagonal stays FP64).
template <class SV, int KT>
__global__ __launch_bounds__(kBlock) void k_assemble_rows(AsmArgs a)
{
    __shared__ int32_t ring_c[kBlock * kRing];
    __shared__ double ring_v[kBlock * kRing];
    const int nu = *a.nu;
    if ((long)blockIdx.x * blockDim.x >= nu) return;          // uniform over the block
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = r < nu;
    const int i = live ? a.urow[r] : 0;
    const int self = a.pos_of ? a.pos_of[i] : i;   // own index among the entries' records
    int32_t* rc = ring_c + threadIdx.x * kRing;
    double* rv = ring_v + threadIdx.x * kRing;
    // rows of positions sit at the particle's position
    const int32_t* __restrict__ row = a.L.ids + sl_index(self, 0, a.L.cap);
    const PosRec me = a.rec[self];
    const double xi = me.x, yi = me.y, zi = me.z;
    const long ob = mx_index(r, 0, a.acap);                  // entry k at ob + 64 k
    double diag = 0.0;
    int k = 0, l0 = 0;
    int first = 0, last = 0;                                  // column steps
    bool bad = false;
    const int cnt = live ? a.L.cnt[self] : 0;
    constexpr int U = 8;
    // the next chunk's ids are loaded while the current chunk's records are
    // gathered and used
    int32_t jn[U];
    if (cnt > 0) load_chunk8(row, 0, jn);
    for (int g = 0;; ++g) {
        // this lane's rows until group g is complete or the row is done; a
        // list chunk adds at most 8 entries, so the ring never holds more
        // than groups g and g + 1
        while (k < 8 * (g + 1) && l0 < cnt) {
            int jj[U];
            PosRec rr[U];
#pragma unroll
            for (int u = 0; u < U; ++u) jj[u] = jn[u];
            if (l0 + U < cnt) load_chunk8(row, (l0 >> 3) + 1, jn);
#pragma unroll
            for (int u = 0; u < U; ++u) jj[u] = (l0 + u < cnt) ? jj[u] : self;
#pragma unroll
            for (int u = 0; u < U; ++u) rr[u] = a.rec[jj[u]];
            // Branch-free per entry (a skipped entry keeps diag unchanged
            // through a select), so no record load is sunk into a branch
            // behind a full wait.
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const PosRec rj = rr[u];
                const int bj = rj.id;
                // z is 0 in 2D records, and adding 0*0 leaves the 2D sum
                // unchanged, so one expression serves both dimensions
                const double dx = rj.x - xi, dy = rj.y - yi, dz = rj.z - zi;
                const double d = sqrt(dx * dx + dy * dy + dz * dz);
                // jj == self: the particle itself, or padding past the row end
                const bool use = (jj[u] != self) && (bj != -1) && !(fabs(d) <= 1.0e-15);
                // (= c0 * W(d) / n0l / rho1 with IEEE divisions, see div_by);
                // matopt: hl / d^3 / n0l / rho1 (functions.cu:903-913)
                const double val = a.matopt
                    ? div_by(div_by(a.hl / (d * d * d), a.n0l, a.inv_n0l), a.rho1, a.inv_rho1)
                    : div_by(div_by(a.c0 * wfun_k<KT>(d, a.ktype, a.re, a.inv_re), a.n0l, a.inv_n0l),
                             a.rho1, a.inv_rho1);
                if (use && bj == 0 && val != 0.0) {
                    rc[k & 15] = rj.aux;
                    rv[k & 15] = -val;
                    ++k;
                }
                diag = use ? diag + val : diag;
            }
            l0 += U;
        }
        const bool has = k > 8 * g;
        if (!__any(has)) break;                                // every row is written
        if (has) {
            const int s0 = (g & 1) * 8;
            // the group's columns also as steps from the previous column
            // (entry 0 from itself: its column is the row's base); steps past
            // the row's end are 0
            int prev = (g == 0) ? rc[s0] : last;
            if (g == 0) first = prev;
            uint32_t w[4] = {0u, 0u, 0u, 0u};
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int kk = 8 * g + u;
                if (kk < k) {
                    const long o = ob + 64L * kk;
                    const int cu = rc[s0 + u];
                    if (a.acol) a.acol[o] = cu;
                    const int d = cu - prev;
                    bad = bad || d < 0 || d > 0xffff;
                    w[u >> 1] |= (uint32_t)(d & 0xffff) << (16 * (u & 1));
                    prev = cu;
                    if constexpr (sizeof(SV) == 8) {
                        a.aval[o] = rv[s0 + u];
                        if (a.aval32) a.aval32[o] = (float)rv[s0 + u];
                    } else {
                        a.aval32[o] = (float)rv[s0 + u];
                    }
                }
            }
            last = prev;
            if (a.dcol) a.dcol[dc_index(r, g, a.acap)] = v4u{w[0], w[1], w[2], w[3]};
        }
    }
    if (a.dcol) {
        // how far the rows' columns reach below and above the row (every lane
        // of the wave is here: the loop above ends for all of them together)
        int below = (live && k > 0) ? r - first : 0;
        int above = (live && k > 0) ? last - r : 0;
        for (int o = kWave / 2; o > 0; o >>= 1) {
            below = max(below, __shfl_xor(below, o));
            above = max(above, __shfl_xor(above, o));
        }
        if ((threadIdx.x & (kWave - 1)) == 0) {
            atomicMax(&a.creach[0], below);
            atomicMax(&a.creach[1], above);
        }
    }
    if (!live) return;
    a.adiag[r] = diag + a.dshift;
    a.rowlen[r] = k;
    a.rhs[r] = a.source[i];
    if (a.dcol) {
        a.cbase[r] = first;
        if (bad) atomicOr(a.dbad, 1);
    }
}

// FP64 copy of FP32-stored values (a solve in FP64 of a system assembled in
// mixed precision: the FP32 values are exact in FP64).
__global__ void k_aval64(int nu, int cap, const int32_t* __restrict__ rowlen,
                         const float* __restrict__ aval32, double* __restrict__ aval)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nu) return;
    const long ob = mx_index(r, 0, cap);
    const int len = rowlen[r];
    for (int k = 0; k < len; ++k) aval[ob + 64L * k] = (double)aval32[ob + 64L * k];
}

// FP32 copy of the off-diagonal values for the mixed-precision inner solves:
// one thread per row, so a wave's 64 rows read and write entry k of their
// slice as one contiguous block.
__global__ void k_aval32(int nu, int cap, const int32_t* __restrict__ rowlen,
                         const double* __restrict__ aval, float* __restrict__ aval32)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nu) return;
    const long ob = mx_index(r, 0, cap);
    const int len = rowlen[r];
    constexpr int U = 8;
    int k = 0;
    for (; k + U <= len; k += U) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = __builtin_nontemporal_load(&aval[ob + 64L * (k + u)]);
#pragma unroll
        for (int u = 0; u < U; ++u) aval32[ob + 64L * (k + u)] = (float)v[u];
    }
    for (; k < len; ++k) aval32[ob + 64L * k] = (float)aval[ob + 64L * k];
}

// ---- higher-order source with error compensation (srcopt, functions.cu:928-988)
struct SrcArgs {
    int n, dim;
    double re, hs_coef, n0s, dt;
    ListView L;
    const double *x, *y, *z, *us, *vs, *ws, *nstar;
    const int32_t *bcon, *ptype;
    double* hs;
};

// The divergence part of row i (bcon == 0): hs_i = (-1.25 / (n0 dt)) * sum_j
// (re / d^3) ((x_i - x_j) du + (y_i - y_j) dv [+ (z_i - z_j) dw]) over list L,
// du = u*_i - u*_j, skipping dummies, self and coincident particles.
__global__ __launch_bounds__(kBlock) void k_src_hs(SrcArgs a)
{
    const int i = xcd_tid();
    if (i >= a.n) return;
    if (a.ptype[i] == -1 || a.bcon[i] != 0) return;
    const int32_t* __restrict__ row = a.L.ids + sl_index(i, 0, a.L.cap);
    const double xi = a.x[i], yi = a.y[i], zi = (a.dim == 3) ? a.z[i] : 0.0;
    const double ui = a.us[i], vi = a.vs[i], wi = (a.dim == 3) ? a.ws[i] : 0.0;
    double sum = 0.0;
    const int cnt = a.L.cnt[i];
    for (int l = 0; l < cnt; ++l) {
        const int j = row[sl_step(l)];
        if (a.ptype[j] <= -1 || i == j) continue;
        const double dx = a.x[j] - xi, dy = a.y[j] - yi;
        double d;
        if (a.dim == 3) {
            const double dz = a.z[j] - zi;
            d = sqrt(dx * dx + dy * dy + dz * dz);
        } else {
            d = sqrt(dx * dx + dy * dy);
        }
        if (fabs(d) <= 1.0e-15) continue;
        const double du = ui - a.us[j], dv = vi - a.vs[j];
        if (a.dim == 3) {
            const double dw = wi - a.ws[j];
            sum += (a.re / (d * d * d)) * ((xi - a.x[j]) * du + (yi - a.y[j]) * dv + (zi - a.z[j]) * dw);
        } else {
            sum += (a.re / (d * d * d)) * ((xi - a.x[j]) * du + (yi - a.y[j]) * dv);
        }
    }
    a.hs[i] = a.hs_coef * sum;
}

// The error-compensating term reads the source just written for particle
// i - 1 (the reference's sequential loop), so rows are chained in particle
// order. A chain starts where particle i - 1 is not an unknown (its source is
// 0); one thread walks each chain. source[] holds 0 for every non-unknown.
__global__ __launch_bounds__(kBlock) void k_src_chain(int n, const int32_t* __restrict__ bcon,
                                                      const int32_t* __restrict__ ptype,
                                                      const double* __restrict__ hs,
                                                      const double* __restrict__ nstar, double n0s,
                                                      double dt, double* source)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    auto unknown = [&](int k) { return bcon[k] == 0 && ptype[k] != -1; };
    if (!unknown(i) || (i > 0 && unknown(i - 1))) return;
    double prev = (i > 0) ? source[i - 1] : 0.0;
    for (int k = i; k < n && unknown(k); ++k) {
        double ecs = 0.0;
        if (k > 0) {
            const double alpha = (nstar[k] - n0s) / n0s;
            const double beta = (dt / n0s) * (prev / (-1.0 / (n0s * dt)));
            ecs = fabs(alpha) * (beta / dt) + fabs(beta) * (alpha / dt);
        }
        double s = hs[k] + ecs;
        if (fabs(s) < 0.99) s = 0;
        source[k] = s;
        prev = s;
    }
}

__global__ void k_rhs_from_source(int nu, const int32_t* __restrict__ urow, const double* __restrict__ source,
                                  double* rhs)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nu) rhs[r] = source[urow[r]];
}

static void launch_assemble_rows(vox_ctx* c, const AsmArgs& a, bool fp32_only)
{
    const int n = c->tp;
    switch (kt_of(c->p.ktype) * 2 + (fp32_only ? 1 : 0)) {
    case 4: k_assemble_rows<double, 2><<<grid_for(n), kBlock, 0, c->stream>>>(a); break;
    case 5: k_assemble_rows<float, 2><<<grid_for(n), kBlock, 0, c->stream>>>(a); break;
    case 12: k_assemble_rows<double, 6><<<grid_for(n), kBlock, 0, c->stream>>>(a); break;
    case 13: k_assemble_rows<float, 6><<<grid_for(n), kBlock, 0, c->stream>>>(a); break;
    case 1: k_assemble_rows<float, 0><<<grid_for(n), kBlock, 0, c->stream>>>(a); break;
    default: k_assemble_rows<double, 0><<<grid_for(n), kBlock, 0, c->stream>>>(a); break;
    }
}

int ppe_assemble(vox_ctx* c)
{
    VOX_TRY(list_l_for_consumer(c));
    ProfScope ps(c, KC_ASSEMBLE);
    const int n = c->tp;
    const DevList& L = list_of(c, VOX_LIST_L);
    if (c->a_cap < L.cap) {
        if (c->acol) (void)hipFree(c->acol);
        if (c->aval) (void)hipFree(c->aval);
        if (c->dcol) (void)hipFree(c->dcol);
        c->acol = nullptr;
        c->aval = nullptr;
        c->dcol = nullptr;
        VOX_HIP(hipMalloc(&c->acol, sizeof(int32_t) * (size_t)sl_size(n, L.cap)));
        VOX_HIP(hipMalloc(&c->aval, sizeof(double) * (size_t)sl_size(n, L.cap)));
        if (c->aval32) (void)hipFree(c->aval32);
        c->aval32 = nullptr;
        c->aval32_ok = false;
        c->a_cap = L.cap;
    }
    // Mixed precision stores the off-diagonal values in FP32 only (computed in
    // FP64 and rounded once; diagonal and right-hand side stay FP64). The FP64
    // mode stores FP64 values, plus an FP32 copy for the learned
    // preconditioner's network.
    const bool fp32_only = (c->p.precision == VOX_PREC_MIXED);
    const bool mixed = fp32_only || (c->p.solver == VOX_SOLVER_LEARNED_FCG && c->lpc);
    if (mixed && !c->aval32)
        VOX_HIP(hipMalloc(&c->aval32, sizeof(float) * (size_t)sl_size(c->tp, c->a_cap)));
    VOX_TRY(number_unknowns(c));
    const vox_params& p = c->p;
    AsmArgs a;
    a.n = n; a.dim = p.dim; a.ktype = p.ktype; a.re = p.re_l;
    a.c0 = 2.0 * p.dim / p.lambda;
    a.matopt = p.matopt;
    a.hl = (p.dim == 2) ? 3.0 * p.re_l : 2.0 * p.re_l;
    a.n0l = c->n0l; a.rho1 = p.rho1;
    a.inv_re = 1.0 / a.re; a.inv_n0l = 1.0 / a.n0l; a.inv_rho1 = 1.0 / a.rho1;
    a.dshift = (1.00e-7) / p.dt / p.dt;
    a.inv_dt_dt = 1.0 / p.dt / p.dt;
    a.n0s = c->n0s;
    a.L = view(c, VOX_LIST_L);
    a.acap = c->a_cap;
    // the CPU path assembles at x with n (main.cu:1149-1150), the CUDA path at
    // x* with n* (main.cu:1165, functions.cu:1007-1133)
    const bool star = p.ref_gpu != 0;
    a.x = c->f[star ? VOX_F_XSTAR : VOX_F_X]; a.y = c->f[star ? VOX_F_YSTAR : VOX_F_Y];
    a.z = c->f[star ? VOX_F_ZSTAR : VOX_F_Z]; a.pnd = c->f[star ? VOX_F_NSTAR : VOX_F_N];
    a.bcon = c->bcon; a.uidx = c->uidx; a.urow = c->urow; a.nu = c->dflags + FLAG_NU;
    // The 32-bit columns are stored where the 16-bit steps may not fit (the
    // unknowns in particle order) or the learned preconditioner reads them;
    // otherwise the steps stand in for them (ensure_acol).
    const bool write_acol = !c->uorder_cells || c->p.solver == VOX_SOLVER_LEARNED_FCG;
    a.acol = write_acol ? c->acol : nullptr; a.rowlen = c->rowlen; a.aval = c->aval;
    if (!c->dcol) VOX_HIP(hipMalloc(&c->dcol, sizeof(uint16_t) * (size_t)sl_size(n, c->a_cap)));
    if (!c->cbase) VOX_HIP(hipMalloc(&c->cbase, sizeof(int32_t) * (size_t)n));
    a.dcol = reinterpret_cast<v4u*>(c->dcol); a.cbase = c->cbase; a.dbad = c->dflags + FLAG_DCOL_BAD;
    a.creach = c->dflags + FLAG_COL_BELOW;
    VOX_TRY(fill_i32(c, a.dbad, 0, 1));
    VOX_TRY(fill_i32(c, a.creach, 0, 2));
    a.aval32 = mixed ? c->aval32 : nullptr;
    a.adiag = c->adiag; a.rhs = c->rhs;
    a.source = c->f[VOX_F_SOURCE];
    a.rec = c->pos_rec;
    // rows of positions (list L's compact rows) index list L's sorted records,
    // whose positions are this assembly's (the CPU path's x, unchanged since
    // the search: list_l_for_consumer)
    a.order = L.pos_rows ? L.rec : nullptr;
    a.pos_of = L.pos_rows ? L.pos_of : nullptr;
    k_pack_asm<<<grid_for(n), kBlock, 0, c->stream>>>(a);
    // rows: a grid for every particle, of which the blocks past nu (read on
    // the device) return at once
    launch_assemble_rows(c, a, fp32_only);
    VOX_HIP(hipGetLastError());
    VOX_TRY(read_flags(c));
    c->nu = c->hflags[FLAG_NU];
    c->dcol_ok = !c->hflags[FLAG_DCOL_BAD];
    c->acol_ok = write_acol;
    if (!c->dcol_ok && !write_acol) {
        // a step did not fit: the rows again, with their 32-bit columns (the
        // same values; the kernel is deterministic)
        a.acol = c->acol;
        a.dcol = nullptr;
        launch_assemble_rows(c, a, fp32_only);
        VOX_HIP(hipGetLastError());
        c->acol_ok = true;
    }
    c->col_below = c->hflags[FLAG_COL_BELOW];
    c->col_above = c->hflags[FLAG_COL_ABOVE];
    c->asm_count += 1;
    c->aval32_ok = mixed;
    c->aval64_ok = !fp32_only;
    if (p.srcopt) {
        SrcArgs s;
        s.n = n; s.dim = p.dim; s.re = p.re_l; s.n0s = c->n0s; s.dt = p.dt;
        s.hs_coef = -1.25 / (c->n0l * p.dt);
        s.L = view(c, VOX_LIST_L);
        s.x = a.x; s.y = a.y; s.z = a.z;
        s.us = c->f[VOX_F_USTAR]; s.vs = c->f[VOX_F_VSTAR]; s.ws = c->f[VOX_F_WSTAR];
        s.nstar = c->f[VOX_F_NSTAR]; s.bcon = c->bcon; s.ptype = c->ptype;
        s.hs = c->scratch_d;
        k_src_hs<<<grid_for(n), kBlock, 0, c->stream>>>(s);
        k_src_chain<<<grid_for(n), kBlock, 0, c->stream>>>(n, c->bcon, c->ptype, c->scratch_d,
                                                           c->f[VOX_F_NSTAR], c->n0s, p.dt,
                                                           c->f[VOX_F_SOURCE]);
        if (c->nu > 0)
            k_rhs_from_source<<<grid_for(c->nu), kBlock, 0, c->stream>>>(c->nu, c->urow,
                                                                         c->f[VOX_F_SOURCE], c->rhs);
        VOX_HIP(hipGetLastError());
    }
    return VOX_OK;
}

// PRESSURECALC, method 3 (functions.cu:2471-2500): equation of state on n*
// for every non-dummy particle; dummies keep their pressure.
__global__ void k_pressure_wc(int gp, int n, const int32_t* __restrict__ ptype, const double* __restrict__ C,
                              const double* __restrict__ nstar, double c01, double c02, double rho1,
                              double rho2, double beta, double n0, double pmin, double pmax, double* pnew)
{
    const int i = gp + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double c0 = (ptype[i] == 2) ? c02 : c01;
    const double rho = rho2 * C[i] + rho1 * (1 - C[i]);
    double p;
    if (nstar[i] < beta * n0) p = 0.0;
    else p = (c0 * c0 * rho / 7.0) * (pow(nstar[i] / n0, 7.0) - 1);
    if (p < pmin) p = pmin;
    if (p > pmax) p = pmax;
    pnew[i] = p;
}

int st_pressure_wc(vox_ctx* c)
{
    ProfScope ps(c, KC_MISC);
    const vox_params& p = c->p;
    const int m = c->tp - c->gp;
    if (m <= 0) return VOX_OK;
    k_pressure_wc<<<grid_for(m), kBlock, 0, c->stream>>>(c->gp, c->tp, c->ptype, c->f[VOX_F_C],
                                                         c->f[VOX_F_NSTAR], p.c01, p.c02, p.rho1, p.rho2,
                                                         p.wc_beta, c->n0s, p.pmin, p.pmax,
                                                         c->f[VOX_F_PNEW]);
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

int ppe_make_aval64(vox_ctx* c)
{
    if (c->nu > 0) {
        k_aval64<<<grid_for(c->nu), kBlock, 0, c->stream>>>(c->nu, c->a_cap, c->rowlen, c->aval32,
                                                           c->aval);
        VOX_HIP(hipGetLastError());
    }
    c->aval64_ok = true;
    return VOX_OK;
}

int ppe_make_aval32(vox_ctx* c)
{
    if (!c->aval64_ok) return fail(VOX_ERR_STATE, "no assembled FP64 values to round");
    if (!c->aval32)
        VOX_HIP(hipMalloc(&c->aval32, sizeof(float) * (size_t)sl_size(c->tp, c->a_cap)));
    if (c->nu > 0) {
        k_aval32<<<grid_for(c->nu), kBlock, 0, c->stream>>>(c->nu, c->a_cap, c->rowlen, c->aval,
                                                           c->aval32);
        VOX_HIP(hipGetLastError());
    }
    c->aval32_ok = true;
    return VOX_OK;
}

// The 32-bit columns from the 16-bit steps, one thread per row.
__global__ void k_acol_from_steps(int nu, int cap, const int32_t* __restrict__ rowlen,
                                  const v4u* __restrict__ dcol, const int32_t* __restrict__ cbase,
                                  int32_t* __restrict__ acol)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nu) return;
    const int len = rowlen[r];
    const v4u* __restrict__ dc = dcol + dc_index(r, 0, cap);
    int32_t* __restrict__ col = acol + mx_index(r, 0, cap);
    int cj = cbase[r];
    for (int k = 0; k < len; k += 8) {
        const v4u w = dc[8L * k];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            if (k + u < len) {
                cj += (int)((w[u >> 1] >> (16 * (u & 1))) & 0xffffu);
                col[64L * (k + u)] = cj;
            }
        }
    }
}

int ensure_acol(vox_ctx* c)
{
    if (c->acol_ok || c->nu <= 0) return VOX_OK;
    if (!c->dcol_ok) return fail(VOX_ERR_STATE, "PPE columns unavailable");
    k_acol_from_steps<<<grid_for(c->nu), kBlock, 0, c->stream>>>(
        c->nu, c->a_cap, c->rowlen, reinterpret_cast<const v4u*>(c->dcol), c->cbase, c->acol);
    VOX_HIP(hipGetLastError());
    c->acol_ok = true;
    return VOX_OK;
}

// CSR over the unknowns: the diagonal first, then the stored entries in list
// order.
int ppe_export(vox_ctx* c, int32_t* nu_out, int32_t* nnz_out, int32_t* rowptr, int32_t* cols,
               double* vals, double* rhs, int32_t* ids)
{
    const int nu = c->nu;
    VOX_TRY(ensure_acol(c));
    std::vector<int32_t> len(nu > 0 ? nu : 1);
    if (nu > 0)
        VOX_HIP(hipMemcpy(len.data(), c->rowlen, sizeof(int32_t) * nu, hipMemcpyDeviceToHost));
    long nnz = 0;
    for (int r = 0; r < nu; ++r) nnz += len[r] + 1;
    if (nu_out) *nu_out = nu;
    if (nnz_out) *nnz_out = (int32_t)nnz;
    if (!rowptr && !cols && !vals && !rhs && !ids) return VOX_OK;
    const long tot = sl_size(nu, c->a_cap);
    std::vector<int32_t> acol(nu > 0 ? tot : 1);
    std::vector<double> aval(nu > 0 ? tot : 1);
    std::vector<double> diag(nu > 0 ? nu : 1), b(nu > 0 ? nu : 1);
    std::vector<int32_t> urow(nu > 0 ? nu : 1);
    if (nu > 0) {
        VOX_HIP(hipMemcpy(acol.data(), c->acol, sizeof(int32_t) * tot, hipMemcpyDeviceToHost));
        if (c->aval64_ok) {
            VOX_HIP(hipMemcpy(aval.data(), c->aval, sizeof(double) * tot, hipMemcpyDeviceToHost));
        } else {            // mixed precision: the values are stored in FP32
            std::vector<float> a32(tot);
            VOX_HIP(hipMemcpy(a32.data(), c->aval32, sizeof(float) * tot, hipMemcpyDeviceToHost));
            for (long q = 0; q < tot; ++q) aval[q] = (double)a32[q];
        }
        VOX_HIP(hipMemcpy(diag.data(), c->adiag, sizeof(double) * nu, hipMemcpyDeviceToHost));
        VOX_HIP(hipMemcpy(b.data(), c->rhs, sizeof(double) * nu, hipMemcpyDeviceToHost));
        VOX_HIP(hipMemcpy(urow.data(), c->urow, sizeof(int32_t) * nu, hipMemcpyDeviceToHost));
    }
    // rows and columns in the reference's numbering (unknowns in particle
    // order): row o of the export is the unknown of the o-th smallest particle
    std::vector<int32_t> row_of(nu > 0 ? nu : 1), num(nu > 0 ? nu : 1);
    for (int r = 0; r < nu; ++r) row_of[r] = r;
    if (c->uorder_cells)
        std::sort(row_of.begin(), row_of.begin() + nu,
                  [&](int32_t a, int32_t b2) { return urow[a] < urow[b2]; });
    for (int o = 0; o < nu; ++o) num[row_of[o]] = o;
    long k = 0;
    for (int o = 0; o < nu; ++o) {
        const int r = row_of[o];
        if (rowptr) rowptr[o] = (int32_t)k;
        if (cols) cols[k] = o;
        if (vals) vals[k] = diag[r];
        ++k;
        for (int s = 0; s < len[r]; ++s) {
            const double v = aval[mx_index(r, s, c->a_cap)];
            if (cols) cols[k] = num[acol[mx_index(r, s, c->a_cap)]];
            if (vals) vals[k] = v;
            ++k;
        }
        if (rhs) rhs[o] = b[r];
        if (ids) ids[o] = urow[r];
    }
    if (rowptr) rowptr[nu] = (int32_t)k;
    return VOX_OK;
}

// Load a PPE given as CSR over nu unknowns (the diagonal first in each row,
// as ppe_export writes it) in place of the assembled one: the inverse of
// ppe_export. The rows go to the same SELL-64 storage the assembly fills.
int ppe_load(vox_ctx* c, int32_t nu, int32_t nnz, const int32_t* rowptr, const int32_t* cols,
             const double* vals, const double* rhs, const int32_t* ids)
{
    if (nu < 0 || nu > c->tp) return fail(VOX_ERR_ARG, "vox_ppe_load: nu outside 0..tp");
    if (nu > 0 && (!rowptr || !cols || !vals || !rhs || !ids))
        return fail(VOX_ERR_ARG, "vox_ppe_load: null array");
    if (nu > 0 && (rowptr[0] != 0 || rowptr[nu] != nnz))
        return fail(VOX_ERR_ARG, "vox_ppe_load: rowptr[0] must be 0 and rowptr[nu] == nnz");
    int maxoff = 0;
    std::vector<char> seen(c->tp, 0);
    for (int r = 0; r < nu; ++r) {
        const int a = rowptr[r], b = rowptr[r + 1];
        if (b <= a) return fail(VOX_ERR_ARG, "vox_ppe_load: row " + std::to_string(r) + " is empty");
        if (cols[a] != r)
            return fail(VOX_ERR_ARG, "vox_ppe_load: row " + std::to_string(r) +
                                         " does not start with its diagonal");
        for (int k = a + 1; k < b; ++k)
            if (cols[k] < 0 || cols[k] >= nu)
                return fail(VOX_ERR_ARG, "vox_ppe_load: column out of range in row " +
                                             std::to_string(r));
        maxoff = std::max(maxoff, b - a - 1);
        if (ids[r] < 0 || ids[r] >= c->tp || seen[ids[r]])
            return fail(VOX_ERR_ARG, "vox_ppe_load: particle ids must be distinct and in 0..tp-1");
        seen[ids[r]] = 1;
    }
    const int cap = round_cap(std::max(maxoff, 1));
    if (c->a_cap < cap) {
        if (c->acol) (void)hipFree(c->acol);
        if (c->aval) (void)hipFree(c->aval);
        c->acol = nullptr;
        c->aval = nullptr;
        VOX_HIP(hipMalloc(&c->acol, sizeof(int32_t) * (size_t)sl_size(c->tp, cap)));
        VOX_HIP(hipMalloc(&c->aval, sizeof(double) * (size_t)sl_size(c->tp, cap)));
        if (c->aval32) (void)hipFree(c->aval32);
        if (c->dcol) (void)hipFree(c->dcol);
        c->aval32 = nullptr;
        c->dcol = nullptr;
        c->a_cap = cap;
    }
    const long tot = sl_size(nu, c->a_cap);
    std::vector<int32_t> len(nu > 0 ? nu : 1), acol(nu > 0 ? tot : 1, 0);
    std::vector<double> aval(nu > 0 ? tot : 1, 0.0), diag(nu > 0 ? nu : 1);
    std::vector<int32_t> uidx(c->tp, -1);
    for (int r = 0; r < nu; ++r) {
        const int a = rowptr[r], b = rowptr[r + 1];
        diag[r] = vals[a];
        len[r] = b - a - 1;
        for (int k = a + 1; k < b; ++k) {
            acol[mx_index(r, k - a - 1, c->a_cap)] = cols[k];
            aval[mx_index(r, k - a - 1, c->a_cap)] = vals[k];
        }
        uidx[ids[r]] = r;
    }
    VOX_HIP(hipStreamSynchronize(c->stream));
    if (nu > 0) {
        VOX_HIP(hipMemcpy(c->rowlen, len.data(), sizeof(int32_t) * nu, hipMemcpyHostToDevice));
        VOX_HIP(hipMemcpy(c->acol, acol.data(), sizeof(int32_t) * tot, hipMemcpyHostToDevice));
        VOX_HIP(hipMemcpy(c->aval, aval.data(), sizeof(double) * tot, hipMemcpyHostToDevice));
        VOX_HIP(hipMemcpy(c->adiag, diag.data(), sizeof(double) * nu, hipMemcpyHostToDevice));
        VOX_HIP(hipMemcpy(c->rhs, rhs, sizeof(double) * nu, hipMemcpyHostToDevice));
        VOX_HIP(hipMemcpy(c->urow, ids, sizeof(int32_t) * nu, hipMemcpyHostToDevice));
    }
    VOX_HIP(hipMemcpy(c->uidx, uidx.data(), sizeof(int32_t) * c->tp, hipMemcpyHostToDevice));
    c->nu = nu;
    c->aval32_ok = false;
    c->aval64_ok = true;
    c->dcol_ok = false;            // the product reads the 32-bit columns
    c->acol_ok = true;
    c->asm_count += 1;
    c->uorder_cells = false;       // the rows keep the loaded order
    return VOX_OK;
}

}  // namespace vox
