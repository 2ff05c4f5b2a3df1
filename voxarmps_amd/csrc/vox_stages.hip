// vox_stages.hip — per-particle stages of the MPS time step on gfx950.
//
// One thread per particle; neighbour rows are read 8 entries at a time from the
// sliced row storage (one 32-byte load pair per lane, one contiguous block per
// wave), and each row is summed sequentially in the reference's order so that,
// built with -ffp-contract=off, every per-particle sum is bit-identical to the
// reference's sequential CPU path.
#include "vox_internal.hpp"

namespace vox {

__device__ __forceinline__ double pdist(int dim, const double* __restrict__ x,
                                        const double* __restrict__ y,
                                        const double* __restrict__ z, int i, int j,
                                        double xi, double yi, double zi)
{
    double dx = x[j] - xi, dy = y[j] - yi;
    if (dim == 3) {
        double dz = z[j] - zi;
        return sqrt(dx * dx + dy * dy + dz * dz);
    }
    return sqrt(dx * dx + dy * dy);
}

// ---- density: main.cu:970-977 -----------------------------------------------------------
__global__ void k_density(int n, const int32_t* __restrict__ ptype, double rho1, double rho2,
                          double* __restrict__ rho)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) rho[i] = (ptype[i] == 2) ? rho2 : rho1;
}

int st_density(vox_ctx* c)
{
    ProfScope ps(c, KC_MISC);
    k_density<<<grid_for(c->tp), kBlock, 0, c->stream>>>(c->tp, c->ptype, c->p.rho1, c->p.rho2,
                                                        c->f[VOX_F_RHO]);
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

// ---- PNUM / PNUMSTAR: functions.cu:94-123, 157-180 -------------------------------------------
// One 32-byte record per particle (x, y, z; z = 0 in 2D), so a neighbour costs
// one gather.
__global__ void k_pack_pos(int n, int dim, const double* __restrict__ x,
                           const double* __restrict__ y, const double* __restrict__ z,
                           PosRec* __restrict__ rec)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    PosRec r;
    r.x = x[i]; r.y = y[i]; r.z = (dim == 3) ? z[i] : 0.0;
    r.id = i; r.aux = 0;
    rec[i] = r;
}

// n_i = sum over the row of W(d_ij), j != i, in row order. Branch-free per
// entry with 8 record loads in flight; the 3-term distance equals the 2D one
// bit for bit (dz = 0).
// rec: the records the row entries index (by particle id, or by position in
// the list's cell-sorted records when self_of is given: self_of[i] is then
// particle i's own position there).
__global__ __launch_bounds__(kBlock) void k_pnd(int n, int ktype, double re, double ncorr,
                                                ListView L, const PosRec* __restrict__ rec,
                                                const int32_t* __restrict__ self_of,
                                                double* __restrict__ out,
                                                const double* __restrict__ p,
                                                double* __restrict__ pnew)
{
    // (threads in the list's cell order read the rows, stored by particle id,
    // from scattered slices: measured 20% slower than particle order here,
    // where the gathers are light and the row reads dominate)
    const int i = xcd_tid();
    if (i >= n) return;
    const int self = self_of ? self_of[i] : i;
    const PosRec me = rec[self];
    double sum = 0.0;
    const double inv_re = 1.0 / re;
    const int cnt = L.cnt[self];                 // rows of positions sit at positions
    const int32_t* __restrict__ rowp = L.ids + sl_index(self, 0, L.cap);
    for (int g = 0; (g << 3) < cnt; ++g) {
        int32_t jc[8];
        load_chunk8(rowp, g, jc);
        PosRec rr[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) jc[u] = ((g << 3) + u < cnt) ? jc[u] : self;
#pragma unroll
        for (int u = 0; u < 8; ++u) rr[u] = rec[jc[u]];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const double dx = rr[u].x - me.x, dy = rr[u].y - me.y, dz = rr[u].z - me.z;
            const double w = wfun_r(sqrt(dx * dx + dy * dy + dz * dz), ktype, re, inv_re);
            sum = (jc[u] != self) ? sum + w : sum;
        }
    }
    if (ktype != 2) sum = sum * ncorr;
    out[i] = sum;
    if (pnew) pnew[i] = p[i];
}

// Rows of positions (list L's compact rows) index the list's own sorted
// records, which hold the positions of its search: the caller passes those
// positions (n0 right after NEIGHBOR).
static void launch_pnd(vox_ctx* c, const DevList& L, double re, double ncorr, const double* x,
                       const double* y, const double* z, double* out, const double* p,
                       double* pnew)
{
    const PosRec* rec = L.rec;
    const int32_t* self_of = L.pos_of;
    if (!L.pos_rows) {
        k_pack_pos<<<grid_for(c->tp), kBlock, 0, c->stream>>>(c->tp, c->p.dim, x, y, z, c->pos_rec);
        rec = c->pos_rec;
        self_of = nullptr;
    }
    // (a kernel specialised on the kernel type holds fewer registers, runs
    // more waves and measured 11% slower: the gathers then share the caches
    // among more rows)
    k_pnd<<<grid_for(c->tp), kBlock, 0, c->stream>>>(c->tp, c->p.ktype, re, ncorr,
                                                      ListView{L.cnt, L.ids, L.cap}, rec, self_of,
                                                      out, p, pnew);
}

int st_pnd(vox_ctx* c, int on_star, int loopstarted)
{
    ProfScope ps(c, KC_PND);
    const double* x = c->f[on_star ? VOX_F_XSTAR : VOX_F_X];
    const double* y = c->f[on_star ? VOX_F_YSTAR : VOX_F_Y];
    const double* z = c->f[on_star ? VOX_F_ZSTAR : VOX_F_Z];
    double* out = c->f[on_star ? VOX_F_NSTAR : VOX_F_N];
    // PNUMSTAR applies no Ncorrection: multiplying by 1.0 is exact
    double ncorr = on_star ? 1.0 : c->p.ncorrection;
    launch_pnd(c, list_of(c, VOX_LIST_S), c->p.re_s, ncorr, x, y, z, out, c->f[VOX_F_P],
               (!on_star && !loopstarted) ? c->f[VOX_F_PNEW] : nullptr);
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

// n0 = max n over list S (main.cu:1011) and the same maximum of the PND on list L.
int st_n0(vox_ctx* c)
{
    std::vector<double> h(c->tp);
    VOX_HIP(hipMemcpyAsync(h.data(), c->f[VOX_F_N], sizeof(double) * c->tp,
                           hipMemcpyDeviceToHost, c->stream));
    VOX_HIP(hipStreamSynchronize(c->stream));
    double mx = 0.0;
    for (double v : h) if (v > mx) mx = v;
    c->n0s = mx;
    if (c->two_lists) {
        VOX_TRY(list_l_for_consumer(c));
        launch_pnd(c, list_of(c, VOX_LIST_L), c->p.re_l, c->p.ncorrection, c->f[VOX_F_X],
                   c->f[VOX_F_Y], c->f[VOX_F_Z], c->scratch_d, nullptr, nullptr);
        VOX_HIP(hipGetLastError());
        VOX_HIP(hipMemcpyAsync(h.data(), c->scratch_d, sizeof(double) * c->tp,
                               hipMemcpyDeviceToHost, c->stream));
        VOX_HIP(hipStreamSynchronize(c->stream));
        double ml = 0.0;
        for (double v : h) if (v > ml) ml = v;
        c->n0l = ml;
    } else {
        c->n0l = mx;
    }
    return VOX_OK;
}

// ---- strain rate: SPS functions.cu:1738-1790, VISCOSITY 2095-2150 ---------------------------------
struct Strain { double s11, s12, s22, s13, s23, s33; };

__device__ __forceinline__ Strain strain_row(int i, int dim, int ktype, double re, double n0,
                                             double dmin, bool use_dmin, ListView L,
                                             const double* __restrict__ x,
                                             const double* __restrict__ y,
                                             const double* __restrict__ z,
                                             const double* __restrict__ u,
                                             const double* __restrict__ v,
                                             const double* __restrict__ w)
{
    double xi = x[i], yi = y[i], zi = (dim == 3) ? z[i] : 0.0;
    double ui = u[i], vi = v[i], wi = (dim == 3) ? w[i] : 0.0;
    double s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0, s6 = 0, s7 = 0, s8 = 0, s9 = 0;
    int cnt = L.cnt[i];
    const int32_t* __restrict__ rowp = L.ids + sl_index(i, 0, L.cap);
    for (int g = 0; (g << 3) < cnt; ++g) {
        int32_t jc[8];
        load_chunk8(rowp, g, jc);
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            if ((g << 3) + e >= cnt) break;
            int j = jc[e];
            if (j == i) continue;
            double dx = x[j] - xi, dy = y[j] - yi, dz = 0.0, d;
            if (dim == 3) { dz = z[j] - zi; d = sqrt(dx * dx + dy * dy + dz * dz); }
            else d = sqrt(dx * dx + dy * dy);
            if (use_dmin && !(d > dmin)) continue;
            double wv = wfun(d, ktype, re);
            double du = u[j] - ui, dv = v[j] - vi;
            s1 = s1 + du * dx * wv / d / d;
            s2 = s2 + dv * dy * wv / d / d;
            s3 = s3 + du * dy * wv / d / d;
            s4 = s4 + dv * dx * wv / d / d;
            if (dim == 3) {
                double dw = w[j] - wi;
                s5 = s5 + dw * dz * wv / d / d;
                s6 = s6 + du * dz * wv / d / d;
                s7 = s7 + dv * dz * wv / d / d;
                s8 = s8 + dw * dx * wv / d / d;
                s9 = s9 + dw * dy * wv / d / d;
            }
        }
    }
    double Uxx = (2.0 / n0) * s1, Uyy = (2.0 / n0) * s2, Uxy = (2.0 / n0) * s3, Uyx = (2.0 / n0) * s4;
    Strain S;
    S.s12 = 0.5 * (Uxy + Uyx);
    S.s11 = 0.5 * (Uxx + Uxx);
    S.s22 = 0.5 * (Uyy + Uyy);
    S.s13 = S.s23 = S.s33 = 0.0;
    if (dim == 3) {
        double Uzz = (2.0 / n0) * s5, Uxz = (2.0 / n0) * s6, Uyz = (2.0 / n0) * s7;
        double Uzx = (2.0 / n0) * s8, Uzy = (2.0 / n0) * s9;
        S.s33 = 0.5 * (Uzz + Uzz);
        S.s13 = 0.5 * (Uxz + Uzx);
        S.s23 = 0.5 * (Uyz + Uzy);
    }
    return S;
}

// ---- SPS: functions.cu:1719-1841 -------------------------------------------------------------
struct SpsArgs {
    int n, first, dim, ktype;
    double re, n0, dmin, csdl2;
    ListView L;
    const double *x, *y, *z, *u, *v, *w;
    double *s11, *s12, *s22, *s13, *s23, *s33, *neut, *t1, *t2, *t3;
};

__global__ __launch_bounds__(kBlock) void k_sps1(SpsArgs a)
{
    int i = a.first + xcd_tid();
    if (i >= a.n) return;
    Strain S = strain_row(i, a.dim, a.ktype, a.re, a.n0, a.dmin, true, a.L, a.x, a.y, a.z, a.u,
                          a.v, a.w);
    a.s11[i] = S.s11; a.s12[i] = S.s12; a.s22[i] = S.s22;
    a.s13[i] = S.s13; a.s23[i] = S.s23; a.s33[i] = S.s33;
    double s = sqrt(2 * S.s11 * S.s11) + 2 * sqrt(2 * S.s12 * S.s12) + sqrt(2 * S.s22 * S.s22);
    if (a.dim == 3)
        s = s + sqrt(2 * S.s33 * S.s33) + 2 * sqrt(2 * S.s13 * S.s13) + 2 * sqrt(2 * S.s23 * S.s23);
    a.neut[i] = a.csdl2 * s;
}

__global__ __launch_bounds__(kBlock) void k_sps2(SpsArgs a)
{
    int i = a.first + xcd_tid();
    if (i >= a.n) return;
    double xi = a.x[i], yi = a.y[i], zi = (a.dim == 3) ? a.z[i] : 0.0, ni = a.neut[i];
    double s1 = 0, s2 = 0, s3 = 0;
    int cnt = a.L.cnt[i];
    const int32_t* __restrict__ rowp = a.L.ids + sl_index(i, 0, a.L.cap);
    for (int g = 0; (g << 3) < cnt; ++g) {
        int32_t jc[8];
        load_chunk8(rowp, g, jc);
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            if ((g << 3) + e >= cnt) break;
            int j = jc[e];
            if (j == i) continue;
            double dx = a.x[j] - xi, dy = a.y[j] - yi, dz = 0.0, d;
            if (a.dim == 3) { dz = a.z[j] - zi; d = sqrt(dx * dx + dy * dy + dz * dz); }
            else d = sqrt(dx * dx + dy * dy);
            if (!(d > a.dmin)) continue;
            double wv = wfun(d, a.ktype, a.re);
            double dn = a.neut[j] - ni;
            s1 = s1 + dn * dx * wv / d / d;
            s2 = s2 + dn * dy * wv / d / d;
            if (a.dim == 3) s3 = s3 + dn * dz * wv / d / d;
        }
    }
    double s11 = a.s11[i], s12 = a.s12[i], s22 = a.s22[i];
    if (a.dim == 3) {
        double s13 = a.s13[i], s23 = a.s23[i], s33 = a.s33[i];
        a.t1[i] = (2.0 / a.n0) * (2 * s11 * s1 + 2 * s12 * s2 + 2 * s13 * s3);
        a.t2[i] = (2.0 / a.n0) * (2 * s12 * s1 + 2 * s22 * s2 + 2 * s23 * s3);
        a.t3[i] = (2.0 / a.n0) * (2 * s13 * s1 + 2 * s23 * s2 + 2 * s33 * s3);
    } else {
        a.t1[i] = (2.0 / a.n0) * (2 * s11 * s1 + 2 * s12 * s2);
        a.t2[i] = (2.0 / a.n0) * (2 * s12 * s1 + 2 * s22 * s2);
    }
}

int st_sps(vox_ctx* c)
{
    if (c->p.precision == VOX_PREC_MIXED) return st_sps32(c);
    ProfScope ps(c, KC_SPS);
    SpsArgs a;
    a.n = c->tp; a.first = c->gp; a.dim = c->p.dim; a.ktype = c->p.ktype;
    a.re = c->p.re_s; a.n0 = c->n0s; a.dmin = c->p.coll / 10;
    a.csdl2 = (c->p.cs * c->p.dl) * (c->p.cs * c->p.dl);
    a.L = view(c, VOX_LIST_S);
    a.x = c->f[VOX_F_X]; a.y = c->f[VOX_F_Y]; a.z = c->f[VOX_F_Z];
    a.u = c->f[VOX_F_UNEW]; a.v = c->f[VOX_F_VNEW]; a.w = c->f[VOX_F_WNEW];
    a.s11 = c->sps_s[0]; a.s12 = c->sps_s[1]; a.s22 = c->sps_s[2];
    a.s13 = c->sps_s[3]; a.s23 = c->sps_s[4]; a.s33 = c->sps_s[5];
    a.neut = c->f[VOX_F_NEUT];
    a.t1 = c->f[VOX_F_TURB1]; a.t2 = c->f[VOX_F_TURB2]; a.t3 = c->f[VOX_F_TURB3];
    int m = c->tp - c->gp;
    if (m <= 0) return VOX_OK;
    k_sps1<<<grid_for(m), kBlock, 0, c->stream>>>(a);
    k_sps2<<<grid_for(m), kBlock, 0, c->stream>>>(a);
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

// ---- V_FRACTION, method 2: functions.cu:2604-2634 ---------------------------------------------
__global__ __launch_bounds__(kBlock) void k_vfraction(int n, int dim, int ktype, double re,
                                                      ListView L, const int32_t* __restrict__ pt,
                                                      const double* __restrict__ x,
                                                      const double* __restrict__ y,
                                                      const double* __restrict__ z,
                                                      double* __restrict__ C)
{
    int i = xcd_tid();
    if (i >= n) return;
    double xi = x[i], yi = y[i], zi = (dim == 3) ? z[i] : 0.0;
    double s1 = 0, s2 = 0;
    int cnt = L.cnt[i];
    const int32_t* __restrict__ rowp = L.ids + sl_index(i, 0, L.cap);
    for (int g = 0; (g << 3) < cnt; ++g) {
        int32_t jc[8];
        load_chunk8(rowp, g, jc);
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            if ((g << 3) + e >= cnt) break;
            int j = jc[e];
            if (j == i) continue;
            int tj = pt[j];
            if (tj <= 0) continue;
            double wv = wfun(pdist(dim, x, y, z, i, j, xi, yi, zi), ktype, re);
            s1 = s1 + wv;
            if (tj == 2) s2 = s2 + wv;
        }
    }
    double cv = s2 / s1;
    if (s1 == 0) cv = 0;
    C[i] = cv;
}

int st_vfraction(vox_ctx* c)
{
    ProfScope ps(c, KC_VFRAC);
    // Without fluid-2 particles every sum of fluid-2 weights is +0, so each C
    // is 0 / s1 = +0 (or the s1 == 0 case, also +0): the same bits as the walk.
    if (c->n_fluid2 == 0) return fill_f64(c, c->f[VOX_F_C], 0.0, c->tp);
    if (c->p.precision == VOX_PREC_MIXED) return st_vfraction32(c);
    k_vfraction<<<grid_for(c->tp), kBlock, 0, c->stream>>>(
        c->tp, c->p.dim, c->p.ktype, c->p.re_s, view(c, VOX_LIST_S), c->ptype, c->f[VOX_F_X],
        c->f[VOX_F_Y], c->f[VOX_F_Z], c->f[VOX_F_C]);
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

// ---- VISCOSITY: functions.cu:2053-2184 ----------------------------------------------------------
__global__ void k_visc_newton(int n, const int32_t* __restrict__ pt, double m1, double m2,
                              double* __restrict__ meu)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int t = pt[i];
    if (t <= 1) meu[i] = m1;
    if (t == 2) meu[i] = m2;
}

struct HbArgs {
    int n, dim, ktype;
    double re, n0, m1, cohes, phi_param, meu0, npl;
    ListView L;
    const int32_t* pt;
    const double *x, *y, *z, *u, *v, *w, *phat;
    double* meu;
};

__global__ __launch_bounds__(kBlock) void k_visc_hb(HbArgs a)
{
    int i = xcd_tid();
    if (i >= a.n) return;
    if (a.pt[i] == 1) { a.meu[i] = a.m1; return; }
    Strain S = strain_row(i, a.dim, a.ktype, a.re, a.n0, 0.0, false, a.L, a.x, a.y, a.z, a.u, a.v,
                          a.w);
    // phi reads C[I] with the caller's I == 0 (main.cu:211): always phi = 0 (2154-2155)
    double cz = 0.0;
    double phi = (cz - 0.25) * a.phi_param / (1 - 0.25);
    if (cz < 0.25) phi = 0;
    double ys = a.cohes * cos(phi) + a.phat[i] * sin(phi);
    double II;
    if (a.dim == 3)
        II = fabs((S.s11 * S.s22 * S.s33 + S.s12 * S.s23 * S.s13 + S.s13 * S.s12 * S.s23) -
                  (S.s13 * S.s22 * S.s13 + S.s11 * S.s23 * S.s23 + S.s12 * S.s12 * S.s33));
    else
        II = fabs(S.s11 * S.s22 - S.s12 * S.s12);
    double m = ys / 2 / sqrt(II) + a.meu0 * pow(4 * II, (a.npl - 1) / 2);
    if (II == 0 || m > 200) m = 200;
    a.meu[i] = m;
}

int st_viscosity(vox_ctx* c)
{
    ProfScope ps(c, KC_VISC);
    const vox_params& p = c->p;
    c->meu_uniform = (p.fluid2_type == 0 && c->n_fluid2 == 0);
    if (p.fluid2_type == 0) {
        k_visc_newton<<<grid_for(c->tp), kBlock, 0, c->stream>>>(c->tp, c->ptype, p.nu1 * p.rho1,
                                                                 p.nu2 * p.rho2, c->f[VOX_F_MEU]);
    } else if (p.precision == VOX_PREC_MIXED) {
        VOX_TRY(st_visc_hb32(c));
    } else {
        HbArgs a;
        a.n = c->tp; a.dim = p.dim; a.ktype = p.ktype; a.re = p.re_s; a.n0 = c->n0s;
        a.m1 = p.nu1 * p.rho1; a.cohes = p.cohes; a.phi_param = p.phi; a.meu0 = p.meu0;
        a.npl = p.n_pl; a.L = view(c, VOX_LIST_S); a.pt = c->ptype;
        a.x = c->f[VOX_F_X]; a.y = c->f[VOX_F_Y]; a.z = c->f[VOX_F_Z];
        a.u = c->f[VOX_F_UNEW]; a.v = c->f[VOX_F_VNEW]; a.w = c->f[VOX_F_WNEW];
        a.phat = c->f[VOX_F_PHAT]; a.meu = c->f[VOX_F_MEU];
        k_visc_hb<<<grid_for(c->tp), kBlock, 0, c->stream>>>(a);
    }
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

// ---- PREDICTION: functions.cu:2283-2369 ------------------------------------------------------------
struct PredArgs {
    int n, dim, ktype;
    const int32_t* pos_of;    // pos: each particle's position in order
    double re_s, re_l, rho1, rho2, gx, gy, gz, dt, lapc, ln0, gcoef_num, n0s, relax1;
    ListView L;
    const int32_t* pt;
    const double *x, *y, *z, *u, *v, *w, *pn, *ph, *meu, *neut, *rho, *t1, *t2, *t3;
    double *xs, *ys, *zs, *us, *vs, *ws;
    PredRec* rec;             // the records the row entries index (see pos)
    const PosRec* order;      // thread t takes particle order[t].id (cell-sorted)
    int32_t* rpt;             // pos: the types in the same order as rec
    bool pos;                 // row entries are positions in order (list L's compact
                              // rows): rec and rpt are in that order, and thread t's
                              // own record is rec[t]; else both are by particle id
};

// One 64-byte record per particle with every field the predictor gathers, at
// the particle's id, or at its position in the cell-sorted order (with its
// type beside it) when the rows hold positions. Threads run in particle
// order (the fields are read in lines) and store whole records.
__global__ void k_pack_pred(PredArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int t = a.pos ? a.pos_of[i] : i;
    if (a.pos) a.rpt[t] = a.pt[i];
    PredRec r;
    r.x = a.x[i]; r.y = a.y[i]; r.z = (a.dim == 3) ? a.z[i] : 0.0;
    r.u = a.u[i]; r.v = a.v[i]; r.w = (a.dim == 3) ? a.w[i] : 0.0;
    r.pn = a.pn[i]; r.meu = a.meu[i];
    a.rec[t] = r;
}

// kUniform: every particle has the same MEU m (Newtonian, no fluid 2) and
// 2 m m / (m + m) rounds back to m, so the harmonic-mean branch of
// functions.cu:2328-2337 gives the same NEU as the same-type branch: the
// neighbour's type is not read (one gathered line less per neighbour).
// R: records gathered together per group (8 ids are loaded per row chunk and
// consumed in groups of R; the terms are added in row order either way, so
// every R gives the same bits), BS: threads per block. The autotuner picks
// (R, BS) by timing.
template <bool kUniform, int R, int BS>
__global__ __launch_bounds__(BS) void k_prediction(PredArgs a)
{
    const int t = xcd_tid();
    if (t >= a.n) return;
    const int i = sorted_particle(a.order, t);
    const int self = a.pos ? t : i;          // own index among the row entries' records
    const int32_t* __restrict__ rpt = a.pos ? a.rpt : a.pt;
    const int dim = a.dim;
    const int ti = a.pt[i];
    if (ti <= 0) {
        a.xs[i] = a.x[i]; a.ys[i] = a.y[i]; a.us[i] = a.u[i]; a.vs[i] = a.v[i];
        if (dim == 3) { a.zs[i] = a.z[i]; a.ws[i] = a.w[i]; }
        return;
    }
    const PredRec me = a.rec[self];
    const double xi = me.x, yi = me.y, zi = me.z, ui = me.u, vi = me.v, wi = me.w;
    const double mi = me.meu, phi_ = a.ph[i];
    double rhoN = (ti == 1) ? a.rho1 : a.rho2;
    const double inv_rhoN = 1.0 / rhoN;
    const double inv_re_s = 1.0 / a.re_s, inv_re_l = 1.0 / a.re_l;
    double s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0, s6 = 0, s7 = 0, s8 = 0, s9 = 0;
    const int cnt = a.L.cnt[self];               // rows of positions sit at positions
    const int32_t* __restrict__ row = a.L.ids + sl_index(self, 0, a.L.cap);
    // One row chunk (8 ids) per outer step, consumed in groups of R: a group's
    // record and type loads are issued together, then its terms are
    // accumulated in row order.
    for (int g = 0; (g << 3) < cnt; ++g) {
    int32_t jc[8];
    load_chunk8(row, g, jc);
#pragma unroll
    for (int h = 0; h < 8 / R; ++h) {
    constexpr int U = R;
    int jj[U], tj[U];
    PredRec rv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) jj[u] = ((g << 3) + U * h + u < cnt) ? jc[U * h + u] : self;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        rv[u] = a.rec[jj[u]];
        tj[u] = kUniform ? ti : rpt[jj[u]];
    }
    // Branch-free per neighbour: every term is computed and a skipped entry
    // (self, padding, coincident particle) keeps each sum through a select, so
    // no gathered field is sunk into a branch behind a full memory wait.
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const PredRec rj = rv[u];
        // z, w are 0 in 2D records: the 3-term distance equals the 2D one
        // bit for bit, and no load is left inside a dimension branch
        const double dx = rj.x - xi, dy = rj.y - yi, dz = rj.z - zi;
        const double D = sqrt(dx * dx + dy * dy + dz * dz);
        const bool use = (jj[u] != self) && (D > 0);
        const double mj = rj.meu;
        const double NEU = (kUniform || ti == tj[u])
                               ? div_by(mi, rhoN, inv_rhoN)
                               : div_by(2 * mi * mj / (mi + mj), rhoN, inv_rhoN);
        const double dp = rj.pn - phi_;
        const double du = rj.u - ui, dv = rj.v - vi, dw = rj.w - wi;
        // The gradient terms carry W(D; re_s), which is 0 for D / re_s > 1
        // (every kernel type): there each term is +-0 for finite fields, and
        // adding it leaves the sum's bits unchanged (a sum that starts at +0
        // is never -0), so it is skipped. On list L (radius re_l = 4 l0 against re_s = 2.1 l0)
        // that is most entries. Every quotient equals the plain IEEE division
        // (div_by); the divisions by D share one reciprocal.
        if (use && div_by(D, a.re_s, inv_re_s) <= 1.0) {
            const double inv_D = 1.0 / D;
            const double wg = wfun_r(D, a.ktype, a.re_s, inv_re_s);
            s1 = s1 + div_by(div_by(dp * dx * wg, D, inv_D), D, inv_D);
            s2 = s2 + div_by(div_by(dp * dy * wg, D, inv_D), D, inv_D);
            s3 = s3 + div_by(div_by(dp * dz * wg, D, inv_D), D, inv_D);   // z sum: only read in 3D
        }
        const double wl = wfun_r(D, a.ktype, a.re_l, inv_re_l);
        s4 = use ? s4 + wl * du * NEU : s4;
        s5 = use ? s5 + wl * dv * NEU : s5;
        s7 = use ? s7 + wl * du : s7;
        s8 = use ? s8 + wl * dv : s8;
        s6 = use ? s6 + wl * dw * NEU : s6;
        s9 = use ? s9 + wl * dw : s9;
    }
    }
    }
    double nt = a.neut[i];
    double gc = a.relax1 * (a.gcoef_num * a.dt / a.n0s / a.rho[i]);
    double dt = a.dt;
    double us = ui + a.gx * dt + a.lapc * dt * (s4 + nt * s7) / a.ln0 + dt * a.t1[i] - gc * s1;
    double vs = vi + a.gy * dt + a.lapc * dt * (s5 + nt * s8) / a.ln0 + dt * a.t2[i] - gc * s2;
    a.us[i] = us;
    a.vs[i] = vs;
    a.xs[i] = xi + dt * us;
    a.ys[i] = yi + dt * vs;
    if (dim == 3) {
        double ws = wi + a.gz * dt + a.lapc * dt * (s6 + nt * s9) / a.ln0 + dt * a.t3[i] - gc * s3;
        a.ws[i] = ws;
        a.zs[i] = zi + dt * ws;
    }
}

// The same predictor with the gradient terms deferred: the row walk adds the
// Laplacian terms of every entry and queues (in LDS, in row order) the
// entries within re_s, whose gradient terms are then added in a second pass
// over the queue, the distance recomputed by the same expression. Each of the
// two groups of sums keeps its terms in row order, so the results are the
// same bits. In the single walk a wave runs the gradient block whenever any
// of its 64 lanes has an entry within re_s at that row index (most of them),
// here once per queued entry of its busiest lane (about 1 in 7 entries).
constexpr int kGradQ = 48;   // queued entries per lane before the queue is drained

template <int BS>
__device__ __forceinline__ void pred_grad_drain(const PredArgs& a, const int32_t* gq, int nq,
                                                double xi, double yi, double zi, double phi_,
                                                double inv_re_s, double& s1, double& s2,
                                                double& s3)
{
    constexpr int G = 8;
    for (int q0 = 0; __any(q0 < nq); q0 += G) {
        int jj[G];
        PredRec rv[G];
#pragma unroll
        for (int u = 0; u < G; ++u) jj[u] = (q0 + u < nq) ? gq[(q0 + u) * BS] : -1;
#pragma unroll
        for (int u = 0; u < G; ++u)
            if (jj[u] >= 0) rv[u] = a.rec[jj[u]];
#pragma unroll
        for (int u = 0; u < G; ++u) {
            if (jj[u] >= 0) {
                const PredRec rj = rv[u];
                const double dx = rj.x - xi, dy = rj.y - yi, dz = rj.z - zi;
                const double D = sqrt(dx * dx + dy * dy + dz * dz);
                const double dp = rj.pn - phi_;
                const double inv_D = 1.0 / D;
                const double wg = wfun_r(D, a.ktype, a.re_s, inv_re_s);
                s1 = s1 + div_by(div_by(dp * dx * wg, D, inv_D), D, inv_D);
                s2 = s2 + div_by(div_by(dp * dy * wg, D, inv_D), D, inv_D);
                s3 = s3 + div_by(div_by(dp * dz * wg, D, inv_D), D, inv_D);
            }
        }
    }
}

// (specialised on the kernel type, 8 records in flight fit in 3 waves per
// SIMD, 166 VGPRs against 188, and measured no faster)
template <bool kUniform, int R, int BS>
__global__ __launch_bounds__(BS) void k_prediction_dg(PredArgs a)
{
    __shared__ int32_t gq_all[kGradQ * BS];                 // slot-major: conflict-free
    int32_t* gq = gq_all + threadIdx.x;                     // slot q at gq[q * BS]
    const int t = xcd_tid();
    if (t >= a.n) return;
    const int i = sorted_particle(a.order, t);
    const int self = a.pos ? t : i;
    const int32_t* __restrict__ rpt = a.pos ? a.rpt : a.pt;
    const int dim = a.dim;
    const int ti = a.pt[i];
    if (ti <= 0) {
        a.xs[i] = a.x[i]; a.ys[i] = a.y[i]; a.us[i] = a.u[i]; a.vs[i] = a.v[i];
        if (dim == 3) { a.zs[i] = a.z[i]; a.ws[i] = a.w[i]; }
        return;
    }
    const PredRec me = a.rec[self];
    const double xi = me.x, yi = me.y, zi = me.z, ui = me.u, vi = me.v, wi = me.w;
    const double mi = me.meu, phi_ = a.ph[i];
    double rhoN = (ti == 1) ? a.rho1 : a.rho2;
    const double inv_rhoN = 1.0 / rhoN;
    const double inv_re_s = 1.0 / a.re_s, inv_re_l = 1.0 / a.re_l;
    double s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0, s6 = 0, s7 = 0, s8 = 0, s9 = 0;
    int nq = 0;
    const int cnt = a.L.cnt[self];
    const int32_t* __restrict__ row = a.L.ids + sl_index(self, 0, a.L.cap);
    // one drain site (the loop runs to the wave's longest row; a lane past
    // its own row only drains)
    for (int g = 0;; ++g) {
    const bool more = (g << 3) < cnt;
    if (more) {
    int32_t jc[8];
    load_chunk8(row, g, jc);
#pragma unroll
    for (int h = 0; h < 8 / R; ++h) {
    constexpr int U = R;
    int jj[U], tj[U];
    PredRec rv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) jj[u] = ((g << 3) + U * h + u < cnt) ? jc[U * h + u] : self;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        rv[u] = a.rec[jj[u]];
        tj[u] = kUniform ? ti : rpt[jj[u]];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const PredRec rj = rv[u];
        const double dx = rj.x - xi, dy = rj.y - yi, dz = rj.z - zi;
        const double D = sqrt(dx * dx + dy * dy + dz * dz);
        const bool use = (jj[u] != self) && (D > 0);
        const double mj = rj.meu;
        const double NEU = (kUniform || ti == tj[u])
                               ? div_by(mi, rhoN, inv_rhoN)
                               : div_by(2 * mi * mj / (mi + mj), rhoN, inv_rhoN);
        const double du = rj.u - ui, dv = rj.v - vi, dw = rj.w - wi;
        // W(D; re_s) is 0 beyond re_s (every kernel type), where the
        // gradient terms are +-0 and leave the sums' bits unchanged
        if (use && div_by(D, a.re_s, inv_re_s) <= 1.0) {
            gq[nq * BS] = jj[u];
            ++nq;
        }
        // a skipped entry adds terms with W = +0: +-0 for finite fields,
        // which leave each sum's bits unchanged (one select instead of six)
        const double wl = use ? wfun_r(D, a.ktype, a.re_l, inv_re_l) : 0.0;
        s4 = s4 + wl * du * NEU;
        s5 = s5 + wl * dv * NEU;
        s7 = s7 + wl * du;
        s8 = s8 + wl * dv;
        s6 = s6 + wl * dw * NEU;
        s9 = s9 + wl * dw;
    }
    }
    }
    const bool last = !__any(more);
    if (last || __any(nq > kGradQ - 8)) {                   // room for the next chunk
        pred_grad_drain<BS>(a, gq, nq, xi, yi, zi, phi_, inv_re_s, s1, s2, s3);
        nq = 0;
    }
    if (last) break;
    }
    double nt = a.neut[i];
    double gc = a.relax1 * (a.gcoef_num * a.dt / a.n0s / a.rho[i]);
    double dt = a.dt;
    double us = ui + a.gx * dt + a.lapc * dt * (s4 + nt * s7) / a.ln0 + dt * a.t1[i] - gc * s1;
    double vs = vi + a.gy * dt + a.lapc * dt * (s5 + nt * s8) / a.ln0 + dt * a.t2[i] - gc * s2;
    a.us[i] = us;
    a.vs[i] = vs;
    a.xs[i] = xi + dt * us;
    a.ys[i] = yi + dt * vs;
    if (dim == 3) {
        double ws = wi + a.gz * dt + a.lapc * dt * (s6 + nt * s9) / a.ln0 + dt * a.t3[i] - gc * s3;
        a.ws[i] = ws;
        a.zs[i] = zi + dt * ws;
    }
}

// Every MEU equals nu1 rho1 (Newtonian viscosity and no PTYPE 2 particle:
// functions.cu:2073-2085) and the harmonic mean 2 m m / (m + m) of that value
// rounds back to it, so NEU does not depend on the neighbour's type.
bool uniform_meu(const vox_ctx* c)
{
    if (!c->meu_uniform || c->p.fluid2_type != 0 || c->n_fluid2 != 0) return false;
    const volatile double m = c->p.nu1 * c->p.rho1;
    const volatile double hm = 2 * m * m / (m + m);
    return hm == m;
}

// Predictor variants for the autotuner: (records per group, block size), the
// gradient terms in the row walk (0-4) or deferred to a queue (5-8).
static const int kPredR[kPredVariants] = {4, 8, 2, 4, 8, 8, 8, 4, 4};
static const int kPredBS[kPredVariants] = {256, 256, 256, 128, 128, 128, 256, 128, 256};

template <bool kU>
static void launch_pred_u(const PredArgs& a, int v, int n, hipStream_t s)
{
    switch (v) {
    case 5: k_prediction_dg<kU, 8, 128><<<grid_for(n, 128), 128, 0, s>>>(a); break;
    case 6: k_prediction_dg<kU, 8, 256><<<grid_for(n, 256), 256, 0, s>>>(a); break;
    case 7: k_prediction_dg<kU, 4, 128><<<grid_for(n, 128), 128, 0, s>>>(a); break;
    case 8: k_prediction_dg<kU, 4, 256><<<grid_for(n, 256), 256, 0, s>>>(a); break;
    case 1: k_prediction<kU, 8, 256><<<grid_for(n, 256), 256, 0, s>>>(a); break;
    case 2: k_prediction<kU, 2, 256><<<grid_for(n, 256), 256, 0, s>>>(a); break;
    case 3: k_prediction<kU, 4, 128><<<grid_for(n, 128), 128, 0, s>>>(a); break;
    case 4: k_prediction<kU, 8, 128><<<grid_for(n, 128), 128, 0, s>>>(a); break;
    default: k_prediction<kU, 4, 256><<<grid_for(n, 256), 256, 0, s>>>(a); break;
    }
}

// (variants specialised on the kernel type hold about 20 registers fewer and
// measured the same or slower: the gathers, not the arithmetic, set the pace)
static int launch_prediction(vox_ctx* c, const PredArgs& a, int variant)
{
    if (uniform_meu(c)) launch_pred_u<true>(a, variant, c->tp, c->stream);
    else launch_pred_u<false>(a, variant, c->tp, c->stream);
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

const char* pred_variant_name(int v)
{
    static char buf[kPredVariants][48];
    if (v < 0 || v >= kPredVariants) return "?";
    std::snprintf(buf[v], sizeof buf[v], "%d records per group, %d threads per block%s", kPredR[v],
                  kPredBS[v], v >= 5 ? ", gradient deferred" : "");
    return buf[v];
}

int st_prediction(vox_ctx* c)
{
    VOX_TRY(list_l_for_consumer(c));
    if (c->p.precision == VOX_PREC_MIXED) return st_prediction32(c);
    ProfScope ps(c, KC_PRED);
    const vox_params& p = c->p;
    PredArgs a;
    a.n = c->tp; a.dim = p.dim; a.ktype = p.ktype; a.re_s = p.re_s; a.re_l = p.re_l;
    a.rho1 = p.rho1; a.rho2 = p.rho2; a.gx = p.gx; a.gy = p.gy; a.gz = p.gz; a.dt = p.dt;
    a.lapc = p.coef2d ? 4.0 : 2.0 * p.dim;
    a.ln0 = p.lambda * c->n0l;
    a.gcoef_num = p.coef2d ? 2.0 : (double)p.dim;
    a.n0s = c->n0s;
    a.relax1 = (1 - p.relaxp);
    a.L = view(c, VOX_LIST_L); a.pt = c->ptype;
    a.x = c->f[VOX_F_X]; a.y = c->f[VOX_F_Y]; a.z = c->f[VOX_F_Z];
    a.u = c->f[VOX_F_U]; a.v = c->f[VOX_F_V]; a.w = c->f[VOX_F_W];
    a.pn = c->f[VOX_F_PNEW]; a.ph = c->f[VOX_F_PHAT]; a.meu = c->f[VOX_F_MEU];
    a.neut = c->f[VOX_F_NEUT]; a.rho = c->f[VOX_F_RHO];
    a.t1 = c->f[VOX_F_TURB1]; a.t2 = c->f[VOX_F_TURB2]; a.t3 = c->f[VOX_F_TURB3];
    a.xs = c->f[VOX_F_XSTAR]; a.ys = c->f[VOX_F_YSTAR]; a.zs = c->f[VOX_F_ZSTAR];
    a.us = c->f[VOX_F_USTAR]; a.vs = c->f[VOX_F_VSTAR]; a.ws = c->f[VOX_F_WSTAR];
    a.rec = c->pred_rec;
    // rows in list L's cell order: a wave's particles share most neighbours
    // (2.5% faster than particle order, whose row reads are contiguous)
    a.order = list_order(c, VOX_LIST_L);
    a.pos = list_of(c, VOX_LIST_L).pos_rows && a.order;
    a.pos_of = list_of(c, VOX_LIST_L).pos_of;
    a.rpt = c->pt_sorted;
    k_pack_pred<<<grid_for(c->tp), kBlock, 0, c->stream>>>(a);
    VOX_TRY(launch_prediction(c, a, c->tune_pred));
    return VOX_OK;
}

// ---- BCON: functions.cu:433-445 -----------------------------------------------------------------
__global__ void k_bcon(int n, const int32_t* __restrict__ pt, const double* __restrict__ ns,
                       double n0, double dirichlet, int32_t* __restrict__ bcon)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int b;
    if (pt[i] <= -1) b = -1;
    else if (ns[i] / n0 < dirichlet) b = 1;
    else b = 0;
    bcon[i] = b;
}

int st_bcon(vox_ctx* c)
{
    ProfScope ps(c, KC_BCON);
    k_bcon<<<grid_for(c->tp), kBlock, 0, c->stream>>>(c->tp, c->ptype, c->f[VOX_F_NSTAR], c->n0s,
                                                     c->p.dirichlet, c->bcon);
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

// ---- pressure clamp of PRESSURECALC: functions.cu:2534-2547 --------------------------------------
__global__ void k_clamp(int n, double pmin, double pmax, double* __restrict__ pn)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = pn[i];
    if (v < pmin) v = pmin;
    if (v > pmax || v * 0.0 != 0.0) v = pmax;
    pn[i] = v;
}

int st_clamp_pressure(vox_ctx* c)
{
    ProfScope ps(c, KC_MISC);
    k_clamp<<<grid_for(c->tp), kBlock, 0, c->stream>>>(c->tp, c->p.pmin, c->p.pmax,
                                                      c->f[VOX_F_PNEW]);
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

// ---- PHATCALC: functions.cu:207-224 -------------------------------------------------------------
__global__ void k_phat(int n, ListView L, const double* __restrict__ pn, double* __restrict__ ph)
{
    int i = xcd_tid();
    if (i >= n) return;
    double mn = 999999999999999.0;
    int cnt = L.cnt[i];
    const int32_t* __restrict__ rowp = L.ids + sl_index(i, 0, L.cap);
    // All 8 loads of a chunk are issued before the compares (padding slots
    // load p_i and are dropped by the select); entries are compared in row
    // order with a strict <, so ties keep the first value as before.
    for (int g = 0; (g << 3) < cnt; ++g) {
        int32_t jc[8];
        double v[8];
        load_chunk8(rowp, g, jc);
#pragma unroll
        for (int e = 0; e < 8; ++e) jc[e] = ((g << 3) + e < cnt) ? jc[e] : i;
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = pn[jc[e]];
#pragma unroll
        for (int e = 0; e < 8; ++e) mn = ((g << 3) + e < cnt && v[e] < mn) ? v[e] : mn;
    }
    ph[i] = mn;
}

int st_phat(vox_ctx* c)
{
    if (c->p.precision == VOX_PREC_MIXED) return st_phat32(c);
    ProfScope ps(c, KC_PHAT);
    k_phat<<<grid_for(c->tp), kBlock, 0, c->stream>>>(c->tp, view(c, VOX_LIST_S),
                                                     c->f[VOX_F_PNEW], c->f[VOX_F_PHAT]);
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

// ---- PRESSGRAD: functions.cu:245-355 --------------------------------------------------------------
struct GradArgs {
    int first, n, dim, ktype, kh;
    double re, gnum, dt, n0, relaxp, vmax;
    ListView L;
    const double *x, *y, *z, *pn, *ph, *rho, *us, *vs, *ws;
    double *un, *vn, *wn;
    struct GradRec* rec;
};

// One 64-byte record per particle with every field the gradient gathers
// (x*, y*, z* with z = 0 in 2D, p, p^).
struct __align__(64) GradRec { double x, y, z, pn, ph, pad0, pad1, pad2; };

__global__ void k_pack_grad(GradArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    GradRec r;
    r.x = a.x[i]; r.y = a.y[i]; r.z = (a.dim == 3) ? a.z[i] : 0.0;
    r.pn = a.pn[i]; r.ph = a.ph[i];
    r.pad0 = r.pad1 = r.pad2 = 0.0;
    a.rec[i] = r;
}

// Branch-free per entry with 8 records in flight; the 3-term distance equals
// the 2D one bit for bit, and the z sum is only read in 3D.
template <int KT>
__global__ __launch_bounds__(kBlock) void k_pressgrad(GradArgs a)
{
    int i = a.first + xcd_tid();
    if (i >= a.n) return;
    const int dim = a.dim;
    const GradRec me = a.rec[i];
    const double xi = me.x, yi = me.y, zi = me.z, pi_ = me.pn, phi_ = me.ph;
    double s1 = 0, s2 = 0, s3 = 0;
    const double inv_re = 1.0 / a.re;
    int cnt = a.L.cnt[i];
    const int32_t* __restrict__ rowp = a.L.ids + sl_index(i, 0, a.L.cap);
    for (int g = 0; (g << 3) < cnt; ++g) {
        int32_t jc[8];
        load_chunk8(rowp, g, jc);
        double px[8], py[8], pz[8], pp[8], ph[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) jc[e] = ((g << 3) + e < cnt) ? jc[e] : i;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const GradRec* r = a.rec + jc[e];
            px[e] = r->x; py[e] = r->y; pz[e] = r->z; pp[e] = r->pn; ph[e] = r->ph;
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const bool use = (jc[e] != i);
            const double dx = px[e] - xi, dy = py[e] - yi, dz = pz[e] - zi;
            const double D = sqrt(dx * dx + dy * dy + dz * dz);
            const double wv = wfun_k<KT>(D, a.ktype, a.re, inv_re);
            // kh 1: p_j + p_i - (p^_i + p^_j) (functions.cu:275-278); kh 2, the
            // reference's CUDA kernel: p_j + p_i - 2 p^_i (382-385); kh 0: p_j - p^_i
            const double dp = (a.kh == 1)   ? (pp[e] + pi_ - (phi_ + ph[e]))
                              : (a.kh == 2) ? (pp[e] + pi_ - 2.0 * phi_)
                                            : (pp[e] - phi_);
            const double inv_D = 1.0 / D;                 // x / D / D via div_by
            s1 = use ? s1 + div_by(div_by(dp * dx * wv, D, inv_D), D, inv_D) : s1;
            s2 = use ? s2 + div_by(div_by(dp * dy * wv, D, inv_D), D, inv_D) : s2;
            s3 = use ? s3 + div_by(div_by(dp * dz * wv, D, inv_D), D, inv_D) : s3;
        }
    }
    double g = a.relaxp * (a.gnum * a.dt / a.n0 / a.rho[i]);
    double un = a.us[i] - g * s1;
    double vn = a.vs[i] - g * s2;
    double cf = 2.0;
    if (fabs(un) > cf * a.vmax) un = a.vmax;                   // damper 345-351
    if (vn > cf * a.vmax) vn = cf * a.vmax;
    if (vn < -cf * a.vmax) vn = -cf * a.vmax;
    a.un[i] = un;
    a.vn[i] = vn;
    if (dim == 3) {
        double wn = a.ws[i] - g * s3;
        if (fabs(wn) > cf * a.vmax) wn = a.vmax;
        a.wn[i] = wn;
    }
}

int st_pressgrad(vox_ctx* c)
{
    if (c->p.precision == VOX_PREC_MIXED) return st_pressgrad32(c);
    ProfScope ps(c, KC_PGRAD);
    const vox_params& p = c->p;
    GradArgs a;
    a.first = c->gp + c->wp; a.n = c->tp; a.dim = p.dim; a.ktype = p.ktype;
    a.kh = (p.ref_gpu && p.kh_correction == 1) ? 2 : p.kh_correction;
    a.re = p.re_s; a.gnum = p.coef2d ? 2.0 : (double)p.dim;
    a.dt = p.dt; a.n0 = c->n0s; a.relaxp = p.relaxp; a.vmax = p.vmax;
    a.L = view(c, VOX_LIST_S);
    a.x = c->f[VOX_F_XSTAR]; a.y = c->f[VOX_F_YSTAR]; a.z = c->f[VOX_F_ZSTAR];
    a.pn = c->f[VOX_F_PNEW]; a.ph = c->f[VOX_F_PHAT]; a.rho = c->f[VOX_F_RHO];
    a.us = c->f[VOX_F_USTAR]; a.vs = c->f[VOX_F_VSTAR]; a.ws = c->f[VOX_F_WSTAR];
    a.un = c->f[VOX_F_UNEW]; a.vn = c->f[VOX_F_VNEW]; a.wn = c->f[VOX_F_WNEW];
    a.rec = reinterpret_cast<GradRec*>(c->pred_rec);     // the predictor's 64-byte slots
    int m = c->tp - a.first;
    if (m <= 0) return VOX_OK;
    k_pack_grad<<<grid_for(c->tp), kBlock, 0, c->stream>>>(a);
    switch (kt_of(a.ktype)) {
    case 2: k_pressgrad<2><<<grid_for(m), kBlock, 0, c->stream>>>(a); break;
    case 6: k_pressgrad<6><<<grid_for(m), kBlock, 0, c->stream>>>(a); break;
    default: k_pressgrad<0><<<grid_for(m), kBlock, 0, c->stream>>>(a); break;
    }
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

// ---- EULERINTEGRATION + BOUNDARIE: functions.cu:1963-2021 -------------------------------------------
struct EulerArgs {
    int first, n, dim, kind;
    double dt, dl, cc, xmin, xmax, ymin, ymax, zmin, zmax;
    double *x, *y, *z, *u, *v, *w;
    const int32_t* region;      // per-particle region (batched scenarios) or null
    const double* rbox;         // 6 bounds per region
};

// The grid box bounds, or those of the particle's region when scenarios are
// batched (each keeps the box of its own run).
struct Box { double xmin, xmax, ymin, ymax, zmin, zmax; };
__device__ __forceinline__ Box box_of(const int32_t* region, const double* rbox, int i, Box g)
{
    if (!region) return g;
    const double* b = rbox + 6L * region[i];
    return Box{b[0], b[1], b[2], b[3], b[4], b[5]};
}

constexpr int kEulerRefGpu = -1;     // EulerArgs::kind of the CUDA-path boundary rule

__global__ void k_euler(EulerArgs a)
{
    int i = a.first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    {
        const Box b = box_of(a.region, a.rbox, i, Box{a.xmin, a.xmax, a.ymin, a.ymax, a.zmin, a.zmax});
        a.xmin = b.xmin; a.xmax = b.xmax; a.ymin = b.ymin; a.ymax = b.ymax; a.zmin = b.zmin;
        a.zmax = b.zmax;
    }
    double x = a.x[i] + a.u[i] * a.dt;
    double y = a.y[i] + a.v[i] * a.dt;
    double u = a.u[i], v = a.v[i];
    double z = 0.0, w = 0.0;
    if (a.dim == 3) { w = a.w[i]; z = a.z[i] + w * a.dt; }
    const double DL = a.dl, CC = a.cc;
    if (a.kind == kEulerRefGpu) {
        // the reference's CUDA kernel (functions.cu:2023-2047): every test
        // kind, upper y bound Ymax - DL/2
        if (y < a.ymin + DL * 3.0) { v = fabs(v) * CC; y = a.ymin + DL * 3.0; }
        if (y > a.ymax - DL / 2.0) { v = -fabs(v) * CC; y = a.ymax - DL / 2.0; }
        if (x < a.xmin + DL * 3.0) { u = fabs(u) * CC; x = a.xmin + DL * 3.0; }
        if (x > a.xmax - DL * 3.0) { u = -fabs(u) * CC; x = a.xmax - DL * 3.0; }
        if (a.dim == 3) {
            if (z < a.zmin + DL * 3.0) { w = fabs(w) * CC; z = a.zmin + DL * 3.0; }
            if (z > a.zmax - DL * 3.0) { w = -fabs(w) * CC; z = a.zmax - DL * 3.0; }
        }
    } else if (a.kind == VOX_TEST_DAM) {
        if (y < a.ymin + DL * 3.0) { v = fabs(v) * CC; y = a.ymin + DL * 3.0; }
        if (y > a.ymax - DL * 3.0) { v = -fabs(v) * CC; y = a.ymax - DL * 3.0; }
        if (x < a.xmin + DL * 3.0) { u = fabs(u) * CC; x = a.xmin + DL * 3.0; }
        if (x > a.xmax - DL * 3.0) { u = -fabs(u) * CC; x = a.xmax - DL * 3.0; }
        if (a.dim == 3) {
            if (z < a.zmin + DL * 3.0) { w = fabs(w) * CC; z = a.zmin + DL * 3.0; }
            if (z > a.zmax - DL * 3.0) { w = -fabs(w) * CC; z = a.zmax - DL * 3.0; }
        }
    } else if (a.kind == VOX_TEST_TURB) {
        if (y < a.ymin + DL * 2.0) { v = fabs(v) * CC; y = a.ymin + DL * 2.0; }
        if (y > a.ymax - DL * 2.0) { v = -fabs(v) * CC; y = a.ymax - DL * 2.0; }
        if (x > a.xmax + DL * 0.0) { u = (u) * CC; x = a.xmin; }
    }
    a.x[i] = x; a.y[i] = y; a.u[i] = u; a.v[i] = v;
    if (a.dim == 3) { a.z[i] = z; a.w[i] = w; }
}

int st_euler(vox_ctx* c)
{
    ProfScope ps(c, KC_EULER);
    ++c->pos_epoch;
    const vox_params& p = c->p;
    EulerArgs a;
    a.first = c->gp + c->wp; a.n = c->tp; a.dim = p.dim;
    // the CUDA path passes the collision factor CC where the CPU path passes
    // velCor (main.cu:1293 and the GPU branch beside it)
    a.kind = p.ref_gpu ? kEulerRefGpu : p.test_kind;
    a.dt = p.dt; a.dl = p.dl; a.cc = p.ref_gpu ? p.cc : p.vel_cor;
    a.xmin = p.xmin; a.xmax = p.xmax; a.ymin = p.ymin; a.ymax = p.ymax; a.zmin = p.zmin; a.zmax = p.zmax;
    a.x = c->f[VOX_F_X]; a.y = c->f[VOX_F_Y]; a.z = c->f[VOX_F_Z];
    a.u = c->f[VOX_F_UNEW]; a.v = c->f[VOX_F_VNEW]; a.w = c->f[VOX_F_WNEW];
    a.region = c->region; a.rbox = c->rbox;
    int m = c->tp - a.first;
    if (m <= 0) return VOX_OK;
    k_euler<<<grid_for(m), kBlock, 0, c->stream>>>(a);
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

// ---- COLLISION2: functions.cu:1523-1601 ------------------------------------------------------------
// The reference updates particles i and j in place while it walks i = 1..TP and
// each row in order, so a later pair sees earlier updates. That order is
// reproduced exactly: a parallel screen lists every (i, j) with d < coll+margin
// in (i, slot) order; the candidates are then replayed in the reference order
// with the live positions, one thread per connected component of the candidate
// pairs (components share no particle, so they are independent). A pair
// outside the screen can only reach d < coll if its particles move by more
// than margin in total during the pass; the replay accumulates each touched
// particle's path length and raises a flag if any exceeds margin/2 — then an
// undo log restores the state and the pass repeats with a wider screen.
// Positions come as packed 32-byte records (z = 0 in 2D: the 3-term distance
// equals the 2D one bit for bit); 8 record loads are in flight per row chunk
// and candidates are emitted in slot order.
__global__ __launch_bounds__(kBlock) void k_coll_screen(int n, double rc, ListView L,
                                                        const PosRec* __restrict__ rec,
                                                        const int32_t* __restrict__ off,
                                                        int32_t* __restrict__ cnt,
                                                        int32_t* __restrict__ ci,
                                                        int32_t* __restrict__ cj, long cap)
{
    int i = xcd_tid();
    if (i >= n) return;
    const PosRec me = rec[i];
    long pos = off ? off[i] : 0;
    int k = 0;
    int rows = L.cnt[i];
    const int32_t* __restrict__ rowp = L.ids + sl_index(i, 0, L.cap);
    for (int g = 0; (g << 3) < rows; ++g) {
        int32_t jc[8];
        load_chunk8(rowp, g, jc);
        double px[8], py[8], pz[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) jc[e] = ((g << 3) + e < rows) ? jc[e] : i;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const PosRec* r = rec + jc[e];
            px[e] = r->x; py[e] = r->y; pz[e] = r->z;
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const double dx = px[e] - me.x, dy = py[e] - me.y, dz = pz[e] - me.z;
            const double d = sqrt(dx * dx + dy * dy + dz * dz);
            const bool hit = (jc[e] != i) && (d < rc);
            if (hit && off && pos + k < cap) { ci[pos + k] = i; cj[pos + k] = jc[e]; }
            k += hit ? 1 : 0;
        }
    }
    if (!off) cnt[i] = k;
}

struct CollArgs {
    int dim;
    double coll, cc, rho1, rho2, dt, half_margin;
    const int32_t* total;
    const int32_t *ci, *cj, *pt;
    long cap;
    double *x, *y, *z, *u, *v, *w;
    int32_t *stamp, *logidx, *logn, *flags;
    double* log;
    int32_t epoch;
};

// The reference loop body's test and impulse for the pair (i, j)
// (functions.cu:1540-1575) on the current state: returns 0 if no impulse
// applies, else fills the impulse (um, vm, wm) and the masses.
__device__ __forceinline__ int coll_eval(const CollArgs& a, int i, int j, double& m1, double& m2,
                                         double& um, double& vm, double& wm)
{
    const int dim = a.dim;
    const int ti = a.pt[i], tj = a.pt[j];
    m1 = (ti == 1) ? a.rho1 : a.rho2;
    m2 = (tj == 1) ? a.rho1 : a.rho2;
    double dx = a.x[j] - a.x[i], dy = a.y[j] - a.y[i], dz = 0.0, d;
    if (dim == 3) { dz = a.z[j] - a.z[i]; d = sqrt(dx * dx + dy * dy + dz * dz); }
    else d = sqrt(dx * dx + dy * dy);
    if (!(d < a.coll)) return 0;
    double ug = (m1 * a.u[i] + m2 * a.u[j]) / (m1 + m2);
    double vg = (m1 * a.v[i] + m2 * a.v[j]) / (m1 + m2);
    double wg = 0, wr = 0;
    wm = 0;
    if (dim == 3) wg = (m1 * a.w[i] + m2 * a.w[j]) / (m1 + m2);
    double ur = m1 * (a.u[i] - ug);
    double vr = m1 * (a.v[i] - vg);
    if (dim == 3) wr = m1 * (a.w[i] - wg);
    double vabs = (dim == 3) ? (ur * dx + vr * dy + wr * dz) / d : (ur * dx + vr * dy) / d;
    um = (1.0 + a.cc) * vabs * dx / d;
    vm = (1.0 + a.cc) * vabs * dy / d;
    if (dim == 3) wm = (1.0 + a.cc) * vabs * dz / d;
    return (vabs > 0) ? 1 : 0;
}

// Applies candidate k exactly as the reference loop body (functions.cu:1540-1595)
// with the live positions; logs each particle the first time it moves (for the
// undo) and accumulates its path length. Returns 1 if an impulse was applied.
__device__ __forceinline__ int coll_apply(const CollArgs& a, int k)
{
    const int dim = a.dim;
    const int i = a.ci[k], j = a.cj[k];
    const int ti = a.pt[i], tj = a.pt[j];
    double m1, m2, um, vm, wm;
    if (!coll_eval(a, i, j, m1, m2, um, vm, wm)) return 0;
    for (int side = 0; side < 2; ++side) {
        const int q = side ? j : i;
        if ((side ? tj : ti) <= 0) continue;
        if (a.stamp[q] != a.epoch) {
            a.stamp[q] = a.epoch;
            const int slot = atomicAdd(a.logn, 1);
            a.logidx[q] = slot;
            double* e = a.log + 8L * slot;
            e[0] = (double)q; e[1] = a.x[q]; e[2] = a.y[q]; e[3] = a.u[q]; e[4] = a.v[q];
            e[5] = (dim == 3) ? a.z[q] : 0.0; e[6] = (dim == 3) ? a.w[q] : 0.0; e[7] = 0.0;
        }
        const double m = side ? m2 : m1;
        double nx, ny, nz = 0.0;
        if (!side) {
            a.u[q] = a.u[q] - um / m;
            a.v[q] = a.v[q] - vm / m;
            nx = a.x[q] - a.dt * um / m;
            ny = a.y[q] - a.dt * vm / m;
        } else {
            a.u[q] = a.u[q] + um / m;
            a.v[q] = a.v[q] + vm / m;
            nx = a.x[q] + a.dt * um / m;
            ny = a.y[q] + a.dt * vm / m;
        }
        double step = fabs(nx - a.x[q]) + fabs(ny - a.y[q]);
        a.x[q] = nx; a.y[q] = ny;
        if (dim == 3) {
            if (!side) { a.w[q] = a.w[q] - wm / m; nz = a.z[q] - a.dt * wm / m; }
            else { a.w[q] = a.w[q] + wm / m; nz = a.z[q] + a.dt * wm / m; }
            step += fabs(nz - a.z[q]);
            a.z[q] = nz;
        }
        a.log[8L * a.logidx[q] + 7] += step;
    }
    return 1;
}

// Single-lane replay of every candidate in (i, slot) order.
__global__ void k_coll_replay(CollArgs a)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int total = *a.total;
    int hits = 0;
    for (int k = 0; k < total && k < a.cap; ++k) hits += coll_apply(a, k);
    a.flags[FLAG_COLL_HITS] = hits;
}

// Candidate components. Candidate pairs are listed in both orders (the screen
// is symmetric), so the candidates of particle p are ci == p at off[p] ..
// off[p] + cnt[p] and its partners are their cj. Particles joined through
// candidate pairs form components; a candidate only reads and moves its two
// particles, so the sequential replay of one component is independent of every
// other component, and the reference order restricted to a component is its
// members in ascending id, each with its own candidates in slot order.
constexpr int kCompMax = 48;   // members one thread collects

// Members of p's component in ascending id (at most kCompMax; returns -1 if
// the component is larger).
__device__ int coll_component(int p, const int32_t* __restrict__ off,
                              const int32_t* __restrict__ cnt, const int32_t* __restrict__ cj,
                              int* mem)
{
    int m = 1, head = 0;
    mem[0] = p;
    while (head < m) {
        const int q = mem[head++];
        const int s = off[q], e = s + cnt[q];
        for (int k = s; k < e; ++k) {
            const int r = cj[k];
            bool seen = false;
            for (int t = 0; t < m; ++t) seen = seen || (mem[t] == r);
            if (seen) continue;
            if (m == kCompMax) return -1;
            mem[m++] = r;
        }
    }
    for (int t = 1; t < m; ++t) {            // insertion sort, ascending id
        const int v = mem[t];
        int u = t - 1;
        while (u >= 0 && mem[u] > v) { mem[u + 1] = mem[u]; --u; }
        mem[u + 1] = v;
    }
    return m;
}

// Pass 1: the smallest member of each particle's component; flags components
// beyond the per-thread bound.
__global__ __launch_bounds__(kBlock) void k_coll_roots(int n, const int32_t* __restrict__ off,
                                                       const int32_t* __restrict__ cnt,
                                                       const int32_t* __restrict__ cj,
                                                       int32_t* __restrict__ root,
                                                       int32_t* __restrict__ flags)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    if (cnt[p] == 0) { root[p] = -2; return; }   // never equal to a candidate end's root
    int mem[kCompMax];
    const int m = coll_component(p, off, cnt, cj, mem);
    if (m < 0) { atomicMax(&flags[FLAG_COLL_BIG], 1); root[p] = -1; return; }
    root[p] = mem[0];
}

// The component construction relies on every candidate pair being listed in
// both orders (symmetric lists); a pair whose two ends found different
// components means it was not, and the replay falls back to a single lane.
__global__ void k_coll_verify(const int32_t* __restrict__ total, const int32_t* __restrict__ ci,
                              const int32_t* __restrict__ cj, const int32_t* __restrict__ root,
                              int32_t* __restrict__ flags)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= *total) return;
    if (root[ci[k]] != root[cj[k]]) flags[FLAG_COLL_BIG] = 2;
}

// Pass 2: one thread per component (its smallest member) replays the
// component's candidates in reference order.
__global__ __launch_bounds__(kBlock) void k_coll_replay_comp(CollArgs a, int n,
                                                             const int32_t* __restrict__ off,
                                                             const int32_t* __restrict__ cnt,
                                                             const int32_t* __restrict__ root)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || cnt[p] == 0 || root[p] != p) return;
    int mem[kCompMax];
    const int m = coll_component(p, off, cnt, a.cj, mem);
    if (m < 0) return;                       // not reached: pass 1 flagged it
    int hits = 0;
    for (int t = 0; t < m; ++t) {
        const int q = mem[t];
        for (int k = off[q], e = off[q] + cnt[q]; k < e; ++k) hits += coll_apply(a, k);
    }
    if (hits) atomicAdd(&a.flags[FLAG_COLL_HITS], hits);
}

// Any touched particle that moved more than half the screen margin.
__global__ void k_coll_check(CollArgs a)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= *a.logn) return;
    if (a.log[8L * k + 7] > a.half_margin) a.flags[FLAG_COLL_VIOLATION] = 1;
}

__global__ void k_coll_undo(int dim, const int32_t* __restrict__ logn, const double* __restrict__ log,
                            double* x, double* y, double* z, double* u, double* v, double* w)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= *logn) return;
    const double* e = log + 8L * k;
    int q = (int)e[0];
    x[q] = e[1]; y[q] = e[2]; u[q] = e[3]; v[q] = e[4];
    if (dim == 3) { z[q] = e[5]; w[q] = e[6]; }
}

// ---- active set: the replay is restricted to the candidates that touch a
// particle which may move. A candidate evaluates on the state its two
// particles have when the sequential loop reaches it; if neither was moved by
// an earlier candidate, that is the state before the pass. So every candidate
// that hits on the state before the pass marks its movable particles active;
// the replay runs, in reference order, exactly the candidates with an active
// end; and if it moves a particle that is not active, that particle joins the
// set and the replay is redone from the state before the pass. When no
// particle outside the set moves, every skipped candidate saw the state before
// the pass, where it does not hit: the result is the full sequential replay's.

// Marks the movable particles of every candidate that hits on the current state.
__global__ void k_coll_mark(CollArgs a, int32_t* __restrict__ act)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= *a.total || k >= a.cap) return;
    const int i = a.ci[k], j = a.cj[k];
    double m1, m2, um, vm, wm;
    if (!coll_eval(a, i, j, m1, m2, um, vm, wm)) return;
    if (a.pt[i] > 0) act[i] = 1;
    if (a.pt[j] > 0) act[j] = 1;
}

// Active candidates per particle (pass 0 counts, pass 1 writes them in order).
__global__ __launch_bounds__(kBlock) void k_coll_filter(int n, const int32_t* __restrict__ off,
                                                        const int32_t* __restrict__ cnt,
                                                        const int32_t* __restrict__ cj,
                                                        const int32_t* __restrict__ act,
                                                        int32_t* __restrict__ cnt2,
                                                        const int32_t* __restrict__ off2,
                                                        int32_t* __restrict__ ci2,
                                                        int32_t* __restrict__ cj2)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const bool ap = act[p] != 0;
    int m = 0;
    const int s = off[p], e = s + cnt[p];
    for (int k = s; k < e; ++k) {
        const int q = cj[k];
        if (!(ap || act[q] != 0)) continue;
        if (off2) { ci2[off2[p] + m] = p; cj2[off2[p] + m] = q; }
        ++m;
    }
    if (!off2) cnt2[p] = m;
}

// A particle moved by the replay that is not active: add it, flag the pass.
__global__ void k_coll_closure(const int32_t* __restrict__ logn, const double* __restrict__ log,
                               int32_t* __restrict__ act, int32_t* __restrict__ flags)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= *logn) return;
    const int q = (int)log[8L * k];
    if (act[q] == 0) {
        act[q] = 1;
        flags[FLAG_COLL_OPEN] = 1;
    }
}

static int coll_reserve(vox_ctx* c, long need)
{
    if (need <= c->coll_cap) return VOX_OK;
    long cap = need + need / 2 + 1024;
    for (int32_t** q : {&c->coll_i, &c->coll_j, &c->coll_i2, &c->coll_j2}) {
        if (*q) (void)hipFree(*q);
        *q = nullptr;
    }
    if (c->coll_log) (void)hipFree(c->coll_log);
    c->coll_log = nullptr;
    VOX_HIP(hipMalloc(&c->coll_i, sizeof(int32_t) * cap));
    VOX_HIP(hipMalloc(&c->coll_j, sizeof(int32_t) * cap));
    VOX_HIP(hipMalloc(&c->coll_i2, sizeof(int32_t) * cap));
    VOX_HIP(hipMalloc(&c->coll_j2, sizeof(int32_t) * cap));
    VOX_HIP(hipMalloc(&c->coll_log, sizeof(double) * 8 * (2 * cap)));
    c->coll_cap = cap;
    return VOX_OK;
}

__global__ void k_coll_reset(int32_t* __restrict__ flags, int32_t* __restrict__ logn)
{
    if (threadIdx.x == 0) {
        flags[FLAG_COLL_HITS] = 0;
        flags[FLAG_COLL_VIOLATION] = 0;
        flags[FLAG_COLL_BIG] = 0;
        flags[FLAG_COLL_OPEN] = 0;
        *logn = 0;
    }
}

// Single-lane replay of the candidates of the components too large for one
// thread's member list (root == -1), in (i, slot) order.
__global__ void k_coll_replay_big(CollArgs a, const int32_t* __restrict__ root)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int total = *a.total;
    int hits = 0;
    for (int k = 0; k < total && k < a.cap; ++k)
        if (root[a.ci[k]] < 0) hits += coll_apply(a, k);
    a.flags[FLAG_COLL_HITS] += hits;
}

constexpr double kCollMarginBase = 0.1;     // screen margin, fraction of coll

int st_collision(vox_ctx* c)
{
    ProfScope ps(c, KC_COLL);
    ++c->pos_epoch;
    const vox_params& p = c->p;
    const int n = c->tp;
    double* X = c->f[VOX_F_X]; double* Y = c->f[VOX_F_Y]; double* Z = c->f[VOX_F_Z];
    double* U = c->f[VOX_F_UNEW]; double* V = c->f[VOX_F_VNEW]; double* W = c->f[VOX_F_WNEW];
    ListView L = view(c, VOX_LIST_S);
    // The margin widens on a violation and relaxes back towards its base over
    // the following steps.
    c->coll_margin = std::max(kCollMarginBase, 0.75 * c->coll_margin);
    // Once coll + margin exceeds the list's own accept radius the screen takes
    // every listed pair (radius +inf) and the motion check is off: the replay
    // then evaluates exactly the pairs the reference loop visits, for any
    // motion, even though the list was built at the step's earlier positions.
    const double list_r = (p.dim == 3 || p.ref_gpu) ? p.re_s + p.delta : p.re_s;
    for (int attempt = 0; attempt < 8; ++attempt) {
        double margin = p.coll * c->coll_margin;
        const bool every_pair = p.coll + margin > list_r;
        const double rc = every_pair ? HUGE_VAL : p.coll + margin;
        k_pack_pos<<<grid_for(n), kBlock, 0, c->stream>>>(n, p.dim, X, Y, Z, c->pos_rec);
        k_coll_screen<<<grid_for(n), kBlock, 0, c->stream>>>(n, rc, L, c->pos_rec, nullptr,
                                                              c->coll_cnt, nullptr, nullptr, 0);
        VOX_TRY(scan_exclusive(c, c->coll_cnt, c->coll_off, n, c->coll_off + n));
        int32_t total = 0;
        VOX_HIP(hipMemcpyAsync(&total, c->coll_off + n, sizeof(int32_t), hipMemcpyDeviceToHost,
                               c->stream));
        VOX_HIP(hipStreamSynchronize(c->stream));
        c->st.collision_cand = total;
        if (total == 0) { c->st.collisions = 0; return VOX_OK; }
        VOX_TRY(coll_reserve(c, total));
        k_coll_screen<<<grid_for(n), kBlock, 0, c->stream>>>(n, rc, L, c->pos_rec, c->coll_off,
                                                              nullptr, c->coll_i, c->coll_j,
                                                              c->coll_cap);
        CollArgs a;
        a.dim = p.dim; a.coll = p.coll; a.cc = p.cc; a.rho1 = p.rho1; a.rho2 = p.rho2; a.dt = p.dt;
        a.half_margin = every_pair ? HUGE_VAL : 0.45 * margin;
        a.total = c->coll_off + n; a.ci = c->coll_i; a.cj = c->coll_j; a.pt = c->ptype;
        a.cap = c->coll_cap;
        a.x = X; a.y = Y; a.z = Z; a.u = U; a.v = V; a.w = W;
        a.stamp = c->coll_stamp; a.logidx = c->coll_logidx; a.logn = c->coll_logn;
        a.flags = c->dflags; a.log = c->coll_log; a.epoch = c->coll_epoch;
        // active set from the state before the pass
        VOX_TRY(fill_i32(c, c->coll_act, 0, n));
        k_coll_mark<<<grid_for(total), kBlock, 0, c->stream>>>(a, c->coll_act);
        VOX_HIP(hipGetLastError());
        for (int round = 0;; ++round) {
            // the active candidates, in reference order
            k_coll_filter<<<grid_for(n), kBlock, 0, c->stream>>>(n, c->coll_off, c->coll_cnt, c->coll_j,
                                                                c->coll_act, c->coll_cnt2, nullptr,
                                                                nullptr, nullptr);
            VOX_TRY(scan_exclusive(c, c->coll_cnt2, c->coll_off2, n, c->coll_off2 + n));
            k_coll_filter<<<grid_for(n), kBlock, 0, c->stream>>>(n, c->coll_off, c->coll_cnt, c->coll_j,
                                                                c->coll_act, c->coll_cnt2, c->coll_off2,
                                                                c->coll_i2, c->coll_j2);
            CollArgs r = a;
            r.total = c->coll_off2 + n; r.ci = c->coll_i2; r.cj = c->coll_j2;
            if (++c->coll_epoch == INT32_MAX) {
                VOX_TRY(fill_i32(c, c->coll_stamp, 0, n));
                c->coll_epoch = 1;
            }
            r.epoch = c->coll_epoch;
            k_coll_reset<<<1, kWave, 0, c->stream>>>(c->dflags, c->coll_logn);
            k_coll_roots<<<grid_for(n), kBlock, 0, c->stream>>>(n, c->coll_off2, c->coll_cnt2, c->coll_j2,
                                                                 c->coll_root, c->dflags);
            k_coll_verify<<<grid_for(total), kBlock, 0, c->stream>>>(c->coll_off2 + n, c->coll_i2,
                                                                     c->coll_j2, c->coll_root, c->dflags);
            VOX_HIP(hipGetLastError());
            VOX_TRY(read_flags(c));
            if (c->hflags[FLAG_COLL_BIG] == 2) {
                k_coll_replay<<<1, kWave, 0, c->stream>>>(r);          // every active candidate, one lane
            } else {
                k_coll_replay_comp<<<grid_for(n), kBlock, 0, c->stream>>>(r, n, c->coll_off2,
                                                                           c->coll_cnt2, c->coll_root);
                if (c->hflags[FLAG_COLL_BIG])
                    k_coll_replay_big<<<1, kWave, 0, c->stream>>>(r, c->coll_root);
            }
            VOX_HIP(hipGetLastError());
            // at most two log entries per candidate
            k_coll_check<<<grid_for(2L * total), kBlock, 0, c->stream>>>(r);
            k_coll_closure<<<grid_for(2L * total), kBlock, 0, c->stream>>>(c->coll_logn, c->coll_log,
                                                                          c->coll_act, c->dflags);
            VOX_HIP(hipGetLastError());
            VOX_TRY(read_flags(c));
            c->st.collisions = c->hflags[FLAG_COLL_HITS];
            const bool violation = c->hflags[FLAG_COLL_VIOLATION] != 0;
            const bool open = c->hflags[FLAG_COLL_OPEN] != 0;
            if (!violation && !open) return VOX_OK;
            // undo; then either widen the screen or replay with the larger set
            k_coll_undo<<<grid_for(2L * total), kBlock, 0, c->stream>>>(p.dim, c->coll_logn, c->coll_log,
                                                                       X, Y, Z, U, V, W);
            VOX_HIP(hipGetLastError());
            if (violation) break;
            if (round > 64) return fail(VOX_ERR_SOLVER, "collision active set did not close");
        }
        c->coll_margin *= 2.0;
    }
    return fail(VOX_ERR_SOLVER, "collision screen could not bound particle motion");
}

// ---- PREPDATA: functions.cu:2688-2746 --------------------------------------------------------------
struct PrepArgs {
    int n, first, dim;
    double xmin, xmax, ymin, ymax, zmin, zmax;
    const double *x, *y, *z, *un, *vn, *wn, *pn;
    double *u, *v, *w, *p;
    int32_t* flags;
    const int32_t* region;
    const double* rbox;
};

__global__ __launch_bounds__(kBlock) void k_prepdata(PrepArgs a)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int nanc = 0, oob = 0;
    if (i < a.n) {
        double pv = a.pn[i];
        a.p[i] = pv;
        double u, v, w = 0.0;
        if (i >= a.first) {
            u = a.un[i]; v = a.vn[i];
            a.u[i] = u; a.v[i] = v;
            if (a.dim == 3) { w = a.wn[i]; a.w[i] = w; }
        } else {
            u = a.u[i]; v = a.v[i];
            if (a.dim == 3) w = a.w[i];
        }
        double x = a.x[i], y = a.y[i];
        const Box b = box_of(a.region, a.rbox, i, Box{a.xmin, a.xmax, a.ymin, a.ymax, a.zmin, a.zmax});
        if (a.dim == 3) {
            double z = a.z[i];
            if (pv * 0 != 0 || u * 0 != 0 || v * 0 != 0 || w * 0 != 0) nanc = 1;
            if (x >= b.xmax || y >= b.ymax || z >= b.zmax || x <= b.xmin || y <= b.ymin || z <= b.zmin)
                oob = 1;
        } else {
            if (pv * 0 != 0 || u * 0 != 0 || v * 0 != 0) nanc = 1;
            if (x >= b.xmax || y >= b.ymax || x <= b.xmin || y <= b.ymin) oob = 1;
        }
    }
    // one atomic per wave
    unsigned long long bn = __ballot(nanc), bo = __ballot(oob);
    if ((threadIdx.x % kWave) == 0) {
        if (bn) atomicAdd(&a.flags[FLAG_NAN], (int)__popcll(bn));
        if (bo) atomicAdd(&a.flags[FLAG_OOB], (int)__popcll(bo));
    }
}

int st_prepdata(vox_ctx* c)
{
    ProfScope ps(c, KC_PREP);
    ++c->pos_epoch;
    const vox_params& p = c->p;
    PrepArgs a;
    a.n = c->tp; a.dim = p.dim;
    a.first = (p.test_kind == VOX_TEST_DROP) ? c->tp - c->fp - 1 : c->tp - c->fp;
    if (p.ref_gpu) a.first = 0;          // the CUDA kernel copies u of every particle (2753)
    a.xmin = p.xmin; a.xmax = p.xmax; a.ymin = p.ymin; a.ymax = p.ymax; a.zmin = p.zmin; a.zmax = p.zmax;
    a.x = c->f[VOX_F_X]; a.y = c->f[VOX_F_Y]; a.z = c->f[VOX_F_Z];
    a.un = c->f[VOX_F_UNEW]; a.vn = c->f[VOX_F_VNEW]; a.wn = c->f[VOX_F_WNEW]; a.pn = c->f[VOX_F_PNEW];
    a.u = c->f[VOX_F_U]; a.v = c->f[VOX_F_V]; a.w = c->f[VOX_F_W]; a.p = c->f[VOX_F_P];
    a.flags = c->dflags;
    a.region = c->region; a.rbox = c->rbox;
    VOX_TRY(fill_i32(c, c->dflags + FLAG_NAN, 0, 2));
    k_prepdata<<<grid_for(c->tp), kBlock, 0, c->stream>>>(a);
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

}  // namespace vox
