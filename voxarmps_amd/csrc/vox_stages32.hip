// vox_stages32.hip — FP32 twins of the particle stages for the mixed
// precision mode: the predictor (viscous + explicit pressure terms), the
// pressure-gradient correction, p-hat, the SPS-LES eddy viscosity and stress
// divergence, the Herschel-Bulkley viscosity and the volume fraction.
//
// The particle state stays FP64: positions are integrated in FP64 and the
// particle number densities, which feed the PPE source (n* - n0)/n0 through a
// cancellation, keep their FP64 kernels. Each FP32 stage packs the fields its
// neighbour loop gathers into one record of fixed-point positions and FP32
// velocities and pressures. A position is stored as an unsigned 32-bit
// fraction of the grid box along each axis: the difference of two, taken
// modulo 2^32 as a signed integer, is exact (neighbours are far closer than
// half the box), so a neighbour offset keeps 2^-32 of the box extent before
// its single rounding to FP32. (A difference of FP32 positions would lose
// |x|/|dx|, 100-300 ulps in a metre-sized tank.) Weights, distances and the
// neighbour sums are FP32; outputs go back to the FP64 fields. Formulas and
// their order follow the FP64 kernels (PREDICTION functions.cu:2283-2369,
// PRESSGRAD 245-355).
#include "vox_internal.hpp"

namespace vox {

struct __align__(16) Pred32 { uint32_t x, y, z; float u, v, w, pn, meu; };   // 32 B
struct __align__(8) Grad32 { uint32_t x, y, z; float pn, ph, pad; };          // 24 B
struct __align__(16) Pos32 { uint32_t x, y, z; float f; };                    // 16 B
static_assert(sizeof(Pred32) == 32 && sizeof(Grad32) == 24 && sizeof(Pos32) == 16, "record sizes");

// Fixed-point frame: origin and 2^32/extent per axis (packing), extent/2^32
// per axis (differences).
struct Quant {
    double ox, oy, oz, ix, iy, iz;
    float sx, sy, sz;
};

static Quant make_quant(const vox_params& p)
{
    Quant q;
    const double two32 = 4294967296.0;
    const double ex = p.xmax - p.xmin, ey = p.ymax - p.ymin, ez = p.zmax - p.zmin;
    q.ox = p.xmin; q.oy = p.ymin; q.oz = p.zmin;
    q.ix = ex > 0 ? two32 / ex : 0.0;
    q.iy = ey > 0 ? two32 / ey : 0.0;
    q.iz = (p.dim == 3 && ez > 0) ? two32 / ez : 0.0;
    q.sx = (float)(ex > 0 ? ex / two32 : 0.0);
    q.sy = (float)(ey > 0 ? ey / two32 : 0.0);
    q.sz = (float)((p.dim == 3 && ez > 0) ? ez / two32 : 0.0);
    return q;
}

__device__ __forceinline__ uint32_t qpos(double v, double o, double inv)
{
    double t = (v - o) * inv;
    t = t < 0.0 ? 0.0 : (t > 4294967295.0 ? 4294967295.0 : t);
    return (uint32_t)rint(t);
}

__device__ __forceinline__ float qdiff(uint32_t a, uint32_t b, float s)
{
    return (float)(int32_t)(a - b) * s;
}

__device__ __forceinline__ float wfun32(float r, int ktype, float re)
{
    float q = r / re;
    if (ktype == 2) return (q < 1.0f) ? (1.0f / q) - 1.0f : 0.0f;
    if (ktype == 6) return (q <= 1.0f) ? (1.0f - q) * (1.0f - q) * (1.0f - q) : 0.0f;
    if (ktype == 1) {
        if (q < 0.5f) return 2.0f - 4.0f * (q * q);
        if (q <= 1.0f) return (2.0f * q - 2.0f) * (2.0f * q - 2.0f);
        return 0.0f;
    }
    if (ktype == 5) return (q <= 1.0f) ? 1.5f * logf(1.0f / (q + 0.000001f)) : 0.0f;
    return 0.0f;
}

// ---- PREDICTION ------------------------------------------------------------------------------
struct Pred32Args {
    int n, dim, ktype;
    double dt64;
    Quant q;
    float re_s, re_l, rho1, rho2, gx, gy, gz, dt, lapc, ln0, gcoef_num, n0s, relax1;
    ListView L;
    const int32_t* pt;
    const double *x, *y, *z, *u, *v, *w, *pn, *ph, *meu, *neut, *rho, *t1, *t2, *t3;
    double *xs, *ys, *zs, *us, *vs, *ws;
    Pred32* rec;              // the records the row entries index (see pos)
    const PosRec* order;      // thread t takes particle order[t].id (cell-sorted)
    int32_t* rpt;             // pos: the types in the same order as rec
    bool pos;                 // row entries are positions in order (list L's compact
                              // rows, stored at positions): rec and rpt are in that
                              // order and thread t's own record is rec[t]
    const int32_t* pos_of;    // pos: each particle's position in order
};

// Threads in particle order (fields read in lines); whole records stored.
__global__ void k_pack_pred32(Pred32Args a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int t = a.pos ? a.pos_of[i] : i;
    if (a.pos) a.rpt[t] = a.pt[i];
    Pred32 r;
    r.x = qpos(a.x[i], a.q.ox, a.q.ix);
    r.y = qpos(a.y[i], a.q.oy, a.q.iy);
    r.z = (a.dim == 3) ? qpos(a.z[i], a.q.oz, a.q.iz) : 0u;
    r.u = (float)a.u[i]; r.v = (float)a.v[i]; r.w = (a.dim == 3) ? (float)a.w[i] : 0.0f;
    r.pn = (float)a.pn[i]; r.meu = (float)a.meu[i];
    a.rec[t] = r;
}

// kUniform: every MEU is the same value (see uniform_meu): NEU does not
// depend on the neighbour's type, which is then not read.
template <bool kUniform>
__global__ __launch_bounds__(kBlock) void k_prediction32(Pred32Args a)
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
    const Pred32 me = a.rec[self];
    const float mi = me.meu, phi_ = (float)a.ph[i];
    const float rhoN = (ti == 1) ? a.rho1 : a.rho2;
    float s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0, s6 = 0, s7 = 0, s8 = 0, s9 = 0;
    const int cnt = a.L.cnt[self];              // rows of positions sit at positions
    const int32_t* __restrict__ row = a.L.ids + sl_index(self, 0, a.L.cap);
    constexpr int U = 8;
    for (int l0 = 0; l0 < cnt; l0 += U) {
        int jj[U], tj[U];
        Pred32 rv[U];
        load_chunk8(row, l0 >> 3, jj);
#pragma unroll
        for (int u = 0; u < U; ++u) jj[u] = (l0 + u < cnt) ? jj[u] : self;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            rv[u] = a.rec[jj[u]];
            tj[u] = kUniform ? ti : rpt[jj[u]];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const Pred32 rj = rv[u];
            const float dx = qdiff(rj.x, me.x, a.q.sx), dy = qdiff(rj.y, me.y, a.q.sy),
                        dz = qdiff(rj.z, me.z, a.q.sz);
            const float D = sqrtf(dx * dx + dy * dy + dz * dz);
            const bool use = (jj[u] != self) && (D > 0.0f);
            const float NEU = (kUniform || ti == tj[u]) ? mi / rhoN
                                                        : 2.0f * mi * rj.meu / (mi + rj.meu) / rhoN;
            const float wl = wfun32(D, a.ktype, a.re_l);
            const float dp = rj.pn - phi_;
            const float du = rj.u - me.u, dv = rj.v - me.v, dw = rj.w - me.w;
            // the gradient terms carry W(D; re_s) = 0 beyond re_s: skipped there
            if (use && D <= a.re_s) {
                const float g = wfun32(D, a.ktype, a.re_s) / D / D;
                s1 = s1 + dp * dx * g;
                s2 = s2 + dp * dy * g;
                s3 = s3 + dp * dz * g;
            }
            s4 = use ? s4 + wl * du * NEU : s4;
            s5 = use ? s5 + wl * dv * NEU : s5;
            s6 = use ? s6 + wl * dw * NEU : s6;
            s7 = use ? s7 + wl * du : s7;
            s8 = use ? s8 + wl * dv : s8;
            s9 = use ? s9 + wl * dw : s9;
        }
    }
    const float nt = (float)a.neut[i];
    const float gc = a.relax1 * (a.gcoef_num * a.dt / a.n0s / (float)a.rho[i]);
    const float dt = a.dt;
    const float us = me.u + a.gx * dt + a.lapc * dt * (s4 + nt * s7) / a.ln0 + dt * (float)a.t1[i] - gc * s1;
    const float vs = me.v + a.gy * dt + a.lapc * dt * (s5 + nt * s8) / a.ln0 + dt * (float)a.t2[i] - gc * s2;
    a.us[i] = (double)us;
    a.vs[i] = (double)vs;
    a.xs[i] = a.x[i] + a.dt64 * (double)us;
    a.ys[i] = a.y[i] + a.dt64 * (double)vs;
    if (dim == 3) {
        const float ws = me.w + a.gz * dt + a.lapc * dt * (s6 + nt * s9) / a.ln0 + dt * (float)a.t3[i] - gc * s3;
        a.ws[i] = (double)ws;
        a.zs[i] = a.z[i] + a.dt64 * (double)ws;
    }
}

int st_prediction32(vox_ctx* c)
{
    ProfScope ps(c, KC_PRED);
    const vox_params& p = c->p;
    Pred32Args a;
    a.n = c->tp; a.dim = p.dim; a.ktype = p.ktype;
    a.dt64 = p.dt;
    a.q = make_quant(p);
    a.re_s = (float)p.re_s; a.re_l = (float)p.re_l;
    a.rho1 = (float)p.rho1; a.rho2 = (float)p.rho2;
    a.gx = (float)p.gx; a.gy = (float)p.gy; a.gz = (float)p.gz; a.dt = (float)p.dt;
    a.lapc = p.coef2d ? 4.0f : 2.0f * p.dim;
    a.ln0 = (float)(p.lambda * c->n0l);
    a.gcoef_num = p.coef2d ? 2.0f : (float)p.dim;
    a.n0s = (float)c->n0s;
    a.relax1 = (float)(1 - p.relaxp);
    a.L = view(c, VOX_LIST_L); a.pt = c->ptype;
    a.x = c->f[VOX_F_X]; a.y = c->f[VOX_F_Y]; a.z = c->f[VOX_F_Z];
    a.u = c->f[VOX_F_U]; a.v = c->f[VOX_F_V]; a.w = c->f[VOX_F_W];
    a.pn = c->f[VOX_F_PNEW]; a.ph = c->f[VOX_F_PHAT]; a.meu = c->f[VOX_F_MEU];
    a.neut = c->f[VOX_F_NEUT]; a.rho = c->f[VOX_F_RHO];
    a.t1 = c->f[VOX_F_TURB1]; a.t2 = c->f[VOX_F_TURB2]; a.t3 = c->f[VOX_F_TURB3];
    a.xs = c->f[VOX_F_XSTAR]; a.ys = c->f[VOX_F_YSTAR]; a.zs = c->f[VOX_F_ZSTAR];
    a.us = c->f[VOX_F_USTAR]; a.vs = c->f[VOX_F_VSTAR]; a.ws = c->f[VOX_F_WSTAR];
    a.rec = reinterpret_cast<Pred32*>(c->pred_rec);          // 64-byte slots hold 32-byte records
    // threads in list L's cell order; rows of positions index records in that order
    a.order = list_order(c, VOX_LIST_L);
    a.pos = list_of(c, VOX_LIST_L).pos_rows && a.order;
    a.pos_of = list_of(c, VOX_LIST_L).pos_of;
    a.rpt = c->pt_sorted;
    k_pack_pred32<<<grid_for(c->tp), kBlock, 0, c->stream>>>(a);
    if (uniform_meu(c)) k_prediction32<true><<<grid_for(c->tp), kBlock, 0, c->stream>>>(a);
    else k_prediction32<false><<<grid_for(c->tp), kBlock, 0, c->stream>>>(a);
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

// ---- PRESSGRAD ------------------------------------------------------------------------------
struct Grad32Args {
    int first, n, dim, ktype, kh;
    Quant q;
    float re, gnum, dt, n0, relaxp, vmax;
    ListView L;
    const double *x, *y, *z, *pn, *ph, *rho, *us, *vs, *ws;
    double *un, *vn, *wn;
    Grad32* rec;
};

__global__ void k_pack_grad32(Grad32Args a)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    Grad32 r;
    r.x = qpos(a.x[i], a.q.ox, a.q.ix);
    r.y = qpos(a.y[i], a.q.oy, a.q.iy);
    r.z = (a.dim == 3) ? qpos(a.z[i], a.q.oz, a.q.iz) : 0u;
    r.pn = (float)a.pn[i]; r.ph = (float)a.ph[i]; r.pad = 0.0f;
    a.rec[i] = r;
}

__global__ __launch_bounds__(kBlock) void k_pressgrad32(Grad32Args a)
{
    const int i = a.first + xcd_tid();
    if (i >= a.n) return;
    const Grad32 me = a.rec[i];
    float s1 = 0, s2 = 0, s3 = 0;
    const int cnt = a.L.cnt[i];
    const int32_t* __restrict__ row = a.L.ids + sl_index(i, 0, a.L.cap);
    constexpr int U = 8;
    for (int l0 = 0; l0 < cnt; l0 += U) {
        int jj[U];
        Grad32 rv[U];
        load_chunk8(row, l0 >> 3, jj);
#pragma unroll
        for (int u = 0; u < U; ++u) jj[u] = (l0 + u < cnt) ? jj[u] : i;
#pragma unroll
        for (int u = 0; u < U; ++u) rv[u] = a.rec[jj[u]];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const Grad32 rj = rv[u];
            const float dx = qdiff(rj.x, me.x, a.q.sx), dy = qdiff(rj.y, me.y, a.q.sy),
                        dz = qdiff(rj.z, me.z, a.q.sz);
            const float D = sqrtf(dx * dx + dy * dy + dz * dz);
            const float wv = wfun32(D, a.ktype, a.re);
            const float dp = (a.kh == 1) ? (rj.pn + me.pn - (me.ph + rj.ph)) : (rj.pn - me.ph);
            const float g = wv / D / D;
            // D == 0: two particles within one position quantum (2^-32 of the
            // box) coincide here; FP64 would see a finite, tiny D. The pair is
            // skipped, as the predictor does, instead of producing 0 * inf.
            const bool use = (jj[u] != i) && (D > 0.0f);
            s1 = use ? s1 + dp * dx * g : s1;
            s2 = use ? s2 + dp * dy * g : s2;
            s3 = use ? s3 + dp * dz * g : s3;
        }
    }
    const float g = a.relaxp * (a.gnum * a.dt / a.n0 / (float)a.rho[i]);
    float un = (float)a.us[i] - g * s1;
    float vn = (float)a.vs[i] - g * s2;
    const float cf = 2.0f;
    if (fabsf(un) > cf * a.vmax) un = a.vmax;                  // damper 345-351
    if (vn > cf * a.vmax) vn = cf * a.vmax;
    if (vn < -cf * a.vmax) vn = -cf * a.vmax;
    a.un[i] = (double)un;
    a.vn[i] = (double)vn;
    if (a.dim == 3) {
        float wn = (float)a.ws[i] - g * s3;
        if (fabsf(wn) > cf * a.vmax) wn = a.vmax;
        a.wn[i] = (double)wn;
    }
}

int st_pressgrad32(vox_ctx* c)
{
    ProfScope ps(c, KC_PGRAD);
    const vox_params& p = c->p;
    Grad32Args a;
    a.first = c->gp + c->wp; a.n = c->tp; a.dim = p.dim; a.ktype = p.ktype;
    a.q = make_quant(p);
    a.kh = p.kh_correction; a.re = (float)p.re_s; a.gnum = p.coef2d ? 2.0f : (float)p.dim;
    a.dt = (float)p.dt; a.n0 = (float)c->n0s; a.relaxp = (float)p.relaxp; a.vmax = (float)p.vmax;
    a.L = view(c, VOX_LIST_S);
    a.x = c->f[VOX_F_XSTAR]; a.y = c->f[VOX_F_YSTAR]; a.z = c->f[VOX_F_ZSTAR];
    a.pn = c->f[VOX_F_PNEW]; a.ph = c->f[VOX_F_PHAT]; a.rho = c->f[VOX_F_RHO];
    a.us = c->f[VOX_F_USTAR]; a.vs = c->f[VOX_F_VSTAR]; a.ws = c->f[VOX_F_WSTAR];
    a.un = c->f[VOX_F_UNEW]; a.vn = c->f[VOX_F_VNEW]; a.wn = c->f[VOX_F_WNEW];
    a.rec = reinterpret_cast<Grad32*>(c->pred_rec);          // 64-byte slots hold 24-byte records
    k_pack_grad32<<<grid_for(c->tp), kBlock, 0, c->stream>>>(a);
    const int m = c->tp - a.first;
    if (m > 0) k_pressgrad32<<<grid_for(m), kBlock, 0, c->stream>>>(a);
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

}  // namespace vox

namespace vox {

// ---- PHATCALC (functions.cu:207-224) ---------------------------------------------------------
// The minimum of the FP32-rounded pressures: rounding is monotone, so this is
// the FP64 p-hat rounded to FP32 (1e15, the reference's start value, when the
// row is empty).
__global__ void k_pack_p32(int n, const double* __restrict__ pn, float* __restrict__ p32)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p32[i] = (float)pn[i];
}

__global__ void k_phat32(int n, ListView L, const float* __restrict__ p32, double* __restrict__ ph)
{
    const int i = xcd_tid();
    if (i >= n) return;
    float mn = 3.0e38f;
    const int cnt = L.cnt[i];
    const int32_t* __restrict__ rowp = L.ids + sl_index(i, 0, L.cap);
    for (int g = 0; (g << 3) < cnt; ++g) {
        int32_t jc[8];
        float v[8];
        load_chunk8(rowp, g, jc);
#pragma unroll
        for (int e = 0; e < 8; ++e) jc[e] = ((g << 3) + e < cnt) ? jc[e] : i;
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = p32[jc[e]];
#pragma unroll
        for (int e = 0; e < 8; ++e) mn = ((g << 3) + e < cnt && v[e] < mn) ? v[e] : mn;
    }
    ph[i] = cnt > 0 ? (double)mn : 999999999999999.0;
}

int st_phat32(vox_ctx* c)
{
    ProfScope ps(c, KC_PHAT);
    float* p32 = reinterpret_cast<float*>(c->scratch_d);      // tp floats in tp doubles
    k_pack_p32<<<grid_for(c->tp), kBlock, 0, c->stream>>>(c->tp, c->f[VOX_F_PNEW], p32);
    k_phat32<<<grid_for(c->tp), kBlock, 0, c->stream>>>(c->tp, view(c, VOX_LIST_S), p32,
                                                        c->f[VOX_F_PHAT]);
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

// ---- strain rate over list S (SPS functions.cu:1743-1788, VISCOSITY 2096-2140) --------------
struct Strain32 { float s11, s12, s22, s13, s23, s33; };

struct Vel32Args {
    int n, first, dim, ktype;
    Quant q;
    float re, n0, dmin, csdl2;
    bool use_dmin;
    ListView L;
    const double *x, *y, *z, *u, *v, *w;
    Pred32* rec;
};

__global__ void k_pack_vel32(Vel32Args a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    Pred32 r;
    r.x = qpos(a.x[i], a.q.ox, a.q.ix);
    r.y = qpos(a.y[i], a.q.oy, a.q.iy);
    r.z = (a.dim == 3) ? qpos(a.z[i], a.q.oz, a.q.iz) : 0u;
    r.u = (float)a.u[i]; r.v = (float)a.v[i]; r.w = (a.dim == 3) ? (float)a.w[i] : 0.0f;
    r.pn = 0.0f; r.meu = 0.0f;
    a.rec[i] = r;
}

// Velocity-gradient sums of row i in FP32; the same terms and row order as
// the FP64 strain_row, branch-free (8 records in flight).
__device__ __forceinline__ Strain32 strain_row32(int i, const Vel32Args& a)
{
    const Pred32 me = a.rec[i];
    float s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0, s6 = 0, s7 = 0, s8 = 0, s9 = 0;
    const int cnt = a.L.cnt[i];
    const int32_t* __restrict__ row = a.L.ids + sl_index(i, 0, a.L.cap);
    for (int l0 = 0; l0 < cnt; l0 += 8) {
        int jj[8];
        Pred32 rv[8];
        load_chunk8(row, l0 >> 3, jj);
#pragma unroll
        for (int u = 0; u < 8; ++u) jj[u] = (l0 + u < cnt) ? jj[u] : i;
#pragma unroll
        for (int u = 0; u < 8; ++u) rv[u] = a.rec[jj[u]];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const Pred32 rj = rv[u];
            const float dx = qdiff(rj.x, me.x, a.q.sx), dy = qdiff(rj.y, me.y, a.q.sy),
                        dz = qdiff(rj.z, me.z, a.q.sz);
            const float d = sqrtf(dx * dx + dy * dy + dz * dz);
            const bool use = (jj[u] != i) && (d > 0.0f) && (!a.use_dmin || d > a.dmin);
            const float g = wfun32(d, a.ktype, a.re) / d / d;
            const float du = rj.u - me.u, dv = rj.v - me.v, dw = rj.w - me.w;
            s1 = use ? s1 + du * dx * g : s1;
            s2 = use ? s2 + dv * dy * g : s2;
            s3 = use ? s3 + du * dy * g : s3;
            s4 = use ? s4 + dv * dx * g : s4;
            s5 = use ? s5 + dw * dz * g : s5;
            s6 = use ? s6 + du * dz * g : s6;
            s7 = use ? s7 + dv * dz * g : s7;
            s8 = use ? s8 + dw * dx * g : s8;
            s9 = use ? s9 + dw * dy * g : s9;
        }
    }
    const float c2 = 2.0f / a.n0;
    const float Uxx = c2 * s1, Uyy = c2 * s2, Uxy = c2 * s3, Uyx = c2 * s4;
    Strain32 S;
    S.s12 = 0.5f * (Uxy + Uyx);
    S.s11 = 0.5f * (Uxx + Uxx);
    S.s22 = 0.5f * (Uyy + Uyy);
    S.s13 = S.s23 = S.s33 = 0.0f;
    if (a.dim == 3) {
        const float Uzz = c2 * s5, Uxz = c2 * s6, Uyz = c2 * s7, Uzx = c2 * s8, Uzy = c2 * s9;
        S.s33 = 0.5f * (Uzz + Uzz);
        S.s13 = 0.5f * (Uxz + Uzx);
        S.s23 = 0.5f * (Uyz + Uzy);
    }
    return S;
}

// ---- SPS (functions.cu:1719-1841) --------------------------------------------------------------
struct Sps32Args {
    Vel32Args v;
    double *s11, *s12, *s22, *s13, *s23, *s33, *neut, *t1, *t2, *t3;
    Pos32* nrec;                 // positions + NEUt, for the second pass
};

__global__ __launch_bounds__(kBlock) void k_sps1_32(Sps32Args a)
{
    const int i = a.v.first + xcd_tid();
    if (i >= a.v.n) return;
    const Strain32 S = strain_row32(i, a.v);
    a.s11[i] = S.s11; a.s12[i] = S.s12; a.s22[i] = S.s22;
    a.s13[i] = S.s13; a.s23[i] = S.s23; a.s33[i] = S.s33;
    float s = sqrtf(2 * S.s11 * S.s11) + 2 * sqrtf(2 * S.s12 * S.s12) + sqrtf(2 * S.s22 * S.s22);
    if (a.v.dim == 3)
        s = s + sqrtf(2 * S.s33 * S.s33) + 2 * sqrtf(2 * S.s13 * S.s13) + 2 * sqrtf(2 * S.s23 * S.s23);
    a.neut[i] = (double)(a.v.csdl2 * s);
}

__global__ void k_pack_neut32(Sps32Args a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.v.n) return;
    const Pred32 r = a.v.rec[i];
    a.nrec[i] = Pos32{r.x, r.y, r.z, (float)a.neut[i]};
}

__global__ __launch_bounds__(kBlock) void k_sps2_32(Sps32Args a)
{
    const int i = a.v.first + xcd_tid();
    if (i >= a.v.n) return;
    const Pos32 me = a.nrec[i];
    float s1 = 0, s2 = 0, s3 = 0;
    const int cnt = a.v.L.cnt[i];
    const int32_t* __restrict__ row = a.v.L.ids + sl_index(i, 0, a.v.L.cap);
    for (int l0 = 0; l0 < cnt; l0 += 8) {
        int jj[8];
        Pos32 rv[8];
        load_chunk8(row, l0 >> 3, jj);
#pragma unroll
        for (int u = 0; u < 8; ++u) jj[u] = (l0 + u < cnt) ? jj[u] : i;
#pragma unroll
        for (int u = 0; u < 8; ++u) rv[u] = a.nrec[jj[u]];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const Pos32 rj = rv[u];
            const float dx = qdiff(rj.x, me.x, a.v.q.sx), dy = qdiff(rj.y, me.y, a.v.q.sy),
                        dz = qdiff(rj.z, me.z, a.v.q.sz);
            const float d = sqrtf(dx * dx + dy * dy + dz * dz);
            const bool use = (jj[u] != i) && (d > 0.0f) && (d > a.v.dmin);
            const float g = wfun32(d, a.v.ktype, a.v.re) / d / d;
            const float dn = rj.f - me.f;
            s1 = use ? s1 + dn * dx * g : s1;
            s2 = use ? s2 + dn * dy * g : s2;
            s3 = use ? s3 + dn * dz * g : s3;
        }
    }
    const float c2 = 2.0f / a.v.n0;
    const float s11 = (float)a.s11[i], s12 = (float)a.s12[i], s22 = (float)a.s22[i];
    if (a.v.dim == 3) {
        const float s13 = (float)a.s13[i], s23 = (float)a.s23[i], s33 = (float)a.s33[i];
        a.t1[i] = c2 * (2 * s11 * s1 + 2 * s12 * s2 + 2 * s13 * s3);
        a.t2[i] = c2 * (2 * s12 * s1 + 2 * s22 * s2 + 2 * s23 * s3);
        a.t3[i] = c2 * (2 * s13 * s1 + 2 * s23 * s2 + 2 * s33 * s3);
    } else {
        a.t1[i] = c2 * (2 * s11 * s1 + 2 * s12 * s2);
        a.t2[i] = c2 * (2 * s12 * s1 + 2 * s22 * s2);
    }
}

static Vel32Args vel_args(vox_ctx* c, bool use_dmin)
{
    const vox_params& p = c->p;
    Vel32Args v;
    v.n = c->tp; v.first = c->gp; v.dim = p.dim; v.ktype = p.ktype; v.q = make_quant(p);
    v.re = (float)p.re_s; v.n0 = (float)c->n0s; v.dmin = (float)(p.coll / 10);
    v.csdl2 = (float)((p.cs * p.dl) * (p.cs * p.dl));
    v.use_dmin = use_dmin;
    v.L = view(c, VOX_LIST_S);
    v.x = c->f[VOX_F_X]; v.y = c->f[VOX_F_Y]; v.z = c->f[VOX_F_Z];
    v.u = c->f[VOX_F_UNEW]; v.v = c->f[VOX_F_VNEW]; v.w = c->f[VOX_F_WNEW];
    v.rec = reinterpret_cast<Pred32*>(c->pred_rec);          // 64-byte slots hold 32-byte records
    return v;
}

int st_sps32(vox_ctx* c)
{
    ProfScope ps(c, KC_SPS);
    Sps32Args a;
    a.v = vel_args(c, true);
    a.s11 = c->sps_s[0]; a.s12 = c->sps_s[1]; a.s22 = c->sps_s[2];
    a.s13 = c->sps_s[3]; a.s23 = c->sps_s[4]; a.s33 = c->sps_s[5];
    a.neut = c->f[VOX_F_NEUT];
    a.t1 = c->f[VOX_F_TURB1]; a.t2 = c->f[VOX_F_TURB2]; a.t3 = c->f[VOX_F_TURB3];
    a.nrec = reinterpret_cast<Pos32*>(c->pos_rec);           // 32-byte slots hold 16-byte records
    const int m = c->tp - c->gp;
    if (m <= 0) return VOX_OK;
    k_pack_vel32<<<grid_for(c->tp), kBlock, 0, c->stream>>>(a.v);
    k_sps1_32<<<grid_for(m), kBlock, 0, c->stream>>>(a);
    k_pack_neut32<<<grid_for(c->tp), kBlock, 0, c->stream>>>(a);
    k_sps2_32<<<grid_for(m), kBlock, 0, c->stream>>>(a);
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

// ---- Herschel-Bulkley VISCOSITY (functions.cu:2053-2184) ---------------------------------------
struct Hb32Args {
    Vel32Args v;
    const int32_t* pt;
    const double* phat;
    double* meu;
    double m1;
    float cohes, phi_param, meu0, npl;
};

__global__ __launch_bounds__(kBlock) void k_visc_hb32(Hb32Args a)
{
    const int i = xcd_tid();
    if (i >= a.v.n) return;
    if (a.pt[i] == 1) { a.meu[i] = a.m1; return; }
    const Strain32 S = strain_row32(i, a.v);
    // phi reads C[I] with the caller's I == 0 (main.cu:211): phi = 0 (2154-2155)
    const float cz = 0.0f;
    float phi = (cz - 0.25f) * a.phi_param / (1 - 0.25f);
    if (cz < 0.25f) phi = 0;
    const float ys = a.cohes * cosf(phi) + (float)a.phat[i] * sinf(phi);
    float II;
    if (a.v.dim == 3)
        II = fabsf((S.s11 * S.s22 * S.s33 + S.s12 * S.s23 * S.s13 + S.s13 * S.s12 * S.s23) -
                   (S.s13 * S.s22 * S.s13 + S.s11 * S.s23 * S.s23 + S.s12 * S.s12 * S.s33));
    else
        II = fabsf(S.s11 * S.s22 - S.s12 * S.s12);
    float m = ys / 2 / sqrtf(II) + a.meu0 * powf(4 * II, (a.npl - 1) / 2);
    if (II == 0 || m > 200) m = 200;
    a.meu[i] = (double)m;
}

int st_visc_hb32(vox_ctx* c)
{
    const vox_params& p = c->p;
    Hb32Args a;
    a.v = vel_args(c, false);
    a.v.first = 0;
    a.pt = c->ptype; a.phat = c->f[VOX_F_PHAT]; a.meu = c->f[VOX_F_MEU];
    a.m1 = p.nu1 * p.rho1; a.cohes = (float)p.cohes; a.phi_param = (float)p.phi;
    a.meu0 = (float)p.meu0; a.npl = (float)p.n_pl;
    k_pack_vel32<<<grid_for(c->tp), kBlock, 0, c->stream>>>(a.v);
    k_visc_hb32<<<grid_for(c->tp), kBlock, 0, c->stream>>>(a);
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

// ---- V_FRACTION, method 2 (functions.cu:2604-2634) ----------------------------------------------
__global__ void k_pack_type32(int n, int dim, Quant q, const double* __restrict__ x,
                              const double* __restrict__ y, const double* __restrict__ z,
                              const int32_t* __restrict__ pt, Pos32* __restrict__ rec)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    rec[i] = Pos32{qpos(x[i], q.ox, q.ix), qpos(y[i], q.oy, q.iy),
                   dim == 3 ? qpos(z[i], q.oz, q.iz) : 0u, (float)pt[i]};
}

__global__ __launch_bounds__(kBlock) void k_vfraction32(int n, int ktype, float re, Quant q,
                                                        ListView L, const Pos32* __restrict__ rec,
                                                        double* __restrict__ C)
{
    const int i = xcd_tid();
    if (i >= n) return;
    const Pos32 me = rec[i];
    float s1 = 0, s2 = 0;
    const int cnt = L.cnt[i];
    const int32_t* __restrict__ row = L.ids + sl_index(i, 0, L.cap);
    for (int l0 = 0; l0 < cnt; l0 += 8) {
        int jj[8];
        Pos32 rv[8];
        load_chunk8(row, l0 >> 3, jj);
#pragma unroll
        for (int u = 0; u < 8; ++u) jj[u] = (l0 + u < cnt) ? jj[u] : i;
#pragma unroll
        for (int u = 0; u < 8; ++u) rv[u] = rec[jj[u]];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const Pos32 rj = rv[u];
            const float dx = qdiff(rj.x, me.x, q.sx), dy = qdiff(rj.y, me.y, q.sy),
                        dz = qdiff(rj.z, me.z, q.sz);
            const float wv = wfun32(sqrtf(dx * dx + dy * dy + dz * dz), ktype, re);
            const bool use = (jj[u] != i) && (rj.f > 0.0f);
            s1 = use ? s1 + wv : s1;
            s2 = (use && rj.f == 2.0f) ? s2 + wv : s2;
        }
    }
    C[i] = (s1 == 0.0f) ? 0.0 : (double)(s2 / s1);
}

int st_vfraction32(vox_ctx* c)
{
    const vox_params& p = c->p;
    Pos32* rec = reinterpret_cast<Pos32*>(c->pos_rec);
    const Quant q = make_quant(p);
    k_pack_type32<<<grid_for(c->tp), kBlock, 0, c->stream>>>(c->tp, p.dim, q, c->f[VOX_F_X],
                                                             c->f[VOX_F_Y], c->f[VOX_F_Z],
                                                             c->ptype, rec);
    k_vfraction32<<<grid_for(c->tp), kBlock, 0, c->stream>>>(c->tp, p.ktype, (float)p.re_s, q,
                                                             view(c, VOX_LIST_S), rec,
                                                             c->f[VOX_F_C]);
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

}  // namespace vox
