/*
 * oracle.c — CPU restatement of the reference's MPS time step (test
 * infrastructure only; see oracle.h). FP64, 0-based arrays. Per-particle
 * loops run on OpenMP threads with results independent of the thread count
 * (each particle's sums in order; max / count reductions; the CG's dot
 * products over fixed index ranges). Collision and the srcopt chain are
 * sequential, as the reference defines them.
 *
 * Arithmetic is written in the reference's operation order with plain
 * IEEE double operations (no FMA: build without -mfma / with
 * -ffp-contract=off), so a device kernel that follows the same order and is
 * built without contraction produces bit-identical per-particle sums.
 * Deviations from the reference text, each at the last-ulp level or below:
 *   - pow(d, 2.0) is written d*d (what the reference's compilers emit);
 *   - pow(a, 0.5) in SPS is written sqrt(a);
 *   - pow(1-q, 3) (ktype 6) is written (1-q)*(1-q)*(1-q).
 * Parity-affecting design choices (documented with the build):
 *   - the PPE is solved by Jacobi-PCG to ||r||/||b|| <= cg_tol from a zero start
 *     (reference: IC(0)-CG, functions.cu:506-716, stop rule 666-687);
 *   - two radii: list S (re_s) feeds PND, p-hat, gradient, SPS, viscosity,
 *     volume fraction and collision; list L (re_l) feeds the Laplacian terms.
 *     With re_s == re_l this is exactly the reference's single list.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct orc_list {
    double   re;
    int      cap;      /* row stride of ids                          */
    int      maxcnt;
    int32_t* cnt;      /* tp                                          */
    int32_t* ids;      /* tp * cap, row-major                         */
    int32_t* cell;     /* tp, 0-based cell number                     */
} orc_list;

struct orc_ctx {
    vox_params p;
    int tp, gp, wp, fp;
    double*  f[VOX_F_COUNT];
    int32_t* ptype;
    int32_t* bcon;
    orc_list list[2];
    int      two_lists;
    double   n0s, n0l;
    int      initialised;
    /* PPE in the reference's slot layout, aligned with list L rows */
    double*  aval;     /* tp * cap_l */
    double*  adiag;    /* tp */
    int32_t* uidx;     /* particle -> unknown index or -1 */
    int      nu;
    vox_status st;
};

static int is_int_field(int f) { return f == VOX_F_PTYPE || f == VOX_F_BCON; }

/* ---- weight function and distances: functions.cu:32-89 ------------------ */
static inline double wfun(double r, int ktype, double re)
{
    double q = r / re;
    if (ktype == 2) return (q < 1.0) ? (1 / q) - 1 : 0.0;            /* functions.cu:42-45 */
    if (ktype == 6) return (q <= 1.0) ? (1 - q) * (1 - q) * (1 - q) : 0.0; /* 64-67 */
    if (ktype == 1) {                                                 /* 37-41 */
        if (q < 0.5) return 2.0 - 4.0 * (q * q);
        if (q <= 1.0) return (2 * q - 2) * (2 * q - 2);
        return 0.0;
    }
    if (ktype == 5) return (q <= 1.0) ? 1.5 * log(1 / (q + 0.000001)) : 0.0; /* 59-62 */
    return 0.0;
}

double orc_weight(double r, int ktype, int dim, double re)
{
    (void)dim;
    return wfun(r, ktype, re);
}

static inline double dist(const orc_ctx* c, const double* x, const double* y, const double* z,
                          int i, int j)
{
    double dx = x[j] - x[i];
    double dy = y[j] - y[i];
    if (c->p.dim == 3) {
        double dz = z[j] - z[i];
        return sqrt(dx * dx + dy * dy + dz * dz);   /* dist3d, functions.cu:84-89 */
    }
    return sqrt(dx * dx + dy * dy);                  /* dist2d, functions.cu:77-82 */
}

/* ---- context --------------------------------------------------------------- */
int orc_create(orc_ctx** out, const vox_params* p, int tp, int gp, int wp, int fp)
{
    if (!out || !p || tp <= 0 || gp < 0 || wp < 0 || fp < 0 || gp + wp + fp != tp)
        return VOX_ERR_ARG;
    if (p->dim != 2 && p->dim != 3) return VOX_ERR_ARG;
    if (p->ref_gpu && p->dim != 2) return VOX_ERR_ARG;
    orc_ctx* c = (orc_ctx*)calloc(1, sizeof(orc_ctx));
    if (!c) return VOX_ERR_ARG;
    c->p = *p;
    if (c->p.pmax <= 0) c->p.pmax = (double)tp * 4.50;         /* main.cu:548 */
    c->tp = tp; c->gp = gp; c->wp = wp; c->fp = fp;
    for (int f = 0; f < VOX_F_COUNT; ++f)
        if (!is_int_field(f)) c->f[f] = (double*)calloc((size_t)tp, sizeof(double));
    c->ptype = (int32_t*)calloc((size_t)tp, sizeof(int32_t));
    c->bcon = (int32_t*)calloc((size_t)tp, sizeof(int32_t));
    c->adiag = (double*)calloc((size_t)tp, sizeof(double));
    c->uidx = (int32_t*)calloc((size_t)tp, sizeof(int32_t));
    c->two_lists = (p->re_l != p->re_s);
    for (int w = 0; w < 2; ++w) {
        c->list[w].cnt = (int32_t*)calloc((size_t)tp, sizeof(int32_t));
        c->list[w].cell = (int32_t*)calloc((size_t)tp, sizeof(int32_t));
    }
    *out = c;
    return VOX_OK;
}

void orc_destroy(orc_ctx* c)
{
    if (!c) return;
    for (int f = 0; f < VOX_F_COUNT; ++f) free(c->f[f]);
    free(c->ptype); free(c->bcon); free(c->adiag); free(c->uidx); free(c->aval);
    for (int w = 0; w < 2; ++w) { free(c->list[w].cnt); free(c->list[w].ids); free(c->list[w].cell); }
    free(c);
}

int orc_set_params(orc_ctx* c, const vox_params* p)
{
    if (!c || !p) return VOX_ERR_ARG;
    if (p->dim != c->p.dim || p->re_s != c->p.re_s || p->re_l != c->p.re_l) return VOX_ERR_ARG;
    double pmax = c->p.pmax;
    c->p = *p;
    if (c->p.pmax <= 0) c->p.pmax = pmax;
    return VOX_OK;
}

int orc_get_params(orc_ctx* c, vox_params* p)
{
    if (!c || !p) return VOX_ERR_ARG;
    *p = c->p;
    return VOX_OK;
}

int orc_upload(orc_ctx* c, int field, const void* host)
{
    if (!c || !host || field < 0 || field >= VOX_F_COUNT) return VOX_ERR_ARG;
    if (field == VOX_F_PTYPE) memcpy(c->ptype, host, sizeof(int32_t) * c->tp);
    else if (field == VOX_F_BCON) memcpy(c->bcon, host, sizeof(int32_t) * c->tp);
    else memcpy(c->f[field], host, sizeof(double) * c->tp);
    return VOX_OK;
}

int orc_download(orc_ctx* c, int field, void* host)
{
    if (!c || !host || field < 0 || field >= VOX_F_COUNT) return VOX_ERR_ARG;
    if (field == VOX_F_PTYPE) memcpy(host, c->ptype, sizeof(int32_t) * c->tp);
    else if (field == VOX_F_BCON) memcpy(host, c->bcon, sizeof(int32_t) * c->tp);
    else memcpy(host, c->f[field], sizeof(double) * c->tp);
    return VOX_OK;
}

/* ---- density: main.cu:970-977 ------------------------------------------- */
int orc_density(orc_ctx* c)
{
    double* rho = c->f[VOX_F_RHO];
    for (int i = 0; i < c->tp; ++i) rho[i] = (c->ptype[i] == 2) ? c->p.rho2 : c->p.rho1;
    return VOX_OK;
}

/* ---- NEIGHBOR: functions.cu:2788-2929 ------------------------------------- */
typedef struct grid_t { int ncx, ncy, ncz; double hx, hz; } grid_t;

static grid_t make_grid(const vox_params* p, double re)
{
    grid_t g;
    g.hx = re;                                                  /* 2790-2791 */
    g.hz = re + p->delta;                                       /* 2796 */
    if (p->ref_gpu) g.hx = re + p->delta;    /* CUDA path: neighbour_parallel.cu:46-47 */
    g.ncx = (int)((p->xmax - p->xmin) / g.hx) + 1;
    g.ncy = (int)((p->ymax - p->ymin) / g.hx) + 1;
    g.ncz = (p->dim == 3) ? (int)((p->zmax - p->zmin) / g.hz) + 1 : 1;
    return g;
}

/* 1-based cell coordinates of functions.cu:2822-2826; returns 0 if outside. */
static int cell_coords(const vox_params* p, const grid_t* g, double x, double y, double z,
                       int* ic, int* jc, int* kc)
{
    if (!isfinite(x) || !isfinite(y) || (p->dim == 3 && !isfinite(z))) return 0;
    double ax = (x - p->xmin) / g->hx, ay = (y - p->ymin) / g->hx;
    if (ax > 2e9 || ay > 2e9 || ax < -2e9 || ay < -2e9) return 0;
    *ic = (int)ax + 1;
    *jc = (int)ay + 1;
    *kc = 1;
    if (p->dim == 3) {
        double az = (z - p->zmin) / g->hz;
        if (az > 2e9 || az < -2e9) return 0;
        *kc = (int)az + 1;
    }
    if (*ic < 1 || *ic > g->ncx || *jc < 1 || *jc > g->ncy || *kc < 1 || *kc > g->ncz) return 0;
    return 1;
}

/* accept rule: 3D d <= re+DELTA with self (2886); 2D I!=J and d < re (2907-2910);
 * the CUDA path's 2D search: R <= re+DELTA with self (neighbour_parallel.cu:180-181) */
static inline int accept(const orc_ctx* c, double re, const double* x, const double* y,
                         const double* z, int i, int j)
{
    if (c->p.dim == 3 || c->p.ref_gpu) return dist(c, x, y, z, i, j) <= re + c->p.delta;
    if (i == j) return 0;
    return dist(c, x, y, z, i, j) < re;
}

static int build_list(orc_ctx* c, int w, double re)
{
    const vox_params* p = &c->p;
    const double *x = c->f[VOX_F_X], *y = c->f[VOX_F_Y], *z = c->f[VOX_F_Z];
    orc_list* L = &c->list[w];
    grid_t g = make_grid(p, re);
    long tnc = (long)g.ncx * g.ncy * g.ncz;
    int tp = c->tp;
    L->re = re;
    /* cell numbers: Cnum = icell + (jcell-1)ncx + (kcell-1)ncx ncy, 2826-2834 */
    for (int i = 0; i < tp; ++i) {
        int ic, jc, kc;
        if (!cell_coords(p, &g, x[i], y[i], z[i], &ic, &jc, &kc)) return VOX_ERR_DOMAIN;
        L->cell[i] = (ic - 1) + (jc - 1) * g.ncx + (kc - 1) * g.ncx * g.ncy;
    }
    /* Binning (2820-2852) inserts particles in id order and keeps each cell's
     * run ascending by id: a stable counting sort by cell. */
    int32_t* start = (int32_t*)calloc((size_t)tnc + 1, sizeof(int32_t));
    int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (size_t)tp);
    int32_t* fill = (int32_t*)calloc((size_t)tnc, sizeof(int32_t));
    if (!start || !order || !fill) { free(start); free(order); free(fill); return VOX_ERR_ARG; }
    for (int i = 0; i < tp; ++i) start[L->cell[i] + 1]++;
    for (long k = 0; k < tnc; ++k) start[k + 1] += start[k];
    for (int i = 0; i < tp; ++i) { int cc = L->cell[i]; order[start[cc] + fill[cc]++] = i; }

    /* search (2854-2921): stencil x-offset outer, y middle, z inner, clipped at
     * the grid edge, ascending id inside each cell. Two passes: count, fill. */
    for (int pass = 0; pass < 2; ++pass) {
        int maxc = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(max : maxc)
        for (int i = 0; i < tp; ++i) {
            int ic, jc, kc;
            cell_coords(p, &g, x[i], y[i], z[i], &ic, &jc, &kc);
            int m1 = (ic == 1) ? 0 : -1, m2 = (ic == g.ncx) ? 0 : 1;
            int m3 = (jc == 1) ? 0 : -1, m4 = (jc == g.ncy) ? 0 : 1;
            int m5 = 0, m6 = 0;
            if (p->dim == 3) { m5 = (kc == 1) ? 0 : -1; m6 = (kc == g.ncz) ? 0 : 1; }
            int k = 0;
            for (int row = m1; row <= m2; ++row)
                for (int col = m3; col <= m4; ++col)
                    for (int dep = m5; dep <= m6; ++dep) {
                        long cn = (ic - 1 + row) + (long)(jc - 1 + col) * g.ncx +
                                  (long)(kc - 1 + dep) * g.ncx * g.ncy;
                        for (int jj = start[cn]; jj < start[cn + 1]; ++jj) {
                            int j = order[jj];
                            if (accept(c, re, x, y, z, i, j)) {
                                if (pass == 1) L->ids[(size_t)i * L->cap + k] = j;
                                ++k;
                            }
                        }
                    }
            L->cnt[i] = k;                     /* neighb[I][1] = k+1, 2919-2920 */
            if (k > maxc) maxc = k;
        }
        if (pass == 0) {
            L->maxcnt = maxc;
            int cap = maxc > 0 ? maxc : 1;
            if (cap > L->cap || !L->ids) {
                free(L->ids);
                L->ids = (int32_t*)malloc(sizeof(int32_t) * (size_t)tp * cap);
                if (!L->ids) { free(start); free(order); free(fill); return VOX_ERR_ARG; }
                L->cap = cap;
            }
        }
    }
    free(start); free(order); free(fill);
    return VOX_OK;
}

int orc_neighbor(orc_ctx* c)
{
    int rc = build_list(c, VOX_LIST_S, c->p.re_s);
    if (rc) return rc;
    if (c->two_lists) return build_list(c, VOX_LIST_L, c->p.re_l);
    return VOX_OK;
}

static const orc_list* LS(const orc_ctx* c) { return &c->list[VOX_LIST_S]; }
static const orc_list* LL(const orc_ctx* c) { return c->two_lists ? &c->list[VOX_LIST_L] : &c->list[VOX_LIST_S]; }
#define ROW(L, i) ((L)->ids + (size_t)(i) * (L)->cap)

int orc_neighbor_bruteforce(const vox_params* p, double re, int tp, const double* x,
                            const double* y, const double* z, int32_t* counts,
                            int32_t* ids, int stride)
{
    orc_ctx tmp;
    memset(&tmp, 0, sizeof(tmp));
    tmp.p = *p;
    grid_t g = make_grid(p, re);
    int32_t* key = (int32_t*)malloc(sizeof(int32_t) * (size_t)tp);
    int32_t* jbuf = (int32_t*)malloc(sizeof(int32_t) * (size_t)tp);
    int* ci = (int*)malloc(sizeof(int) * 3 * (size_t)tp);
    if (!key || !jbuf || !ci) { free(key); free(jbuf); free(ci); return VOX_ERR_ARG; }
    for (int i = 0; i < tp; ++i)
        if (!cell_coords(p, &g, x[i], y[i], p->dim == 3 ? z[i] : 0.0, &ci[3 * i], &ci[3 * i + 1], &ci[3 * i + 2])) {
            free(key); free(jbuf); free(ci); return VOX_ERR_DOMAIN;
        }
    int rc = VOX_OK;
    for (int i = 0; i < tp; ++i) {
        int n = 0;
        for (int j = 0; j < tp; ++j) {
            int ox = ci[3 * j] - ci[3 * i], oy = ci[3 * j + 1] - ci[3 * i + 1], oz = ci[3 * j + 2] - ci[3 * i + 2];
            if (ox < -1 || ox > 1 || oy < -1 || oy > 1 || oz < -1 || oz > 1) continue;
            if (!accept(&tmp, re, x, y, z, i, j)) continue;
            key[n] = ((ox + 1) * 3 + (oy + 1)) * 3 + (oz + 1);
            jbuf[n] = j;
            ++n;
        }
        /* order by stencil offset, then id (jbuf is ascending already: stable insertion) */
        for (int a = 1; a < n; ++a) {
            int kk = key[a], jv = jbuf[a], b = a - 1;
            while (b >= 0 && key[b] > kk) { key[b + 1] = key[b]; jbuf[b + 1] = jbuf[b]; --b; }
            key[b + 1] = kk; jbuf[b + 1] = jv;
        }
        counts[i] = n;
        if (n > stride) { rc = VOX_ERR_OVERFLOW; n = stride; }
        for (int a = 0; a < n; ++a) ids[(size_t)i * stride + a] = jbuf[a];
    }
    free(key); free(jbuf); free(ci);
    return rc;
}

/* ---- PNUM: functions.cu:94-123 -------------------------------------------- */
static double pnd_row(const orc_ctx* c, const orc_list* L, double re, const double* x,
                      const double* y, const double* z, int i)
{
    double sum = 0.0;
    const int32_t* r = ROW(L, i);
    for (int l = 0; l < L->cnt[i]; ++l) {
        int j = r[l];
        double d = dist(c, x, y, z, i, j);
        if (i != j) sum = sum + wfun(d, c->p.ktype, re);
    }
    if (c->p.ktype != 2) sum = sum * c->p.ncorrection;
    return sum;
}

int orc_pnum(orc_ctx* c, int loopstarted)
{
    const double *x = c->f[VOX_F_X], *y = c->f[VOX_F_Y], *z = c->f[VOX_F_Z];
    double* n = c->f[VOX_F_N];
    double mx = 0.0;                           /* MAX = new double[1]() (main.cu:376) */
#pragma omp parallel for schedule(dynamic, 256) reduction(max : mx)
    for (int i = 0; i < c->tp; ++i) {
        n[i] = pnd_row(c, LS(c), c->p.re_s, x, y, z, i);
        if (!loopstarted) {
            c->f[VOX_F_PNEW][i] = c->f[VOX_F_P][i];   /* 115-119 */
            if (n[i] > mx) mx = n[i];
        }
    }
    if (!loopstarted) {
        c->n0s = mx;                                  /* main.cu:1011 */
        if (c->two_lists) {
            double ml = 0.0;
#pragma omp parallel for schedule(dynamic, 256) reduction(max : ml)
            for (int i = 0; i < c->tp; ++i) {
                double v = pnd_row(c, LL(c), c->p.re_l, x, y, z, i);
                if (v > ml) ml = v;
            }
            c->n0l = ml;
        } else {
            c->n0l = mx;
        }
    }
    return VOX_OK;
}

/* ---- PNUMSTAR: functions.cu:157-180 --------------------------------------- */
int orc_pnumstar(orc_ctx* c)
{
    const double *x = c->f[VOX_F_XSTAR], *y = c->f[VOX_F_YSTAR], *z = c->f[VOX_F_ZSTAR];
    double* ns = c->f[VOX_F_NSTAR];
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < c->tp; ++i) {
        double sum = 0.0;
        const int32_t* r = ROW(LS(c), i);
        for (int l = 0; l < LS(c)->cnt[i]; ++l) {
            int j = r[l];
            double d = dist(c, x, y, z, i, j);
            if (i != j) sum = sum + wfun(d, c->p.ktype, c->p.re_s);
        }
        ns[i] = sum;
    }
    return VOX_OK;
}

/* ---- strain rate sums shared by SPS (1738-1790) and VISCOSITY (2095-2150) ---- */
typedef struct strain_t { double s11, s12, s22, s13, s23, s33; } strain_t;

static strain_t strain_row(const orc_ctx* c, int i, double dmin, int use_dmin)
{
    const double *x = c->f[VOX_F_X], *y = c->f[VOX_F_Y], *z = c->f[VOX_F_Z];
    const double *u = c->f[VOX_F_UNEW], *v = c->f[VOX_F_VNEW], *w = c->f[VOX_F_WNEW];
    const orc_list* L = LS(c);
    const int32_t* r = ROW(L, i);
    double re = c->p.re_s;
    double s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0, s6 = 0, s7 = 0, s8 = 0, s9 = 0;
    for (int l = 0; l < L->cnt[i]; ++l) {
        int j = r[l];
        double d = dist(c, x, y, z, i, j);
        if (i != j && (!use_dmin || d > dmin)) {
            double wv = wfun(d, c->p.ktype, re);
            double du = u[j] - u[i], dv = v[j] - v[i];
            double dx = x[j] - x[i], dy = y[j] - y[i];
            s1 = s1 + du * dx * wv / d / d;
            s2 = s2 + dv * dy * wv / d / d;
            s3 = s3 + du * dy * wv / d / d;
            s4 = s4 + dv * dx * wv / d / d;
            if (c->p.dim == 3) {
                double dw = w[j] - w[i], dz = z[j] - z[i];
                s5 = s5 + dw * dz * wv / d / d;
                s6 = s6 + du * dz * wv / d / d;
                s7 = s7 + dv * dz * wv / d / d;
                s8 = s8 + dw * dx * wv / d / d;
                s9 = s9 + dw * dy * wv / d / d;
            }
        }
    }
    double n0 = c->n0s;
    double Uxx = (2.0 / n0) * s1, Uyy = (2.0 / n0) * s2, Uxy = (2.0 / n0) * s3, Uyx = (2.0 / n0) * s4;
    strain_t S;
    S.s12 = 0.5 * (Uxy + Uyx);
    S.s11 = 0.5 * (Uxx + Uxx);
    S.s22 = 0.5 * (Uyy + Uyy);
    S.s13 = S.s23 = S.s33 = 0.0;
    if (c->p.dim == 3) {
        double Uzz = (2.0 / n0) * s5, Uxz = (2.0 / n0) * s6, Uyz = (2.0 / n0) * s7;
        double Uzx = (2.0 / n0) * s8, Uzy = (2.0 / n0) * s9;
        S.s33 = 0.5 * (Uzz + Uzz);
        S.s13 = 0.5 * (Uxz + Uzx);
        S.s23 = 0.5 * (Uyz + Uzy);
    }
    return S;
}

/* ---- SPS: functions.cu:1719-1841 ------------------------------------------ */
int orc_sps(orc_ctx* c)
{
    int tp = c->tp;
    strain_t* S = (strain_t*)calloc((size_t)tp, sizeof(strain_t));
    if (!S) return VOX_ERR_ARG;
    double* neut = c->f[VOX_F_NEUT];
    double csdl = c->p.cs * c->p.dl;
    #pragma omp parallel for schedule(dynamic, 256)
    for (int i = c->gp; i < tp; ++i) {
        S[i] = strain_row(c, i, c->p.coll / 10, 1);
        double s = sqrt(2 * S[i].s11 * S[i].s11) + 2 * sqrt(2 * S[i].s12 * S[i].s12) +
                   sqrt(2 * S[i].s22 * S[i].s22);
        if (c->p.dim == 3)
            s = s + sqrt(2 * S[i].s33 * S[i].s33) + 2 * sqrt(2 * S[i].s13 * S[i].s13) +
                2 * sqrt(2 * S[i].s23 * S[i].s23);
        neut[i] = (csdl * csdl) * s;                     /* 1794 */
    }
    const double *x = c->f[VOX_F_X], *y = c->f[VOX_F_Y], *z = c->f[VOX_F_Z];
    const orc_list* L = LS(c);
    double n0 = c->n0s;
    #pragma omp parallel for schedule(dynamic, 256)
    for (int i = c->gp; i < tp; ++i) {
        double s1 = 0, s2 = 0, s3 = 0;
        const int32_t* r = ROW(L, i);
        for (int l = 0; l < L->cnt[i]; ++l) {
            int j = r[l];
            double d = dist(c, x, y, z, i, j);
            if (i != j && d > c->p.coll / 10) {
                double wv = wfun(d, c->p.ktype, c->p.re_s);
                double dn = neut[j] - neut[i];
                s1 = s1 + dn * (x[j] - x[i]) * wv / d / d;
                s2 = s2 + dn * (y[j] - y[i]) * wv / d / d;
                if (c->p.dim == 3) s3 = s3 + dn * (z[j] - z[i]) * wv / d / d;
            }
        }
        if (c->p.dim == 3) {
            c->f[VOX_F_TURB1][i] = (2.0 / n0) * (2 * S[i].s11 * s1 + 2 * S[i].s12 * s2 + 2 * S[i].s13 * s3);
            c->f[VOX_F_TURB2][i] = (2.0 / n0) * (2 * S[i].s12 * s1 + 2 * S[i].s22 * s2 + 2 * S[i].s23 * s3);
            c->f[VOX_F_TURB3][i] = (2.0 / n0) * (2 * S[i].s13 * s1 + 2 * S[i].s23 * s2 + 2 * S[i].s33 * s3);
        } else {
            c->f[VOX_F_TURB1][i] = (2.0 / n0) * (2 * S[i].s11 * s1 + 2 * S[i].s12 * s2);
            c->f[VOX_F_TURB2][i] = (2.0 / n0) * (2 * S[i].s12 * s1 + 2 * S[i].s22 * s2);
        }
    }
    free(S);
    return VOX_OK;
}

/* ---- V_FRACTION, method 2: functions.cu:2604-2634 -------------------------- */
int orc_vfraction(orc_ctx* c)
{
    const double *x = c->f[VOX_F_X], *y = c->f[VOX_F_Y], *z = c->f[VOX_F_Z];
    const orc_list* L = LS(c);
    #pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < c->tp; ++i) {
        double s1 = 0, s2 = 0;
        const int32_t* r = ROW(L, i);
        for (int l = 0; l < L->cnt[i]; ++l) {
            int j = r[l];
            double d = dist(c, x, y, z, i, j);
            if (i != j && c->ptype[j] > 0) {
                double wv = wfun(d, c->p.ktype, c->p.re_s);
                s1 = s1 + wv;
                if (c->ptype[j] == 2) s2 = s2 + wv;
            }
        }
        double cv = s2 / s1;
        if (s1 == 0) cv = 0;
        c->f[VOX_F_C][i] = cv;
    }
    return VOX_OK;
}

/* ---- VISCOSITY: functions.cu:2053-2184 ------------------------------------- */
int orc_viscosity(orc_ctx* c)
{
    const vox_params* p = &c->p;
    double* meu = c->f[VOX_F_MEU];
    if (p->fluid2_type == 0) {                               /* 2073-2085 */
        for (int i = 0; i < c->tp; ++i) {
            if (c->ptype[i] <= 1) meu[i] = p->nu1 * p->rho1;
            if (c->ptype[i] == 2) meu[i] = p->nu2 * p->rho2;
        }
        return VOX_OK;
    }
    #pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < c->tp; ++i) {                        /* 2087-2170 */
        if (c->ptype[i] == 1) { meu[i] = p->nu1 * p->rho1; continue; }
        strain_t S = strain_row(c, i, 0.0, 0);
        /* phi reads C[I] with the caller's I == 0 (main.cu:211), i.e. C[0] of a
         * 1-based array, which is never written: phi = 0 (2154-2155). */
        double cz = 0.0;
        double phi = (cz - 0.25) * p->phi / (1 - 0.25);
        if (cz < 0.25) phi = 0;
        double ys = p->cohes * cos(phi) + c->f[VOX_F_PHAT][i] * sin(phi);
        double II;
        if (p->dim == 3)
            II = fabs((S.s11 * S.s22 * S.s33 + S.s12 * S.s23 * S.s13 + S.s13 * S.s12 * S.s23) -
                      (S.s13 * S.s22 * S.s13 + S.s11 * S.s23 * S.s23 + S.s12 * S.s12 * S.s33));
        else
            II = fabs(S.s11 * S.s22 - S.s12 * S.s12);
        double m = ys / 2 / sqrt(II) + p->meu0 * pow(4 * II, (p->n_pl - 1) / 2);
        if (II == 0 || m > 200) m = 200;
        meu[i] = m;
    }
    return VOX_OK;
}

/* ---- PREDICTION: functions.cu:2283-2369 ------------------------------------ */
int orc_prediction(orc_ctx* c)
{
    const vox_params* p = &c->p;
    const double *x = c->f[VOX_F_X], *y = c->f[VOX_F_Y], *z = c->f[VOX_F_Z];
    const double *u = c->f[VOX_F_U], *v = c->f[VOX_F_V], *w = c->f[VOX_F_W];
    const double *pn = c->f[VOX_F_PNEW], *ph = c->f[VOX_F_PHAT], *meu = c->f[VOX_F_MEU];
    const double *neut = c->f[VOX_F_NEUT], *rho = c->f[VOX_F_RHO];
    double *xs = c->f[VOX_F_XSTAR], *ys = c->f[VOX_F_YSTAR], *zs = c->f[VOX_F_ZSTAR];
    double *us = c->f[VOX_F_USTAR], *vs = c->f[VOX_F_VSTAR], *ws = c->f[VOX_F_WSTAR];
    const orc_list* L = LL(c);
    int dim = p->dim;
    double lapc = p->coef2d ? 4.0 : 2.0 * dim;      /* the literal 4 of 2354-2358 */
    double grdc = p->coef2d ? 2.0 : (double)dim;    /* the literal 2.0 of 2354-2358 */
    double dt = p->dt;
    #pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < c->tp; ++i) {
        if (c->ptype[i] <= 0) {
            xs[i] = x[i]; ys[i] = y[i]; us[i] = u[i]; vs[i] = v[i];
            if (dim == 3) { zs[i] = z[i]; ws[i] = w[i]; }
            continue;
        }
        double s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0, s6 = 0, s7 = 0, s8 = 0, s9 = 0;
        const int32_t* r = ROW(L, i);
        for (int l = 0; l < L->cnt[i]; ++l) {
            int j = r[l];
            double D = dist(c, x, y, z, i, j);
            if (i != j && D > 0) {
                double NEU;
                if (c->ptype[i] == c->ptype[j]) {
                    NEU = (c->ptype[i] == 1) ? meu[i] / p->rho1 : meu[i] / p->rho2;
                } else {
                    NEU = (c->ptype[i] == 1) ? 2 * meu[i] * meu[j] / (meu[i] + meu[j]) / p->rho1
                                             : 2 * meu[i] * meu[j] / (meu[i] + meu[j]) / p->rho2;
                }
                double wg = wfun(D, p->ktype, p->re_s);
                double wl = wfun(D, p->ktype, p->re_l);
                double dp = pn[j] - ph[i];
                s1 = s1 + dp * (x[j] - x[i]) * wg / D / D;
                s2 = s2 + dp * (y[j] - y[i]) * wg / D / D;
                if (dim == 3) s3 = s3 + dp * (z[j] - z[i]) * wg / D / D;
                s4 = s4 + wl * (u[j] - u[i]) * NEU;
                s5 = s5 + wl * (v[j] - v[i]) * NEU;
                if (dim == 3) s6 = s6 + wl * (w[j] - w[i]) * NEU;
                s7 = s7 + wl * (u[j] - u[i]);
                s8 = s8 + wl * (v[j] - v[i]);
                if (dim == 3) s9 = s9 + wl * (w[j] - w[i]);
            }
        }
        double ln0 = p->lambda * c->n0l;
        double gc = (1 - p->relaxp) * (grdc * dt / c->n0s / rho[i]);
        us[i] = u[i] + p->gx * dt + lapc * dt * (s4 + neut[i] * s7) / ln0 + dt * c->f[VOX_F_TURB1][i] - gc * s1;
        vs[i] = v[i] + p->gy * dt + lapc * dt * (s5 + neut[i] * s8) / ln0 + dt * c->f[VOX_F_TURB2][i] - gc * s2;
        if (dim == 3)
            ws[i] = w[i] + p->gz * dt + lapc * dt * (s6 + neut[i] * s9) / ln0 + dt * c->f[VOX_F_TURB3][i] - gc * s3;
        xs[i] = x[i] + dt * us[i];
        ys[i] = y[i] + dt * vs[i];
        if (dim == 3) zs[i] = z[i] + dt * ws[i];
    }
    return VOX_OK;
}

/* ---- BCON: functions.cu:433-445 -------------------------------------------- */
int orc_bcon(orc_ctx* c)
{
    const double* ns = c->f[VOX_F_NSTAR];
    for (int i = 0; i < c->tp; ++i) {
        if (c->ptype[i] <= -1) c->bcon[i] = -1;
        else if (ns[i] / c->n0s < c->p.dirichlet) c->bcon[i] = 1;
        else c->bcon[i] = 0;
    }
    return VOX_OK;
}

/* ---- MATRIX, Method 1, plain Laplacian + default source: 869-925, 990-1001 -- */
int orc_ppe_assemble(orc_ctx* c)
{
    const vox_params* p = &c->p;
    const orc_list* L = LL(c);
    /* CPU path at x with n (main.cu:1149-1150); CUDA path at x* with n*
     * (main.cu:1165, matrixCalc/sourceCalc functions.cu:1007-1133) */
    const int star = p->ref_gpu != 0;
    const double *x = c->f[star ? VOX_F_XSTAR : VOX_F_X], *y = c->f[star ? VOX_F_YSTAR : VOX_F_Y];
    const double *z = c->f[star ? VOX_F_ZSTAR : VOX_F_Z];
    const double* n = c->f[star ? VOX_F_NSTAR : VOX_F_N];
    double* src = c->f[VOX_F_SOURCE];
    int tp = c->tp;
    free(c->aval);
    c->aval = (double*)calloc((size_t)tp * L->cap, sizeof(double));   /* memset 2509-2510 */
    if (!c->aval) return VOX_ERR_ARG;
    memset(c->adiag, 0, sizeof(double) * tp);
    double dt = p->dt;
    int nu = 0;
    for (int i = 0; i < tp; ++i) c->uidx[i] = (c->bcon[i] == 0) ? nu++ : -1;
    c->nu = nu;
    #pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < tp; ++i) {
        if (c->bcon[i] != 0) continue;
        double diag = 0.0;
        const int32_t* r = ROW(L, i);
        double* a = c->aval + (size_t)i * L->cap;
        for (int l = 0; l < L->cnt[i]; ++l) {
            int j = r[l];
            if (i == j) continue;
            if (c->bcon[j] == -1) { a[l] = 0.0; continue; }
            double d = dist(c, x, y, z, i, j);
            if (fabs(d) <= 1.0e-15) continue;
            double val;
            if (p->matopt) {                                           /* 901-914 */
                val = (p->dim == 2) ? 3.0 * p->re_l : 2.0 * p->re_l;
                val = val / (d * d * d);
                val = val / c->n0l;
                val = val / p->rho1;
            } else {
                val = 2.0 * p->dim / p->lambda * wfun(d, p->ktype, p->re_l) / c->n0l / p->rho1;
            }
            a[l] = -val;
            diag += val;
        }
        diag += (1.00e-7) / dt / dt;                                   /* 925 */
        c->adiag[i] = diag;
    }
    memset(src, 0, sizeof(double) * tp);                                /* 2512 */
    if (p->srcopt) {
        /* 928-988, sequentially: source[i-1] is the value this loop has just
         * written (0 for particles it skips or sets to 0). The velocity
         * divergence uses the Laplacian list and n0, the PND terms n0 of n*. */
        const double* ns = c->f[VOX_F_NSTAR];
        const double *us = c->f[VOX_F_USTAR], *vs = c->f[VOX_F_VSTAR], *ws = c->f[VOX_F_WSTAR];
        const double re = p->re_l, n0l = c->n0l, n0s = c->n0s;
        for (int i = 0; i < tp; ++i) {
            if (c->ptype[i] == -1) continue;
            if (c->bcon[i] == 0) {
                double alpha = 0.0, beta = 0.0, ecs = 0.0;
                if (i > 0) {
                    alpha = (ns[i] - n0s) / n0s;
                    beta = (dt / n0s) * (src[i - 1] / (-1.0 / (n0s * dt)));
                    ecs = fabs(alpha) * (beta / dt) + fabs(beta) * (alpha / dt);
                }
                double sum = 0.0;
                const int32_t* r = ROW(L, i);
                for (int l = 0; l < L->cnt[i]; ++l) {
                    int j = r[l];
                    if (c->ptype[j] <= -1 || i == j) continue;
                    double d = dist(c, x, y, z, i, j);
                    if (fabs(d) <= 1.0e-15) continue;
                    double du = us[i] - us[j], dv = vs[i] - vs[j];
                    if (p->dim == 3) {
                        double dw = ws[i] - ws[j];
                        sum += (re / (d * d * d)) * ((x[i] - x[j]) * du + (y[i] - y[j]) * dv + (z[i] - z[j]) * dw);
                    } else {
                        sum += (re / (d * d * d)) * ((x[i] - x[j]) * du + (y[i] - y[j]) * dv);
                    }
                }
                src[i] = (-1.25 / (n0l * dt)) * sum + ecs;
                if (fabs(src[i]) < 0.99) src[i] = 0;
            } else if (c->bcon[i] == 1) {
                src[i] = 0.0;
            }
        }
        return VOX_OK;
    }
    for (int i = 0; i < tp; ++i) {                                      /* 990-1001 */
        if (c->ptype[i] == -1) continue;
        if (c->bcon[i] == 0) src[i] = 1.0 / dt / dt * (n[i] - c->n0s) / c->n0s;
        else if (c->bcon[i] == 1) src[i] = 0.0;
    }
    return VOX_OK;
}

static void spmv(const orc_ctx* c, const double* xin, double* s)
{
    const orc_list* L = LL(c);
    #pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < c->tp; ++i) {                     /* 577-591 */
        if (c->bcon[i] != 0) continue;
        double acc = c->adiag[i] * xin[i];
        const int32_t* r = ROW(L, i);
        const double* a = c->aval + (size_t)i * L->cap;
        for (int l = 0; l < L->cnt[i]; ++l) {
            int j = r[l];
            if (c->bcon[j] == -1) continue;
            acc = acc + a[l] * xin[j];
        }
        s[i] = acc;
    }
}

/* ---- INCDECOM + CGM, the reference's IC(0)-CG: functions.cu:1139-1200, 506-716.
 * Sequential, in the reference's loop order over list-L slots; a 3D row lists
 * the particle itself, so its slot takes part as the loops let it (the
 * triangular solves then read q[i] of the previous application there). */
static void ic_solve(const orc_ctx* c, const double* ic, const double* icd, const double* r,
                     double* aux, double* q)
{
    const orc_list* L = LL(c);
    const int tp = c->tp;
    memcpy(aux, r, sizeof(double) * tp);
    for (int i = 0; i < tp; ++i) {                                      /* 531-546 */
        if (c->bcon[i] != 0) continue;
        const int32_t* row = ROW(L, i);
        const double* a = ic + (size_t)i * L->cap;
        for (int l = 0; l < L->cnt[i]; ++l) {
            int j = row[l];
            if (j > i) continue;
            if (c->bcon[j] != 0) continue;
            aux[i] = aux[i] - a[l] * q[j];
        }
        q[i] = aux[i] / icd[i];
    }
    memcpy(aux, q, sizeof(double) * tp);
    for (int i = tp - 1; i >= 0; --i) {                                 /* 548-563 */
        if (c->bcon[i] != 0) continue;
        const int32_t* row = ROW(L, i);
        const double* a = ic + (size_t)i * L->cap;
        for (int l = 0; l < L->cnt[i]; ++l) {
            int j = row[l];
            if (j < i) continue;
            if (c->bcon[j] != 0) continue;
            aux[i] = aux[i] - a[l] * q[j];
        }
        q[i] = aux[i] / icd[i];
    }
}

static void a_times(const orc_ctx* c, const double* xin, double* s)
{
    const orc_list* L = LL(c);
    #pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < c->tp; ++i) {                                   /* 518-528, 578-588 */
        if (c->bcon[i] != 0) continue;
        s[i] = c->adiag[i] * xin[i];
        const int32_t* row = ROW(L, i);
        const double* a = c->aval + (size_t)i * L->cap;
        for (int l = 0; l < L->cnt[i]; ++l) {
            int j = row[l];
            if (c->bcon[j] == -1) continue;
            s[i] = s[i] + a[l] * xin[j];
        }
    }
}

int orc_ppe_solve_iccg(orc_ctx* c, int imin, int imax, double eps, vox_solve_info* info)
{
    if (!c || !c->aval) return VOX_ERR_STATE;
    const orc_list* L = LL(c);
    const int tp = c->tp, cap = L->cap;
    const double dt = c->p.dt;
    double* x = c->f[VOX_F_PNEW];
    const double* b = c->f[VOX_F_SOURCE];
    double* ic = (double*)calloc((size_t)tp * cap, sizeof(double));
    double* icd = (double*)calloc((size_t)tp, sizeof(double));
    double* r = (double*)calloc((size_t)tp, sizeof(double));
    double* p = (double*)calloc((size_t)tp, sizeof(double));
    double* q = (double*)calloc((size_t)tp, sizeof(double));
    double* s = (double*)calloc((size_t)tp, sizeof(double));
    double* aux = (double*)calloc((size_t)tp, sizeof(double));
    if (!ic || !icd || !r || !p || !q || !s || !aux) {
        free(ic); free(icd); free(r); free(p); free(q); free(s); free(aux);
        return VOX_ERR_ARG;
    }
    /* INCDECOM, 1144-1197 */
    for (int i = 0; i < tp; ++i) {
        if (c->bcon[i] != 0) continue;
        const int32_t* ri = ROW(L, i);
        double* ici = ic + (size_t)i * cap;
        const double* ai = c->aval + (size_t)i * cap;
        double sum = c->adiag[i];
        for (int l = 0; l < L->cnt[i]; ++l) {
            int j = ri[l];
            if (j > i) continue;
            if (c->bcon[j] != 0) continue;
            sum = sum - ici[l] * ici[l];
        }
        icd[i] = sqrt(sum);
        for (int l = 0; l < L->cnt[i]; ++l) {
            int j = ri[l];
            if (j < i) continue;
            if (c->bcon[j] != 0) continue;
            sum = ai[l];
            const int32_t* rj = ROW(L, j);
            double* icj = ic + (size_t)j * cap;
            for (int mj = 0; mj < L->cnt[j]; ++mj) {
                int kj = rj[mj];
                if (kj >= i) continue;
                if (c->bcon[kj] != 0) continue;
                for (int mi = 0; mi < L->cnt[i]; ++mi) {
                    if (ri[mi] == kj) {
                        sum = sum - ici[mi] * icj[mj];
                        break;
                    }
                }
            }
            ici[l] = sum / icd[i];
            for (int mj = 0; mj < L->cnt[j]; ++mj) {
                if (rj[mj] == i) {
                    icj[mj] = ici[l];
                    break;
                }
            }
        }
    }
    /* CGM, 506-716, from x = 0 (pnew was cleared by PRESSURECALC) */
    memset(x, 0, sizeof(double) * tp);
    a_times(c, x, s);
    for (int i = 0; i < tp; ++i) r[i] = b[i] - s[i];
    ic_solve(c, ic, icd, r, aux, q);
    memcpy(p, q, sizeof(double) * tp);
    double rqo = 0.0;
    for (int i = 0; i < tp; ++i)
        if (c->bcon[i] == 0) rqo += r[i] * q[i];
    int it = 0, conv = 0;
    for (int k = 0; k < imax; ++k) {
        a_times(c, p, s);
        double ps = 0.0;
        for (int i = 0; i < tp; ++i)
            if (c->bcon[i] == 0) ps += p[i] * s[i];
        double aa = rqo / ps;
        for (int i = 0; i < tp; ++i) x[i] = x[i] + aa * p[i];
        for (int i = 0; i < tp; ++i) r[i] = r[i] - aa * s[i];
        ic_solve(c, ic, icd, r, aux, q);
        double rqn = 0.0;
        for (int i = 0; i < tp; ++i)
            if (c->bcon[i] == 0) rqn += r[i] * q[i];
        double bb = rqn / rqo;
        rqo = rqn;
        for (int i = 0; i < tp; ++i) p[i] = q[i] + bb * p[i];
        int nbad = 0;                                                   /* 684-699 */
        for (int i = 0; i < tp; ++i) {
            if (c->bcon[i] != 0) continue;
            double err1 = (r[i] > 0) ? r[i] * dt * dt : -r[i] * dt * dt;
            if (err1 > eps) nbad++;
        }
        it = k + 1;
        if (nbad == 0 && k >= imin) { conv = 1; break; }
    }
    if (info) {
        double rr = 0.0, bb2 = 0.0;
        a_times(c, x, s);
        for (int i = 0; i < tp; ++i)
            if (c->bcon[i] == 0) {
                double e = b[i] - s[i];
                rr += e * e;
                bb2 += b[i] * b[i];
            }
        memset(info, 0, sizeof(*info));
        info->iterations = it;
        info->converged = conv;
        info->unknowns = c->nu;
        info->relres = bb2 > 0 ? sqrt(rr / bb2) : 0.0;
        info->bnorm = sqrt(bb2);
    }
    free(ic); free(icd); free(r); free(p); free(q); free(s); free(aux);
    return VOX_OK;
}

/* Dot products of the CG below: ORC_NCHUNK fixed index ranges summed in
 * parallel, each in index order, then the partial sums in chunk order. The
 * order depends on tp only, not on the thread count. */
#define ORC_NCHUNK 256
static inline int chunk_lo(int tp, int k) { return (int)((long long)tp * k / ORC_NCHUNK); }
static double sum_parts(const double* part)
{
    double s = 0.0;
    for (int k = 0; k < ORC_NCHUNK; ++k) s += part[k];
    return s;
}

/* Jacobi-preconditioned CG over the bcon == 0 rows, zero start (2503). */
int orc_ppe_solve(orc_ctx* c, vox_solve_info* info)
{
    int tp = c->tp;
    double* xs = c->f[VOX_F_PNEW];
    const double* b = c->f[VOX_F_SOURCE];
    const int32_t* bc = c->bcon;
    const double* ad = c->adiag;
    double* r = (double*)calloc((size_t)tp, sizeof(double));
    double* z = (double*)calloc((size_t)tp, sizeof(double));
    double* pv = (double*)calloc((size_t)tp, sizeof(double));
    double* q = (double*)calloc((size_t)tp, sizeof(double));
    if (!r || !z || !pv || !q) { free(r); free(z); free(pv); free(q); return VOX_ERR_ARG; }
    double pa[ORC_NCHUNK], pb[ORC_NCHUNK];
    for (int i = 0; i < tp; ++i) xs[i] = 0.0;
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < ORC_NCHUNK; ++k) {
        double sb = 0.0, sr = 0.0;
        for (int i = chunk_lo(tp, k); i < chunk_lo(tp, k + 1); ++i) {
            if (bc[i] != 0) continue;
            r[i] = b[i];
            z[i] = r[i] / ad[i];
            pv[i] = z[i];
            sb += b[i] * b[i];
            sr += r[i] * z[i];
        }
        pa[k] = sb;
        pb[k] = sr;
    }
    double bb = sum_parts(pa), rz = sum_parts(pb);
    double bnorm = sqrt(bb);
    int it = 0, conv = 0, rc = VOX_OK;
    double relres = 0.0;
    if (bnorm == 0.0) {
        conv = 1;
    } else {
        for (it = 1; it <= c->p.cg_maxit; ++it) {
            spmv(c, pv, q);
            #pragma omp parallel for schedule(static)
            for (int k = 0; k < ORC_NCHUNK; ++k) {
                double sp = 0.0;
                for (int i = chunk_lo(tp, k); i < chunk_lo(tp, k + 1); ++i)
                    if (bc[i] == 0) sp += pv[i] * q[i];
                pa[k] = sp;
            }
            double pq = sum_parts(pa);
            if (!(pq > 0.0) || !isfinite(pq)) { rc = VOX_ERR_SOLVER; break; }
            double alpha = rz / pq;
            #pragma omp parallel for schedule(static)
            for (int k = 0; k < ORC_NCHUNK; ++k) {
                double sr = 0.0;
                for (int i = chunk_lo(tp, k); i < chunk_lo(tp, k + 1); ++i) {
                    if (bc[i] != 0) continue;
                    xs[i] = xs[i] + alpha * pv[i];
                    r[i] = r[i] - alpha * q[i];
                    sr += r[i] * r[i];
                }
                pa[k] = sr;
            }
            double rr = sum_parts(pa);
            relres = sqrt(rr) / bnorm;
            if (relres <= c->p.cg_tol) { conv = 1; break; }
            #pragma omp parallel for schedule(static)
            for (int k = 0; k < ORC_NCHUNK; ++k) {
                double sz = 0.0;
                for (int i = chunk_lo(tp, k); i < chunk_lo(tp, k + 1); ++i) {
                    if (bc[i] != 0) continue;
                    z[i] = r[i] / ad[i];
                    sz += r[i] * z[i];
                }
                pa[k] = sz;
            }
            double rzn = sum_parts(pa);
            double beta = rzn / rz;
            rz = rzn;
            #pragma omp parallel for schedule(static)
            for (int i = 0; i < tp; ++i) if (bc[i] == 0) pv[i] = z[i] + beta * pv[i];
        }
        if (it > c->p.cg_maxit) it = c->p.cg_maxit;
    }
    if (info) {
        info->iterations = it;
        info->converged = conv;
        info->unknowns = c->nu;
        info->relres = relres;
        info->bnorm = bnorm;
        int nnz = 0;
        orc_ppe_export(c, NULL, &nnz, NULL, NULL, NULL, NULL, NULL);
        info->nnz = nnz;
    }
    c->st.last_iters = it;
    c->st.last_relres = relres;
    free(r); free(z); free(pv); free(q);
    if (rc) return rc;
    return conv ? VOX_OK : VOX_ERR_SOLVER;
}

/* ---- PRESSURECALC, Method 1: functions.cu:2501-2547 ------------------------ */
int orc_pressurecalc(orc_ctx* c, vox_solve_info* info)
{
    double* pn = c->f[VOX_F_PNEW];
    if (c->p.method == 3) {                                   /* 2471-2500 */
        const double* ns = c->f[VOX_F_NSTAR];
        const double* C = c->f[VOX_F_C];
        const vox_params* p = &c->p;
        #pragma omp parallel for schedule(static)
        for (int i = c->gp; i < c->tp; ++i) {
            const double c0 = (c->ptype[i] == 2) ? p->c02 : p->c01;
            const double rho = p->rho2 * C[i] + p->rho1 * (1 - C[i]);
            if (ns[i] < p->wc_beta * c->n0s) pn[i] = 0.0;
            else pn[i] = (c0 * c0 * rho / 7.0) * (pow(ns[i] / c->n0s, 7.0) - 1);
            if (pn[i] < p->pmin) pn[i] = p->pmin;
            if (pn[i] > p->pmax) pn[i] = p->pmax;
        }
        if (info) memset(info, 0, sizeof(*info));
        return VOX_OK;
    }
    memset(pn, 0, sizeof(double) * c->tp);                  /* 2503 */
    int rc = orc_ppe_assemble(c);
    if (rc) return rc;
    rc = orc_ppe_solve(c, info);
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < c->tp; ++i) {                         /* 2534-2547 */
        if (pn[i] < c->p.pmin) pn[i] = c->p.pmin;
        if (pn[i] > c->p.pmax || pn[i] * 0.0 != 0.0) pn[i] = c->p.pmax;
    }
    return rc;
}

/* ---- PHATCALC: functions.cu:207-224 ---------------------------------------- */
int orc_phat(orc_ctx* c)
{
    const orc_list* L = LS(c);
    const double* pn = c->f[VOX_F_PNEW];
    #pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < c->tp; ++i) {
        double mn = 999999999999999;
        const int32_t* r = ROW(L, i);
        for (int l = 0; l < L->cnt[i]; ++l) {
            int j = r[l];
            if (pn[j] < mn) mn = pn[j];
        }
        c->f[VOX_F_PHAT][i] = mn;
    }
    return VOX_OK;
}

/* ---- PRESSGRAD: functions.cu:245-355 (KH 1 and 0) -------------------------- */
int orc_pressgrad(orc_ctx* c)
{
    const vox_params* p = &c->p;
    const double *x = c->f[VOX_F_XSTAR], *y = c->f[VOX_F_YSTAR], *z = c->f[VOX_F_ZSTAR];
    const double *pn = c->f[VOX_F_PNEW], *ph = c->f[VOX_F_PHAT], *rho = c->f[VOX_F_RHO];
    const orc_list* L = LS(c);
    int dim = p->dim;
    double grdc = p->coef2d ? 2.0 : (double)dim;    /* literal 2.0 of 339-343 */
    #pragma omp parallel for schedule(dynamic, 256)
    for (int i = c->gp + c->wp; i < c->tp; ++i) {
        double s1 = 0, s2 = 0, s3 = 0;
        const int32_t* r = ROW(L, i);
        for (int l = 0; l < L->cnt[i]; ++l) {
            int j = r[l];
            double D = dist(c, x, y, z, i, j);
            if (i == j) continue;
            double wv = wfun(D, p->ktype, p->re_s);
            double dp;
            if (p->kh_correction == 1 && p->ref_gpu) dp = pn[j] + pn[i] - 2.0 * ph[i];   /* 382-385 */
            else if (p->kh_correction == 1) dp = pn[j] + pn[i] - (ph[i] + ph[j]);
            else dp = pn[j] - ph[i];
            s1 = s1 + dp * (x[j] - x[i]) * wv / D / D;
            s2 = s2 + dp * (y[j] - y[i]) * wv / D / D;
            if (dim == 3) s3 = s3 + dp * (z[j] - z[i]) * wv / D / D;
        }
        double g = p->relaxp * (grdc * p->dt / c->n0s / rho[i]);
        double un = c->f[VOX_F_USTAR][i] - g * s1;
        double vn = c->f[VOX_F_VSTAR][i] - g * s2;
        double wn = 0.0;
        if (dim == 3) wn = c->f[VOX_F_WSTAR][i] - g * s3;
        double cf = 2.0;                                     /* damper 345-351 */
        if (fabs(un) > cf * p->vmax) un = p->vmax;
        if (vn > cf * p->vmax) vn = cf * p->vmax;
        if (vn < -cf * p->vmax) vn = -cf * p->vmax;
        if (dim == 3 && fabs(wn) > cf * p->vmax) wn = p->vmax;
        c->f[VOX_F_UNEW][i] = un;
        c->f[VOX_F_VNEW][i] = vn;
        if (dim == 3) c->f[VOX_F_WNEW][i] = wn;
    }
    return VOX_OK;
}

/* ---- EULERINTEGRATION + BOUNDARIE: functions.cu:1963-2021 ------------------ */
int orc_euler(orc_ctx* c)
{
    const vox_params* p = &c->p;
    double *x = c->f[VOX_F_X], *y = c->f[VOX_F_Y], *z = c->f[VOX_F_Z];
    double *u = c->f[VOX_F_UNEW], *v = c->f[VOX_F_VNEW], *w = c->f[VOX_F_WNEW];
    /* the CUDA path passes the collision factor CC instead of velCor (main.cu:1293) */
    double DL = p->dl, CC = p->ref_gpu ? p->cc : p->vel_cor;
    #pragma omp parallel for schedule(static)
    for (int i = c->gp + c->wp; i < c->tp; ++i) {
        x[i] = x[i] + u[i] * p->dt;
        y[i] = y[i] + v[i] * p->dt;
        if (p->dim == 3) z[i] = z[i] + w[i] * p->dt;
        if (p->ref_gpu) {                                   /* eulerIntegration 2023-2047 */
            if (y[i] < p->ymin + DL * 3.0) { v[i] = fabs(v[i]) * CC; y[i] = p->ymin + DL * 3.0; }
            if (y[i] > p->ymax - DL / 2.0) { v[i] = -fabs(v[i]) * CC; y[i] = p->ymax - DL / 2.0; }
            if (x[i] < p->xmin + DL * 3.0) { u[i] = fabs(u[i]) * CC; x[i] = p->xmin + DL * 3.0; }
            if (x[i] > p->xmax - DL * 3.0) { u[i] = -fabs(u[i]) * CC; x[i] = p->xmax - DL * 3.0; }
            continue;                                       /* 2D only */
        }
        if (p->test_kind == VOX_TEST_DROP) continue;
        if (p->test_kind != VOX_TEST_TURB) {
            if (y[i] < p->ymin + DL * 3.0) { v[i] = fabs(v[i]) * CC; y[i] = p->ymin + DL * 3.0; }
            if (y[i] > p->ymax - DL * 3.0) { v[i] = -fabs(v[i]) * CC; y[i] = p->ymax - DL * 3.0; }
            if (x[i] < p->xmin + DL * 3.0) { u[i] = fabs(u[i]) * CC; x[i] = p->xmin + DL * 3.0; }
            if (x[i] > p->xmax - DL * 3.0) { u[i] = -fabs(u[i]) * CC; x[i] = p->xmax - DL * 3.0; }
            if (p->dim == 3) {
                if (z[i] < p->zmin + DL * 3.0) { w[i] = fabs(w[i]) * CC; z[i] = p->zmin + DL * 3.0; }
                if (z[i] > p->zmax - DL * 3.0) { w[i] = -fabs(w[i]) * CC; z[i] = p->zmax - DL * 3.0; }
            }
        } else {
            if (y[i] < p->ymin + DL * 2.0) { v[i] = fabs(v[i]) * CC; y[i] = p->ymin + DL * 2.0; }
            if (y[i] > p->ymax - DL * 2.0) { v[i] = -fabs(v[i]) * CC; y[i] = p->ymax - DL * 2.0; }
            if (x[i] > p->xmax + DL * 0.0) { u[i] = (u[i]) * CC; x[i] = p->xmin; }
        }
    }
    return VOX_OK;
}

/* ---- COLLISION2: functions.cu:1523-1601, sequential in-place ---------------- */
int orc_collision(orc_ctx* c)
{
    const vox_params* p = &c->p;
    double *x = c->f[VOX_F_X], *y = c->f[VOX_F_Y], *z = c->f[VOX_F_Z];
    double *u = c->f[VOX_F_UNEW], *v = c->f[VOX_F_VNEW], *w = c->f[VOX_F_WNEW];
    const orc_list* L = LS(c);
    int dim = p->dim, hits = 0;
    double dt = p->dt;
    for (int i = 0; i < c->tp; ++i) {
        double m1 = (c->ptype[i] == 1) ? p->rho1 : p->rho2;
        const int32_t* r = ROW(L, i);
        for (int l = 0; l < L->cnt[i]; ++l) {
            int j = r[l];
            double m2 = (c->ptype[j] == 1) ? p->rho1 : p->rho2;
            double d = dist(c, x, y, z, i, j);
            if (!(i != j && d < p->coll)) continue;
            double ug = (m1 * u[i] + m2 * u[j]) / (m1 + m2);
            double vg = (m1 * v[i] + m2 * v[j]) / (m1 + m2);
            double wg = 0, wr = 0, wm = 0;
            if (dim == 3) wg = (m1 * w[i] + m2 * w[j]) / (m1 + m2);
            double ur = m1 * (u[i] - ug);
            double vr = m1 * (v[i] - vg);
            if (dim == 3) wr = m1 * (w[i] - wg);
            double dx = x[j] - x[i], dy = y[j] - y[i], dz = 0;
            if (dim == 3) dz = z[j] - z[i];
            double vabs = (dim == 3) ? (ur * dx + vr * dy + wr * dz) / d : (ur * dx + vr * dy) / d;
            double um = (1.0 + p->cc) * vabs * dx / d;
            double vm = (1.0 + p->cc) * vabs * dy / d;
            if (dim == 3) wm = (1.0 + p->cc) * vabs * dz / d;
            if (vabs > 0) {
                ++hits;
                if (c->ptype[i] > 0) {
                    u[i] = u[i] - um / m1;
                    v[i] = v[i] - vm / m1;
                    x[i] = x[i] - dt * um / m1;
                    y[i] = y[i] - dt * vm / m1;
                    if (dim == 3) { w[i] = w[i] - wm / m1; z[i] = z[i] - dt * wm / m1; }
                }
                if (c->ptype[j] > 0) {
                    u[j] = u[j] + um / m2;
                    v[j] = v[j] + vm / m2;
                    x[j] = x[j] + dt * um / m2;
                    y[j] = y[j] + dt * vm / m2;
                    if (dim == 3) { w[j] = w[j] + wm / m2; z[j] = z[j] + dt * wm / m2; }
                }
            }
        }
    }
    c->st.collisions = hits;
    return VOX_OK;
}

/* ---- PREPDATA: functions.cu:2688-2746 --------------------------------------- */
int orc_prepdata(orc_ctx* c)
{
    const vox_params* p = &c->p;
    double *x = c->f[VOX_F_X], *y = c->f[VOX_F_Y], *z = c->f[VOX_F_Z];
    double *u = c->f[VOX_F_U], *v = c->f[VOX_F_V], *w = c->f[VOX_F_W], *pp = c->f[VOX_F_P];
    int tp = c->tp, dim = p->dim, nanc = 0, oob = 0;
    int first = (p->test_kind == VOX_TEST_DROP) ? tp - c->fp - 1 : tp - c->fp;
    if (p->ref_gpu) first = 0;                 /* prepData copies u of every particle (2753) */
    #pragma omp parallel for schedule(static) reduction(+ : nanc, oob)
    for (int i = 0; i < tp; ++i) {
        pp[i] = c->f[VOX_F_PNEW][i];
        if (i >= first) {
            u[i] = c->f[VOX_F_UNEW][i];
            v[i] = c->f[VOX_F_VNEW][i];
            if (dim == 3) w[i] = c->f[VOX_F_WNEW][i];
        }
        if (dim == 3) {
            if (pp[i] * 0 != 0 || u[i] * 0 != 0 || v[i] * 0 != 0 || w[i] * 0 != 0) ++nanc;
            if (x[i] >= p->xmax || y[i] >= p->ymax || z[i] >= p->zmax || x[i] <= p->xmin ||
                y[i] <= p->ymin || z[i] <= p->zmin) ++oob;
        } else {
            if (pp[i] * 0 != 0 || u[i] * 0 != 0 || v[i] * 0 != 0) ++nanc;
            if (x[i] >= p->xmax || y[i] >= p->ymax || x[i] <= p->xmin || y[i] <= p->ymin) ++oob;
        }
    }
    c->st.nan_count = nanc;
    c->st.out_of_box = oob;
    return VOX_OK;
}

/* Threads the parallel loops above run on (1 without OpenMP). */
int orc_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Sets that thread count (n >= 1); returns the count now in effect. */
int orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n >= 1) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* ---- DTcalculation: functions.cu:1684-1698 ---------------------------------- */
int orc_dtcalc(orc_ctx* c)
{
    vox_params* p = &c->p;
    double c0 = (p->c01 > p->c02) ? p->c01 : p->c02;
    double courant = p->courant * p->dl / c0;
    if (p->dt > courant) p->dt = courant;
    else if (courant < p->dt_max) p->dt = courant;
    else p->dt = p->dt_max;
    return VOX_OK;
}

/* ---- initialisation and the time loop: main.cu:966-1012, 1027-1363 --------- */
int orc_init(orc_ctx* c)
{
    int rc = orc_density(c);
    if (!rc) rc = orc_neighbor(c);
    if (!rc) rc = orc_pnum(c, 0);
    if (rc) return rc;
    c->initialised = 1;
    c->st.steps = 0;
    c->st.time = 0.0;
    return VOX_OK;
}

int orc_step(orc_ctx* c, int nsteps)
{
    if (!c->initialised) return VOX_ERR_STATE;
    for (int s = 0; s < nsteps; ++s) {
        int rc = VOX_OK;
        if (c->p.dt_mode == 1) orc_dtcalc(c);
        if (!rc && c->p.method != 3) rc = orc_pnum(c, 1);     /* 1036: not in method 3 */
        if (!rc) rc = orc_neighbor(c);
        if (!rc && c->p.turb) rc = orc_sps(c);
        if (!rc) rc = orc_vfraction(c);
        if (!rc) rc = orc_viscosity(c);
        if (!rc) rc = orc_prediction(c);
        if (!rc) rc = orc_pnumstar(c);
        if (!rc && c->p.method != 3) rc = orc_bcon(c);        /* 1133 */
        if (!rc) rc = orc_pressurecalc(c, NULL);
        if (!rc) rc = orc_phat(c);
        if (!rc) rc = orc_pressgrad(c);
        if (!rc) rc = orc_euler(c);
        if (!rc) rc = orc_collision(c);
        if (!rc) rc = orc_prepdata(c);
        if (rc) return rc;
        c->st.steps += 1;
        c->st.time += c->p.dt;
    }
    return VOX_OK;
}

int orc_get_status(orc_ctx* c, vox_status* st)
{
    if (!c || !st) return VOX_ERR_ARG;
    *st = c->st;
    st->n0_s = c->n0s;
    st->n0_l = c->n0l;
    st->max_count_s = LS(c)->maxcnt;
    st->max_count_l = LL(c)->maxcnt;
    return VOX_OK;
}

int orc_get_n0(orc_ctx* c, double* n0s, double* n0l)
{
    if (!c) return VOX_ERR_ARG;
    if (n0s) *n0s = c->n0s;
    if (n0l) *n0l = c->n0l;
    return VOX_OK;
}

int orc_set_n0(orc_ctx* c, double n0s, double n0l)
{
    if (!c) return VOX_ERR_ARG;
    c->n0s = n0s;
    c->n0l = n0l;
    return VOX_OK;
}

int orc_max_count(orc_ctx* c, int which)
{
    if (!c || which < 0 || which > 1) return -1;
    return (which == VOX_LIST_S ? LS(c) : LL(c))->maxcnt;
}

int orc_get_neighbors(orc_ctx* c, int which, int32_t* counts, int32_t* ids, int stride)
{
    if (!c || which < 0 || which > 1) return VOX_ERR_ARG;
    const orc_list* L = (which == VOX_LIST_S) ? LS(c) : LL(c);
    if (!L->ids) return VOX_ERR_STATE;
    if (ids && stride < L->maxcnt) return VOX_ERR_OVERFLOW;
    for (int i = 0; i < c->tp; ++i) {
        if (counts) counts[i] = L->cnt[i];
        if (ids) for (int l = 0; l < L->cnt[i]; ++l) ids[(size_t)i * stride + l] = ROW(L, i)[l];
    }
    return VOX_OK;
}

int orc_get_cells(orc_ctx* c, int which, int32_t* cell)
{
    if (!c || !cell || which < 0 || which > 1) return VOX_ERR_ARG;
    const orc_list* L = (which == VOX_LIST_S) ? LS(c) : LL(c);
    memcpy(cell, L->cell, sizeof(int32_t) * c->tp);
    return VOX_OK;
}

/* CSR of the bcon == 0 rows: diagonal first, then the nonzero off-diagonals
 * whose column is an unknown, in list-L order. */
int orc_ppe_export(orc_ctx* c, int32_t* nu, int32_t* nnz, int32_t* rowptr, int32_t* cols,
                   double* vals, double* rhs, int32_t* ids)
{
    if (!c || !c->aval) return VOX_ERR_STATE;
    const orc_list* L = LL(c);
    int k = 0, r = 0;
    for (int i = 0; i < c->tp; ++i) {
        if (c->bcon[i] != 0) continue;
        if (rowptr) rowptr[r] = k;
        if (cols) cols[k] = r;
        if (vals) vals[k] = c->adiag[i];
        ++k;
        const double* a = c->aval + (size_t)i * L->cap;
        for (int l = 0; l < L->cnt[i]; ++l) {
            int j = ROW(L, i)[l];
            if (j == i || c->bcon[j] != 0 || a[l] == 0.0) continue;
            if (cols) cols[k] = c->uidx[j];
            if (vals) vals[k] = a[l];
            ++k;
        }
        if (rhs) rhs[r] = c->f[VOX_F_SOURCE][i];
        if (ids) ids[r] = i;
        ++r;
    }
    if (rowptr) rowptr[r] = k;
    if (nu) *nu = r;
    if (nnz) *nnz = k;
    return VOX_OK;
}
