// vox_api.hip — C ABI of libvoxarmps (declared in include/voxarmps.h).
//
// The context owns every device array; entry points mirror the reference's
// per-stage host routines (functions.cuh:56-111) and vox_step replays the
// body of its time loop (main.cu:1027-1363) in the same order.
#include <algorithm>
#include <cstring>
#include <mutex>

#include "vox_internal.hpp"

namespace vox {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }
int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

static const char* kKNames[KC_COUNT] = {
    "cell", "sort", "list", "pnd", "sps", "vfrac", "visc", "pred", "bcon", "assemble",
    "cg_init", "spmv", "cg_update", "cg_p", "phat", "pgrad", "euler", "collision", "prep", "misc",
    "ir_residual", "list_l", "lpc_net", "lpc_apply"};

const char* kclass_name(int k) { return (k >= 0 && k < KC_COUNT) ? kKNames[k] : "?"; }
int kclass_from_name(const char* name)
{
    for (int k = 0; k < KC_COUNT; ++k)
        if (std::strcmp(kKNames[k], name) == 0) return k;
    return -1;
}

static hipEvent_t prof_event(vox_ctx* c)
{
    if (!c->prof_free.empty()) {
        hipEvent_t e = c->prof_free.back();
        c->prof_free.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void prof_begin(vox_ctx* c, int kclass, hipEvent_t* a)
{
    (void)kclass;
    *a = prof_event(c);
    if (*a) (void)hipEventRecord(*a, c->stream);
}

void prof_end(vox_ctx* c, int kclass, hipEvent_t a)
{
    if (!a) return;
    hipEvent_t b = prof_event(c);
    if (!b) return;
    (void)hipEventRecord(b, c->stream);
    c->prof_pending.push_back(ProfRec{kclass, a, b});
    if (c->prof_pending.size() > 4096) prof_drain(c);
}

int prof_drain(vox_ctx* c)
{
    if (c->prof_pending.empty()) return VOX_OK;
    VOX_HIP(hipStreamSynchronize(c->stream));
    for (auto& r : c->prof_pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            c->prof_ms[r.kclass] += ms;
            c->prof_n[r.kclass] += 1;
        }
        c->prof_free.push_back(r.a);
        c->prof_free.push_back(r.b);
    }
    c->prof_pending.clear();
    return VOX_OK;
}

int read_flags(vox_ctx* c)
{
    VOX_HIP(hipMemcpyAsync(c->hflags, c->dflags, sizeof(int32_t) * FLAG_COUNT,
                           hipMemcpyDeviceToHost, c->stream));
    VOX_HIP(hipMemcpyAsync(c->hscal, c->dscal, sizeof(double) * SC_COUNT, hipMemcpyDeviceToHost,
                           c->stream));
    VOX_HIP(hipStreamSynchronize(c->stream));
    return VOX_OK;
}

// Lattice count of neighbours inside the accept radius, used to size rows
// (incl: the inclusive rule of 3D and of the CUDA path, self counted).
static int lattice_count(int dim, int incl, double dl, double re, double delta)
{
    double R = incl ? re + delta : re;
    int m = (int)std::ceil(R / dl) + 1;
    int cnt = 0;
    for (int a = -m; a <= m; ++a)
        for (int b = -m; b <= m; ++b)
            for (int cc = (dim == 3 ? -m : 0); cc <= (dim == 3 ? m : 0); ++cc) {
                double d = std::sqrt((double)(a * a + b * b + cc * cc)) * dl;
                if (incl ? d <= R : (d < R && (a || b))) ++cnt;
            }
    return cnt;
}

static int auto_cap(int dim, int incl, double dl, double re, double delta)
{
    int n = lattice_count(dim, incl, dl, re, delta);
    int cap = n + n / 3 + 8;
    return (cap + 7) & ~7;
}

template <class T>
static int dalloc(T** p, size_t n)
{
    VOX_HIP(hipMalloc((void**)p, sizeof(T) * (n > 0 ? n : 1)));
    VOX_HIP(hipMemset(*p, 0, sizeof(T) * (n > 0 ? n : 1)));
    return VOX_OK;
}

template <class T>
static void dfree(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

static int check_ctx(vox_ctx* c)
{
    if (!c) return fail(VOX_ERR_ARG, "null context");
    return VOX_OK;
}

static int need_init(vox_ctx* c)
{
    VOX_TRY(check_ctx(c));
    if (!c->initialised) return fail(VOX_ERR_STATE, "vox_init has not run on this context");
    return VOX_OK;
}

}  // namespace vox

using namespace vox;

namespace vox {

// Kernel variants of the autotuner: every variant gives the same bits, so
// the choice is by time only. The step's stages up to MATRIX run once on the
// current state (each field they write is rewritten by the next step before
// it is read; the pressure is not touched), then each predictor variant
// (FP64 path) and each PPE product variant is timed over `reps` launches.
int autotune_kernels(vox_ctx* c, int reps)
{
    if (reps < 1) reps = 1;
    const bool wc = (c->p.method == 3);
    if (!wc) VOX_TRY(st_pnd(c, 0, 1));
    VOX_TRY(build_neighbors(c));
    if (c->p.turb) VOX_TRY(st_sps(c));
    VOX_TRY(st_vfraction(c));
    VOX_TRY(st_viscosity(c));
    hipEvent_t e0, e1;
    VOX_HIP(hipEventCreate(&e0));
    VOX_HIP(hipEventCreate(&e1));
    int rc = VOX_OK;
    if (c->p.precision == VOX_PREC_FP64) {
        float best = 0.f;
        int pick = 0;
        for (int v = 0; v < kPredVariants && rc == VOX_OK; ++v) {
            c->tune_pred = v;
            if ((rc = st_prediction(c)) != VOX_OK) break;                // warm
            (void)hipEventRecord(e0, c->stream);
            for (int k = 0; k < reps && rc == VOX_OK; ++k) rc = st_prediction(c);
            (void)hipEventRecord(e1, c->stream);
            if (rc != VOX_OK) break;
            if (hipEventSynchronize(e1) != hipSuccess) { rc = fail(VOX_ERR_HIP, "autotune sync"); break; }
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (v == 0 || ms < best) { best = ms; pick = v; }
        }
        c->tune_pred = pick;
    }
    if (rc == VOX_OK) rc = st_prediction(c);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    VOX_TRY(rc);
    VOX_TRY(st_pnd(c, 1, 1));
    if (wc) return VOX_OK;
    VOX_TRY(st_bcon(c));
    VOX_TRY(ppe_assemble(c));
    if (c->nu <= 0) return VOX_OK;
    float ms[kSpmvVariants] = {};
    VOX_TRY(time_spmv_variants(c, reps, ms));
    int pick = 0;
    for (int v = 1; v < kSpmvVariants; ++v)
        if (ms[v] < ms[pick]) pick = v;
    c->tune_spmv = pick;
    return VOX_OK;
}

}  // namespace vox

extern "C" {

const char* vox_last_error(void) { return g_err.c_str(); }
const char* vox_version(void) { return "voxarmps-amd 0.1 (gfx950)"; }

int vox_abi_sizes(int32_t* params, int32_t* solve_info, int32_t* status)
{
    if (params) *params = (int32_t)sizeof(vox_params);
    if (solve_info) *solve_info = (int32_t)sizeof(vox_solve_info);
    if (status) *status = (int32_t)sizeof(vox_status);
    return VOX_OK;
}

int vox_device_count(int* count)
{
    if (!count) return fail(VOX_ERR_ARG, "null count");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return VOX_OK;
}

int vox_params_default(vox_params* p, int dim, double dl)
{
    if (!p || (dim != 2 && dim != 3) || !(dl > 0)) return fail(VOX_ERR_ARG, "bad default request");
    std::memset(p, 0, sizeof(*p));
    p->dim = dim;
    p->ktype = 2;                                   // main.cu:288-292
    p->dl = dl;
    double re = (dim == 3) ? 3.2 * dl : 4.0 * dl;   // main.cu:252-260
    p->re_s = re;
    p->re_l = re;
    p->delta = dl / 5.0;                            // main.cu:282
    p->lambda = 0.16832 * (re * re);                // main.cu:492
    p->rho1 = 1000.0; p->nu1 = 0.000001;            // GUI defaults (Form1.Designer.cs)
    p->rho2 = 500.0;  p->nu2 = 0.001;
    p->gx = 0.0; p->gy = -9.8; p->gz = 0.0;         // main.cu:321-330
    p->courant = 1.0; p->c01 = 20.0; p->c02 = 20.0; p->dt_max = 0.005;
    p->dt = p->courant * dl / 20.0;                 // value DTcalculation forces
    p->dt_mode = 0;
    p->turb = 0;
    p->dirichlet = 0.97;                            // main.cu:228
    p->relaxp = 0.2;                                // main.cu:263-265
    p->coll = std::sqrt(dl * dl * 0.5 * 0.5);       // main.cu:269-273
    p->cc = 0.20;                                   // main.cu:277-279
    p->vel_cor = 1.0;                               // main.cu:275
    p->cs = 0.18;                                   // main.cu:281
    p->ncorrection = 0.99;                          // main.cu:284
    p->vmax = 2.50;                                 // main.cu:331
    p->pmin = 0.0;
    p->pmax = -1.0;                                 // -1: 4.5*TP at create (main.cu:548)
    p->fluid2_type = 0;
    p->kh_correction = 1;                           // main.cu:344
    p->n_pl = 1.3; p->meu0 = 0.04; p->phi = 0.0; p->cohes = 0.0;   // main.cu:365-372
    p->test_kind = VOX_TEST_DAM;
    p->coef2d = 1;
    p->precision = VOX_PREC_FP64;
    p->solver = VOX_SOLVER_JACOBI_CG;
    p->cg_tol = 1e-9;
    p->cg_maxit = 5000;
    p->cg_check = 1;
    p->ir_eta = 1e-4;
    p->ir_maxouter = 8;
    p->warm_start = 0;
    p->list_mode = 0;
    p->lpc_cap_pct = 200;
    p->method = 1;
    p->matopt = 0;
    p->srcopt = 0;
    p->ref_gpu = 0;
    p->wc_beta = 0.92;                              // main.cu:262
    return VOX_OK;
}

int vox_params_box(vox_params* p, int tp, const double* x, const double* y, const double* z)
{
    if (!p || tp <= 0 || !x || !y || (p->dim == 3 && !z)) return fail(VOX_ERR_ARG, "bad box request");
    double xmn = 99999999, ymn = 99999999, zmn = 99999999;        // main.cu:625-630
    double xmx = -99999999, ymx = -99999999, zmx = -99999999;
    for (int i = 0; i < tp; ++i) {
        if (x[i] < xmn) xmn = x[i];
        if (x[i] > xmx) xmx = x[i];
        if (y[i] < ymn) ymn = y[i];
        if (y[i] > ymx) ymx = y[i];
        if (p->dim == 3) {
            if (z[i] < zmn) zmn = z[i];
            if (z[i] > zmx) zmx = z[i];
        }
    }
    p->xmin = xmn - p->dl; p->ymin = ymn - p->dl;                   // main.cu:704-715
    p->xmax = xmx + p->dl; p->ymax = ymx + p->dl;
    if (p->dim == 3) { p->zmin = zmn - p->dl; p->zmax = zmx + p->dl; }
    else { p->zmin = 0.0; p->zmax = 0.0; }
    return VOX_OK;
}

// The run main.cu sets up from its 16 command-line arguments: the test and
// option words (main.cu:88-209), the constants they select (211-494), the
// input's box (692-715) and PMAX = 4.5 TP (548).
int vox_params_main(vox_params* p, const char* test, int visc, int turb, int comp, int matopt,
                    int srcopt, const char* code_opt, int threedim, double rho1, double nu1,
                    double rho2, double nu2, double dt, int tp, const double* x, const double* y,
                    const double* z)
{
    if (!p || !test || !code_opt) return fail(VOX_ERR_ARG, "null argument");
    std::string t(test), code(code_opt);
    if (t != "oil" && t != "drop" && t != "shear" && t != "turb") t = "dam";      // main.cu:88-108
    if (code != "openmp" && code != "gpu" && code != "sequential")                // main.cu:154-171
        return fail(VOX_ERR_ARG, "invalid code optimization option (openmp, gpu or sequential)");
    if (!(dt > 0.0) || dt > 0.005)                                              // main.cu:396-402
        return fail(VOX_ERR_ARG, "time-step duration must be in (0, 0.005] s");
    const bool compressible = comp || t == "oil";                               // main.cu:128-138
    const bool three = threedim && t == "dam";                                  // main.cu:173-183
    if (three && code == "gpu")
        return fail(VOX_ERR_ARG, "the reference's CUDA path is 2D only");
    const double dl = compressible ? 0.0040 : 0.010;                            // main.cu:248-250
    VOX_TRY(vox_params_default(p, three ? 3 : 2, dl));
    const double re = p->re_s;                                                  // 3.2 DL / 4.0 DL
    p->relaxp = compressible ? 0.5 : 0.2;                                       // main.cu:263-265
    p->cc = compressible ? 0.50 : 0.20;                                         // main.cu:277-279
    p->ktype = compressible ? 6 : 2;                                            // main.cu:288-292
    p->method = compressible ? 3 : 1;                                           // main.cu:335-339
    p->turb = (t == "turb" || turb) ? 1 : 0;                                    // main.cu:303-312
    p->gy = (t == "turb" || t == "shear") ? 0.0 : -9.8;                          // main.cu:316-328
    p->fluid2_type = visc ? 1 : 0;                                              // main.cu:362-377
    if (visc) { p->n_pl = 1.3; p->meu0 = 0.04; p->phi = 0.0; p->cohes = 0.0; }
    else { p->n_pl = 0.0; p->meu0 = 0.0; p->phi = 0.0; p->cohes = 0.0; }
    p->rho1 = rho1; p->nu1 = nu1; p->rho2 = rho2; p->nu2 = nu2;                 // main.cu:244-247
    p->lambda = (p->ktype == 6 ? 0.14288 : 0.16832) * (re * re);                // main.cu:490-495
    p->dt = dt;
    p->dt_mode = 1;                                                             // main.cu:1030
    p->vel_cor = (t == "turb") ? 0.8 : 1.0;                                     // main.cu:509-512
    p->test_kind = (t == "drop") ? VOX_TEST_DROP : (t == "turb") ? VOX_TEST_TURB : VOX_TEST_DAM;
    p->matopt = matopt ? 1 : 0;
    p->srcopt = srcopt ? 1 : 0;
    p->ref_gpu = (code == "gpu") ? 1 : 0;
    if (tp > 0 && x && y) {
        VOX_TRY(vox_params_box(p, tp, x, y, z));                                // main.cu:704-715
        if (t == "drop") {                                                      // main.cu:692-703
            double xmn = x[0], xmx = x[0], ymn = y[0], ymx = y[0];
            for (int i = 1; i < tp; ++i) {
                xmn = std::min(xmn, x[i]); xmx = std::max(xmx, x[i]);
                ymn = std::min(ymn, y[i]); ymx = std::max(ymx, y[i]);
            }
            p->xmin = xmn - 100 * p->dl;
            p->ymin = ymn - 10000 * p->dl;
            p->xmax = xmx + 100 * p->dl;
            p->ymax = ymx + 100 * p->dl;
        }
        p->pmax = 4.50 * (double)tp;                                            // main.cu:548
    }
    return VOX_OK;
}

const char* vox_main_input(const char* test, int threedim)
{
    const std::string t = test ? test : "";
    if (threedim && t != "oil" && t != "drop" && t != "shear" && t != "turb") return "input-dam3d.grd";
    if (t == "turb") return "input-turb.grd";                                   // main.cu:505-533
    if (t == "shear") return "input-shear.grd";
    if (t == "drop") return "input-drop.grd";
    if (t == "oil") return "input-oil.grd";
    return "input-dam.grd";
}

static int validate(const vox_params* p)
{
    if (p->dim != 2 && p->dim != 3) return fail(VOX_ERR_ARG, "dim must be 2 or 3");
    if (!(p->re_s > 0) || !(p->re_l > 0) || !(p->dl > 0) || !(p->lambda > 0))
        return fail(VOX_ERR_ARG, "radii, spacing and lambda must be positive");
    if (p->re_l < p->re_s) return fail(VOX_ERR_ARG, "re_l must be >= re_s");
    if (!(p->dt > 0)) return fail(VOX_ERR_ARG, "dt must be positive");
    if (!(p->xmax > p->xmin) || !(p->ymax > p->ymin) || (p->dim == 3 && !(p->zmax > p->zmin)))
        return fail(VOX_ERR_ARG, "grid box is empty: call vox_params_box");
    if (p->precision != VOX_PREC_FP64 && p->precision != VOX_PREC_MIXED)
        return fail(VOX_ERR_NOLIB, "unknown precision mode");
    if (p->solver != VOX_SOLVER_JACOBI_CG && p->solver != VOX_SOLVER_LEARNED_FCG)
        return fail(VOX_ERR_NOLIB, "unknown PPE solver");
    if (!(p->ir_eta >= 0.0 && p->ir_eta < 1.0) || p->ir_maxouter < 1 ||
        (p->warm_start != 0 && p->warm_start != 1))
        return fail(VOX_ERR_ARG, "ir_eta must be in [0, 1), ir_maxouter >= 1, warm_start 0 or 1");
    if (p->list_mode < 0 || p->list_mode > 2) return fail(VOX_ERR_ARG, "list_mode must be 0, 1 or 2");
    if (p->lpc_cap_pct < 1) return fail(VOX_ERR_ARG, "lpc_cap_pct must be >= 1");
    if (p->method != 1 && p->method != 3)
        return fail(VOX_ERR_NOLIB, "PRESSURECALC method must be 1 (PPE) or 3 (equation of state)");
    if ((p->matopt != 0 && p->matopt != 1) || (p->srcopt != 0 && p->srcopt != 1))
        return fail(VOX_ERR_ARG, "matopt and srcopt must be 0 or 1");
    if (p->ref_gpu != 0 && p->ref_gpu != 1) return fail(VOX_ERR_ARG, "ref_gpu must be 0 or 1");
    if (p->ref_gpu && (p->dim != 2 || p->precision != VOX_PREC_FP64))
        return fail(VOX_ERR_ARG, "ref_gpu (the reference's CUDA path) is 2D and FP64 only");
    if (p->ktype != 2 && p->ktype != 6 && p->ktype != 1 && p->ktype != 5)
        return fail(VOX_ERR_ARG, "unsupported kernel type");
    return VOX_OK;
}

int vox_create(vox_ctx** out, const vox_params* pin, int tp, int gp, int wp, int fp)
{
    if (!out || !pin) return fail(VOX_ERR_ARG, "null argument");
    *out = nullptr;
    if (tp <= 0 || gp < 0 || wp < 0 || fp < 0 || gp + wp + fp != tp)
        return fail(VOX_ERR_ARG, "particle counts must satisfy gp + wp + fp == tp > 0");
    VOX_TRY(validate(pin));
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(VOX_ERR_HIP, "no HIP device available");
    vox_ctx* c = new vox_ctx();
    c->p = *pin;
    if (c->p.pmax <= 0) c->p.pmax = (double)tp * 4.50;              // main.cu:548
    c->tp = tp; c->gp = gp; c->wp = wp; c->fp = fp;
    c->two_lists = (c->p.re_l != c->p.re_s);
    int rc = VOX_OK;
    auto bail = [&](int code) { vox_destroy(c); return code; };
    if (hipGetDevice(&c->device) != hipSuccess) return bail(fail(VOX_ERR_HIP, "hipGetDevice"));
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(VOX_ERR_HIP, "hipStreamCreate"));
    if (c->two_lists) {
        // the side stream of the neighbour build (default priority, like the main one)
        if (hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess)
            return bail(fail(VOX_ERR_HIP, "side stream of the neighbour build"));
    }
    const size_t n = (size_t)tp;
    for (int f = 0; f < VOX_F_COUNT; ++f) {
        if (f == VOX_F_PTYPE || f == VOX_F_BCON) continue;
        if ((rc = dalloc(&c->f[f], n))) return bail(rc);
    }
    if ((rc = dalloc(&c->ptype, n)) || (rc = dalloc(&c->bcon, n))) return bail(rc);
    for (int w = 0; w < 2; ++w) {
        if (w == 1 && !c->two_lists) break;
        DevList& L = c->list[w];
        if ((rc = dalloc(&L.cnt, n)) || (rc = dalloc(&L.cell, n)) || (rc = dalloc(&L.order_raw, n)) ||
            (rc = dalloc(&L.rec, n)) || (rc = dalloc(&L.pos_of, n)))
            return bail(rc);
        double re = (w == 0) ? c->p.re_s : c->p.re_l;
        int want = (w == 0 ? c->p.cap_s : c->p.cap_l);
        if (want <= 0)
            want = auto_cap(c->p.dim, c->p.dim == 3 || c->p.ref_gpu, c->p.dl, re, c->p.delta);
        if ((rc = ensure_list_storage(c, w, want))) return bail(rc);
    }
    for (int k = 0; k < 6; ++k)
        if ((rc = dalloc(&c->sps_s[k], n))) return bail(rc);
    if ((rc = dalloc(&c->scratch_d, n)) || (rc = dalloc(&c->pred_rec, n)) || (rc = dalloc(&c->pt_sorted, n)) ||
        (rc = dalloc(&c->pos_rec, n)) || (rc = dalloc(&c->blockmax, (size_t)grid_for(n))) ||
        (rc = dalloc(&c->blockmax2, (size_t)grid_for(n))))
        return bail(rc);
    if ((rc = dalloc(&c->uorder, n)) || (rc = dalloc(&c->uscan, n))) return bail(rc);
    if ((rc = dalloc(&c->uidx, n)) || (rc = dalloc(&c->urow, n)) || (rc = dalloc(&c->uflag, n)) ||
        (rc = dalloc(&c->rowlen, n)) || (rc = dalloc(&c->adiag, n)) || (rc = dalloc(&c->rhs, n)) ||
        (rc = dalloc(&c->cg_x, n)) || (rc = dalloc(&c->cg_r, n)) || (rc = dalloc(&c->cg_p, n)) ||
        (rc = dalloc(&c->cg_q, n)) || (rc = dalloc(&c->cg_x32, n)) || (rc = dalloc(&c->cg_r32, n)) ||
        (rc = dalloc(&c->cg_p32, n)) || (rc = dalloc(&c->cg_q32, n)) || (rc = dalloc(&c->cg_d, n)) ||
        (rc = dalloc(&c->cg_e, n)))
        return bail(rc);
    if ((rc = dalloc(&c->coll_cnt, n)) || (rc = dalloc(&c->coll_off, n + 1)) ||
        (rc = dalloc(&c->coll_stamp, n)) || (rc = dalloc(&c->coll_root, n)) || (rc = dalloc(&c->coll_logidx, n)) ||
        (rc = dalloc(&c->coll_act, n)) || (rc = dalloc(&c->coll_cnt2, n)) || (rc = dalloc(&c->coll_off2, n + 1)) ||
        (rc = dalloc(&c->coll_logn, 1)))
        return bail(rc);
    if ((rc = dalloc(&c->partials, kPartialSlots * (size_t)kMaxPartials)) || (rc = dalloc(&c->dscal, SC_COUNT)) ||
        (rc = dalloc(&c->dflags, FLAG_COUNT)))
        return bail(rc);
    if (hipHostMalloc((void**)&c->hflags, sizeof(int32_t) * FLAG_COUNT, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&c->hscal, sizeof(double) * SC_COUNT, hipHostMallocDefault) != hipSuccess)
        return bail(fail(VOX_ERR_HIP, "hipHostMalloc"));
    std::memset(c->hflags, 0, sizeof(int32_t) * FLAG_COUNT);
    std::memset(c->hscal, 0, sizeof(double) * SC_COUNT);
    if (hipDeviceSynchronize() != hipSuccess) return bail(fail(VOX_ERR_HIP, "initial sync"));
    *out = c;
    return VOX_OK;
}

int vox_destroy(vox_ctx* c)
{
    if (!c) return VOX_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (int f = 0; f < VOX_F_COUNT; ++f) dfree(c->f[f]);
    dfree(c->ptype); dfree(c->bcon);
    for (int w = 0; w < 2; ++w) {
        DevList& L = c->list[w];
        dfree(L.cnt); dfree(L.ids); dfree(L.cell); dfree(L.cell_start); dfree(L.cell_fill);
        dfree(L.order_raw); dfree(L.rec); dfree(L.pos_of);
    }
    for (int k = 0; k < 6; ++k) dfree(c->sps_s[k]);
    dfree(c->scratch_d); dfree(c->pred_rec); dfree(c->pt_sorted); dfree(c->pos_rec); dfree(c->blockmax); dfree(c->blockmax2);
    dfree(c->uidx); dfree(c->urow); dfree(c->uflag); dfree(c->uorder); dfree(c->uscan); dfree(c->zcell); dfree(c->rowlen); dfree(c->acol); dfree(c->dcol); dfree(c->cbase); dfree(c->wcol); dfree(c->wuniq); dfree(c->wn);
    dfree(c->aval); dfree(c->adiag); dfree(c->rhs);
    dfree(c->cg_x); dfree(c->cg_r); dfree(c->cg_p); dfree(c->cg_q); dfree(c->cg_d); dfree(c->cg_e);
    dfree(c->cg_x32); dfree(c->cg_r32); dfree(c->cg_p32); dfree(c->cg_q32); dfree(c->aval32);
    dfree(c->coll_cnt); dfree(c->coll_off); dfree(c->coll_i); dfree(c->coll_j);
    dfree(c->coll_act); dfree(c->coll_cnt2); dfree(c->coll_off2); dfree(c->coll_i2); dfree(c->coll_j2);
    dfree(c->coll_stamp); dfree(c->coll_root); dfree(c->region); dfree(c->rbox); dfree(c->coll_logidx); dfree(c->coll_log); dfree(c->coll_logn);
    dfree(c->scan_blocks); dfree(c->scan_blocks2); dfree(c->partials); dfree(c->dscal); dfree(c->dflags);
    lpc_free_ctx(c);
    if (c->hflags) (void)hipHostFree(c->hflags);
    if (c->hscal) (void)hipHostFree(c->hscal);
    for (auto& r : c->prof_pending) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
    for (auto e : c->prof_free) (void)hipEventDestroy(e);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return VOX_OK;
}

int vox_get_params(vox_ctx* c, vox_params* p)
{
    VOX_TRY(check_ctx(c));
    if (!p) return fail(VOX_ERR_ARG, "null params");
    *p = c->p;
    return VOX_OK;
}

int vox_set_params(vox_ctx* c, const vox_params* p)
{
    VOX_TRY(check_ctx(c));
    if (!p) return fail(VOX_ERR_ARG, "null params");
    if (p->dim != c->p.dim || p->re_s != c->p.re_s || p->re_l != c->p.re_l || p->dl != c->p.dl ||
        p->delta != c->p.delta || p->xmin != c->p.xmin || p->xmax != c->p.xmax ||
        p->ymin != c->p.ymin || p->ymax != c->p.ymax || p->zmin != c->p.zmin || p->zmax != c->p.zmax)
        return fail(VOX_ERR_ARG, "dim, spacing, radii and grid box are fixed at create");
    VOX_TRY(validate(p));
    double pmax = c->p.pmax;
    c->p = *p;
    if (c->p.pmax <= 0) c->p.pmax = pmax;
    c->meu_uniform = false;      // MEU was written under the previous parameters
    return VOX_OK;
}

static bool int_field(int f) { return f == VOX_F_PTYPE || f == VOX_F_BCON; }

int vox_upload(vox_ctx* c, int field, const void* host)
{
    VOX_TRY(check_ctx(c));
    if (!host || field < 0 || field >= VOX_F_COUNT) return fail(VOX_ERR_ARG, "bad field upload");
    void* dst = field == VOX_F_PTYPE ? (void*)c->ptype : field == VOX_F_BCON ? (void*)c->bcon
                                                                             : (void*)c->f[field];
    size_t bytes = (int_field(field) ? sizeof(int32_t) : sizeof(double)) * (size_t)c->tp;
    VOX_HIP(hipMemcpyAsync(dst, host, bytes, hipMemcpyHostToDevice, c->stream));
    VOX_HIP(hipStreamSynchronize(c->stream));
    if (field == VOX_F_MEU) c->meu_uniform = false;
    if (field == VOX_F_X || field == VOX_F_Y || field == VOX_F_Z) ++c->pos_epoch;
    if (field == VOX_F_PTYPE) {          // types never change during a run
        const int32_t* t = static_cast<const int32_t*>(host);
        long n2 = 0;
        for (int i = 0; i < c->tp; ++i) n2 += (t[i] == 2);
        c->n_fluid2 = n2;
    }
    return VOX_OK;
}

int vox_download(vox_ctx* c, int field, void* host)
{
    VOX_TRY(check_ctx(c));
    if (!host || field < 0 || field >= VOX_F_COUNT) return fail(VOX_ERR_ARG, "bad field download");
    const void* src = field == VOX_F_PTYPE ? (const void*)c->ptype
                      : field == VOX_F_BCON ? (const void*)c->bcon
                                            : (const void*)c->f[field];
    size_t bytes = (int_field(field) ? sizeof(int32_t) : sizeof(double)) * (size_t)c->tp;
    VOX_HIP(hipMemcpyAsync(host, src, bytes, hipMemcpyDeviceToHost, c->stream));
    VOX_HIP(hipStreamSynchronize(c->stream));
    return VOX_OK;
}

int vox_upload_state(vox_ctx* c, const double* x, const double* y, const double* z,
                     const double* u, const double* v, const double* w, const double* p,
                     const int32_t* ptype)
{
    VOX_TRY(check_ctx(c));
    if (!x || !y || !u || !v || !p || !ptype) return fail(VOX_ERR_ARG, "null state array");
    VOX_TRY(vox_upload(c, VOX_F_X, x));
    VOX_TRY(vox_upload(c, VOX_F_Y, y));
    VOX_TRY(vox_upload(c, VOX_F_U, u));
    VOX_TRY(vox_upload(c, VOX_F_V, v));
    VOX_TRY(vox_upload(c, VOX_F_P, p));
    VOX_TRY(vox_upload(c, VOX_F_PTYPE, ptype));
    if (c->p.dim == 3) {
        if (!z || !w) return fail(VOX_ERR_ARG, "3D state needs z and w");
        VOX_TRY(vox_upload(c, VOX_F_Z, z));
        VOX_TRY(vox_upload(c, VOX_F_W, w));
    }
    return VOX_OK;
}

int vox_density(vox_ctx* c) { VOX_TRY(check_ctx(c)); return st_density(c); }

int vox_neighbor(vox_ctx* c)
{
    VOX_TRY(check_ctx(c));
    return build_neighbors(c);
}

int vox_pnum(vox_ctx* c, int loopstarted)
{
    VOX_TRY(check_ctx(c));
    VOX_TRY(st_pnd(c, 0, loopstarted));
    if (!loopstarted) VOX_TRY(st_n0(c));
    return VOX_OK;
}

int vox_pnumstar(vox_ctx* c) { VOX_TRY(need_init(c)); return st_pnd(c, 1, 1); }
int vox_sps(vox_ctx* c) { VOX_TRY(need_init(c)); return st_sps(c); }
int vox_vfraction(vox_ctx* c) { VOX_TRY(need_init(c)); return st_vfraction(c); }
int vox_viscosity(vox_ctx* c) { VOX_TRY(need_init(c)); return st_viscosity(c); }
int vox_prediction(vox_ctx* c) { VOX_TRY(need_init(c)); return st_prediction(c); }
int vox_bcon(vox_ctx* c) { VOX_TRY(need_init(c)); return st_bcon(c); }
int vox_ppe_assemble(vox_ctx* c) { VOX_TRY(need_init(c)); return ppe_assemble(c); }

int vox_ppe_solve(vox_ctx* c, vox_solve_info* info)
{
    VOX_TRY(need_init(c));
    return ppe_solve(c, info);
}

int vox_pressurecalc(vox_ctx* c, vox_solve_info* info)
{
    VOX_TRY(need_init(c));
    if (c->p.method == 3) {                                 // functions.cu:2471-2500
        if (info) *info = vox_solve_info{};
        return st_pressure_wc(c);
    }
    VOX_TRY(fill_f64(c, c->f[VOX_F_PNEW], 0.0, c->tp));    // functions.cu:2503
    VOX_TRY(ppe_assemble(c));
    int rc = ppe_solve(c, info);
    VOX_TRY(st_clamp_pressure(c));
    return rc;
}

int vox_phat(vox_ctx* c) { VOX_TRY(need_init(c)); return st_phat(c); }
int vox_pressgrad(vox_ctx* c) { VOX_TRY(need_init(c)); return st_pressgrad(c); }
int vox_euler(vox_ctx* c) { VOX_TRY(need_init(c)); return st_euler(c); }
int vox_collision(vox_ctx* c) { VOX_TRY(need_init(c)); return st_collision(c); }

int vox_prepdata(vox_ctx* c)
{
    VOX_TRY(need_init(c));
    VOX_TRY(st_prepdata(c));
    VOX_TRY(read_flags(c));
    c->st.nan_count = c->hflags[FLAG_NAN];
    c->st.out_of_box = c->hflags[FLAG_OOB];
    return VOX_OK;
}

int vox_dtcalc(vox_ctx* c)
{
    VOX_TRY(check_ctx(c));
    vox_params& p = c->p;                                   // functions.cu:1684-1698
    double c0 = (p.c01 > p.c02) ? p.c01 : p.c02;
    double courant = p.courant * p.dl / c0;
    if (p.dt > courant) p.dt = courant;
    else if (courant < p.dt_max) p.dt = courant;
    else p.dt = p.dt_max;
    return VOX_OK;
}

int vox_init(vox_ctx* c)
{
    VOX_TRY(check_ctx(c));
    VOX_TRY(st_density(c));                                 // main.cu:966-978
    VOX_TRY(build_neighbors(c));                            // main.cu:987-995
    VOX_TRY(st_pnd(c, 0, 0));                               // main.cu:1000-1010
    VOX_TRY(st_n0(c));                                      // main.cu:1011
    c->initialised = true;
    c->st.steps = 0;
    c->st.time = 0.0;
    return VOX_OK;
}

int vox_step(vox_ctx* c, int nsteps)
{
    VOX_TRY(need_init(c));
    for (int s = 0; s < nsteps; ++s) {
        if (c->p.dt_mode == 1) VOX_TRY(vox_dtcalc(c));      // main.cu:1030
        const bool wc = (c->p.method == 3);
        if (!wc) VOX_TRY(st_pnd(c, 0, 1));                  // 1036-1041 (previous list)
        VOX_TRY(build_neighbors(c));                        // 1050-1055
        if (c->p.turb) VOX_TRY(st_sps(c));                  // 1066-1071
        VOX_TRY(st_vfraction(c));                           // 1084-1087
        VOX_TRY(st_viscosity(c));                           // 1096-1099
        VOX_TRY(st_prediction(c));                          // 1108-1111
        VOX_TRY(st_pnd(c, 1, 1));                           // 1121-1124
        if (wc) {
            VOX_TRY(st_pressure_wc(c));                     // 1149-1151 -> 2471-2500
        } else {
            VOX_TRY(st_bcon(c));                            // 1133-1138
            VOX_TRY(fill_f64(c, c->f[VOX_F_PNEW], 0.0, c->tp)); // 1147-1151 -> 2503
            VOX_TRY(ppe_assemble(c));
            VOX_TRY(ppe_solve(c, nullptr));
            VOX_TRY(st_clamp_pressure(c));
        }
        VOX_TRY(st_phat(c));                                // 1267-1270
        VOX_TRY(st_pressgrad(c));                           // 1279-1282
        VOX_TRY(st_euler(c));                               // 1291-1294
        VOX_TRY(st_collision(c));                           // 1303-1306
        VOX_TRY(st_prepdata(c));                            // 1315-1318
        c->st.steps += 1;
        c->st.time += c->p.dt;
    }
    VOX_TRY(read_flags(c));
    c->st.nan_count = c->hflags[FLAG_NAN];
    c->st.out_of_box = c->hflags[FLAG_OOB];
    return VOX_OK;
}

int vox_sync(vox_ctx* c)
{
    VOX_TRY(check_ctx(c));
    VOX_HIP(hipStreamSynchronize(c->stream));
    return VOX_OK;
}

int vox_get_status(vox_ctx* c, vox_status* st)
{
    VOX_TRY(check_ctx(c));
    if (!st) return fail(VOX_ERR_ARG, "null status");
    *st = c->st;
    st->n0_s = c->n0s;
    st->n0_l = c->n0l;
    return VOX_OK;
}

int vox_get_n0(vox_ctx* c, double* n0s, double* n0l)
{
    VOX_TRY(check_ctx(c));
    if (n0s) *n0s = c->n0s;
    if (n0l) *n0l = c->n0l;
    return VOX_OK;
}

int vox_set_regions(vox_ctx* c, int nregion, const int32_t* region, const double* boxes)
{
    VOX_TRY(check_ctx(c));
    if (c->region) (void)hipFree(c->region);
    if (c->rbox) (void)hipFree(c->rbox);
    c->region = nullptr;
    c->rbox = nullptr;
    c->nregion = 0;
    if (nregion == 0) return VOX_OK;
    if (nregion < 0 || !region || !boxes) return fail(VOX_ERR_ARG, "bad region arguments");
    for (int i = 0; i < c->tp; ++i)
        if (region[i] < 0 || region[i] >= nregion) return fail(VOX_ERR_ARG, "region id out of range");
    VOX_HIP(hipMalloc(&c->region, sizeof(int32_t) * c->tp));
    VOX_HIP(hipMalloc(&c->rbox, sizeof(double) * 6 * (size_t)nregion));
    VOX_HIP(hipMemcpy(c->region, region, sizeof(int32_t) * c->tp, hipMemcpyHostToDevice));
    VOX_HIP(hipMemcpy(c->rbox, boxes, sizeof(double) * 6 * (size_t)nregion, hipMemcpyHostToDevice));
    c->nregion = nregion;
    return VOX_OK;
}

int vox_autotune(vox_ctx* c, int reps, int32_t* modes)
{
    VOX_TRY(need_init(c));
    VOX_TRY(autotune_lists(c, reps, modes));
    return autotune_kernels(c, reps);
}

int vox_get_tuning(vox_ctx* c, int32_t* choice)
{
    VOX_TRY(check_ctx(c));
    if (!choice) return fail(VOX_ERR_ARG, "null tuning array");
    choice[0] = c->tuned_mode[0];
    choice[1] = c->tuned_mode[1];
    choice[2] = c->tune_pred;
    choice[3] = c->tune_spmv;
    return VOX_OK;
}

int vox_set_tuning(vox_ctx* c, const int32_t* choice)
{
    VOX_TRY(check_ctx(c));
    if (!choice) return fail(VOX_ERR_ARG, "null tuning array");
    if (choice[0] < 0 || choice[0] > 2 || choice[1] < 0 || choice[1] > 2 || choice[2] < 0 ||
        choice[2] >= kPredVariants || choice[3] < 0 || choice[3] >= kSpmvVariants)
        return fail(VOX_ERR_ARG, "tuning choice out of range");
    c->tuned_mode[0] = choice[0];
    c->tuned_mode[1] = choice[1];
    c->tune_pred = choice[2];
    c->tune_spmv = choice[3];
    return VOX_OK;
}

int vox_set_n0(vox_ctx* c, double n0s, double n0l)
{
    VOX_TRY(check_ctx(c));
    c->n0s = n0s;
    c->n0l = n0l;
    return VOX_OK;
}

int vox_get_neighbors(vox_ctx* c, int which, int32_t* counts, int32_t* ids, int stride)
{
    VOX_TRY(check_ctx(c));
    if (which != VOX_LIST_S && which != VOX_LIST_L) return fail(VOX_ERR_ARG, "bad list id");
    if (which == VOX_LIST_L && c->two_lists) VOX_TRY(ensure_reference_list(c, 1));
    const DevList& L = list_of(c, which);
    const size_t n = (size_t)c->tp;
    std::vector<int32_t> cnt(n);
    VOX_HIP(hipMemcpy(cnt.data(), L.cnt, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    int mx = 0;
    for (size_t i = 0; i < n; ++i) mx = std::max(mx, (int)cnt[i]);
    if (counts) std::memcpy(counts, cnt.data(), sizeof(int32_t) * n);
    if (!ids) return VOX_OK;
    if (stride < mx) return fail(VOX_ERR_OVERFLOW, "stride smaller than the longest row (" +
                                                       std::to_string(mx) + ")");
    if (mx == 0) return VOX_OK;
    std::vector<int32_t> sm((size_t)sl_size((long)n, L.cap));
    VOX_HIP(hipMemcpy(sm.data(), L.ids, sizeof(int32_t) * sm.size(), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i)
        for (int l = 0; l < cnt[i]; ++l) ids[i * stride + l] = sm[sl_index((long)i, l, L.cap)];
    return VOX_OK;
}

int vox_get_neighbors_slot(vox_ctx* c, int which, int32_t* neighb, int stride)
{
    VOX_TRY(check_ctx(c));
    if (!neighb) return fail(VOX_ERR_ARG, "null buffer");
    const size_t n = (size_t)c->tp;
    std::vector<int32_t> cnt(n);
    VOX_TRY(vox_get_neighbors(c, which, cnt.data(), nullptr, 0));
    int mx = 0;
    for (size_t i = 0; i < n; ++i) mx = std::max(mx, (int)cnt[i]);
    if (stride < mx + 2) return fail(VOX_ERR_OVERFLOW, "slot stride must exceed the longest row + 1");
    std::vector<int32_t> rows(n * (size_t)(mx > 0 ? mx : 1));
    VOX_TRY(vox_get_neighbors(c, which, nullptr, rows.data(), mx > 0 ? mx : 1));
    std::memset(neighb, 0, sizeof(int32_t) * (n + 1) * stride);
    for (size_t i = 0; i < n; ++i) {
        int32_t* r = neighb + (i + 1) * stride;           // 1-based row I = i+1
        r[1] = cnt[i] + 1;                                // functions.cu:2919-2920
        for (int l = 0; l < cnt[i]; ++l) r[2 + l] = rows[i * (mx > 0 ? mx : 1) + l] + 1;
    }
    return VOX_OK;
}

int vox_get_cells(vox_ctx* c, int which, int32_t* cell)
{
    VOX_TRY(check_ctx(c));
    if (!cell || (which != VOX_LIST_S && which != VOX_LIST_L)) return fail(VOX_ERR_ARG, "bad args");
    VOX_HIP(hipMemcpy(cell, list_of(c, which).cell, sizeof(int32_t) * c->tp, hipMemcpyDeviceToHost));
    return VOX_OK;
}

int vox_ppe_export(vox_ctx* c, int32_t* nu, int32_t* nnz, int32_t* rowptr, int32_t* cols,
                   double* vals, double* rhs, int32_t* ids)
{
    VOX_TRY(need_init(c));
    VOX_HIP(hipStreamSynchronize(c->stream));
    return ppe_export(c, nu, nnz, rowptr, cols, vals, rhs, ids);
}

int vox_ppe_load(vox_ctx* c, int32_t nu, int32_t nnz, const int32_t* rowptr, const int32_t* cols,
                 const double* vals, const double* rhs, const int32_t* ids)
{
    VOX_TRY(need_init(c));
    return ppe_load(c, nu, nnz, rowptr, cols, vals, rhs, ids);
}

int vox_prof_enable(vox_ctx* c, int on)
{
    VOX_TRY(check_ctx(c));
    if (!on) VOX_TRY(prof_drain(c));
    c->prof_on = (on != 0);
    return VOX_OK;
}

int vox_prof_only(vox_ctx* c, const char* kernel)
{
    VOX_TRY(check_ctx(c));
    if (!kernel) { c->prof_only = -1; return VOX_OK; }
    int k = kclass_from_name(kernel);
    if (k < 0) return fail(VOX_ERR_ARG, std::string("unknown kernel class ") + kernel);
    c->prof_only = k;
    return VOX_OK;
}

int vox_prof_get(vox_ctx* c, const char* kernel, double* total_ms, int64_t* launches)
{
    VOX_TRY(check_ctx(c));
    if (!kernel) return fail(VOX_ERR_ARG, "null kernel name");
    int k = kclass_from_name(kernel);
    if (k < 0) return fail(VOX_ERR_ARG, std::string("unknown kernel class ") + kernel);
    VOX_TRY(prof_drain(c));
    if (total_ms) *total_ms = c->prof_ms[k];
    if (launches) *launches = c->prof_n[k];
    return VOX_OK;
}

int vox_prof_reset(vox_ctx* c)
{
    VOX_TRY(check_ctx(c));
    VOX_TRY(prof_drain(c));
    for (int k = 0; k < 32; ++k) { c->prof_ms[k] = 0.0; c->prof_n[k] = 0; }
    return VOX_OK;
}

}  // extern "C"
