/*
 * voxarmps.h — C ABI of the MI355X-native MPS time-step pipeline.
 *
 * This boundary replaces the per-stage host routines of the reference
 * (VoxarMPS/VoxarMPS/functions.cuh:56-111, neighbour_parallel.h:6-7) and the
 * kernel-launch branch of its time loop (VoxarMPS/VoxarMPS/main.cu:1027-1363).
 * The reference routines take 1-based host arrays of length TP+1 and a
 * std::string mode argument; here the state lives in device memory owned by a
 * context, host arrays are 0-based with TP elements (a reference caller passes
 * `x + 1`), and every entry point returns a status code (VOX_OK == 0) with a
 * message available from vox_last_error().
 *
 * Plain C types only: no torch, no HIP types in any signature.
 */
#ifndef VOXARMPS_H
#define VOXARMPS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define VOX_OK               0
#define VOX_ERR_ARG          1  /* invalid argument or unsupported option      */
#define VOX_ERR_HIP          2  /* HIP runtime failure (allocation, launch)    */
#define VOX_ERR_OVERFLOW     3  /* neighbour row exceeded its capacity         */
#define VOX_ERR_DOMAIN       4  /* particle outside the neighbour grid box     */
#define VOX_ERR_SOLVER       5  /* PPE solver breakdown / no convergence       */
#define VOX_ERR_NOLIB        6  /* requested path not built                    */
#define VOX_ERR_STATE        7  /* call order violated (e.g. step before init) */

/* ---- field identifiers (names follow main.cu:211-232, 554-770) ---------- */
enum vox_field {
    VOX_F_X = 0, VOX_F_Y, VOX_F_Z,            /* positions                  */
    VOX_F_U, VOX_F_V, VOX_F_W,                /* velocities                 */
    VOX_F_P,                                  /* pressure                   */
    VOX_F_PTYPE,                              /* int32: -1 dummy, 0 wall, 1/2 fluid */
    VOX_F_XSTAR, VOX_F_YSTAR, VOX_F_ZSTAR,    /* predicted positions        */
    VOX_F_USTAR, VOX_F_VSTAR, VOX_F_WSTAR,    /* predicted velocities       */
    VOX_F_UNEW, VOX_F_VNEW, VOX_F_WNEW,       /* corrected velocities       */
    VOX_F_PNEW,                               /* solved pressure            */
    VOX_F_PHAT,                               /* minimum neighbour pressure */
    VOX_F_N, VOX_F_NSTAR,                     /* PND at x and at x*         */
    VOX_F_NEUT,                               /* SPS eddy viscosity         */
    VOX_F_TURB1, VOX_F_TURB2, VOX_F_TURB3,    /* SPS stress divergence      */
    VOX_F_C,                                  /* volume fraction of fluid 2 */
    VOX_F_MEU,                                /* dynamic viscosity          */
    VOX_F_RHO,                                /* density                    */
    VOX_F_BCON,                               /* int32: -1 / 0 / 1          */
    VOX_F_SOURCE,                             /* PPE right-hand side        */
    VOX_F_COUNT
};

/* neighbour list selector */
#define VOX_LIST_S 0   /* radius re_s: PND, gradient, SPS, collision, p-hat */
#define VOX_LIST_L 1   /* radius re_l: Laplacian (viscous term, PPE matrix)  */

/* precision modes */
#define VOX_PREC_FP64  0   /* every kernel in FP64                           */
#define VOX_PREC_MIXED 1   /* FP32 particle kernels + SpMV, FP64 dots + IR   */

/* PPE solvers */
#define VOX_SOLVER_JACOBI_CG   0
#define VOX_SOLVER_LEARNED_FCG 1  /* flexible CG with the learned preconditioner of an
                                     attached vox_lpc; Jacobi-CG is the fallback */

/* boundary-clamp variants of BOUNDARIE (functions.cu:1963-2002) */
#define VOX_TEST_DAM  0
#define VOX_TEST_TURB 1
#define VOX_TEST_DROP 2

/*
 * Every constant main.cu hard-codes (main.cu:211-494), plus the two-radius,
 * time-step, solver and precision extensions. Fill with vox_params_default()
 * and override. Only int32 and double members, in this order (ctypes mirrors it).
 */
typedef struct vox_params {
    int32_t dim;          /* 2 or 3 (main.cu:309-317)                            */
    int32_t ktype;        /* weight function: 2 rational, 6 cubic (main.cu:288-292) */
    double  dl;           /* particle spacing DL (main.cu:248-250)               */
    double  re_s;         /* radius of PND / gradient stages                    */
    double  re_l;         /* radius of Laplacian stages (== re_s: reference)     */
    double  delta;        /* grid expansion DELTA = DL/5 (main.cu:282)           */
    double  lambda;       /* Laplacian coefficient (main.cu:491-494)             */
    double  rho1, nu1, rho2, nu2;             /* fluid properties (main.cu:244-247) */
    double  gx, gy, gz;                       /* gravity (main.cu:321-330)       */
    double  dt;           /* time step                                           */
    int32_t dt_mode;      /* 0 fixed dt, 1 DTcalculation (functions.cu:1684-1698) */
    int32_t turb;         /* SPS-LES on (main.cu:294-303)                        */
    double  courant, c01, c02, dt_max;        /* DTcalculation inputs             */
    double  dirichlet;    /* free-surface threshold on nstar/n0 (main.cu:228)   */
    double  relaxp;       /* pressure relaxation (main.cu:263-265)               */
    double  coll;         /* collision distance (main.cu:269-273)                */
    double  cc;           /* collision coefficient (main.cu:277-279)             */
    double  vel_cor;      /* BOUNDARIE velocity factor (main.cu:275, 506)        */
    double  cs;           /* Smagorinsky constant (main.cu:281)                  */
    double  ncorrection;  /* PND correction for ktype != 2 (main.cu:284)         */
    double  vmax;         /* damper speed (main.cu:331)                          */
    double  pmin, pmax;   /* pressure clamp (main.cu:351, 548)                   */
    int32_t fluid2_type;  /* 0 Newtonian, 1 Herschel-Bulkley (main.cu:357-377)   */
    int32_t kh_correction;/* 1: symmetric p-hat gradient (main.cu:344)           */
    double  n_pl, meu0, phi, cohes;           /* rheology (main.cu:358-372)       */
    double  xmin, xmax, ymin, ymax, zmin, zmax; /* grid box (main.cu:692-715)    */
    int32_t test_kind;    /* VOX_TEST_* (BOUNDARIE / PREPDATA variants)          */
    int32_t coef2d;       /* 1: keep the 2D constants 4 and 2 of PREDICTION /
                             PRESSGRAD in 3D as the reference does              */
    int32_t precision;    /* VOX_PREC_*                                          */
    int32_t solver;       /* VOX_SOLVER_*                                        */
    double  cg_tol;       /* relative residual target ||r||/||b||               */
    int32_t cg_maxit;     /* iteration cap                                       */
    int32_t cg_check;     /* iterations between host convergence checks         */
    int32_t cap_s;        /* neighbour row capacity for list S (0 = auto)        */
    int32_t cap_l;        /* neighbour row capacity for list L (0 = auto)        */
    double  ir_eta;       /* mixed precision: relative tolerance of each FP32
                             inner solve of the FP64 refinement (0 = 1e-4)       */
    int32_t ir_maxouter;  /* refinement steps cap (mixed precision, warm start)  */
    int32_t warm_start;   /* 1: the PPE solve starts from each unknown's pressure
                             of the previous step (field P) and refines in FP64;
                             0: from zero, as the reference (functions.cu:2503) */
    int32_t list_mode;    /* neighbour search: 0 auto, 1 one thread per particle,
                             2 one wave per 64 cell-sorted particles (same
                             lists either way)                                 */
    int32_t lpc_cap_pct;  /* learned FCG: iteration cap in percent of the Jacobi-CG
                             baseline before the fallback (200: twice)         */
    int32_t method;       /* PRESSURECALC: 1 PPE (incompressible), 3 weakly
                             compressible equation of state (main.cu:335-339)  */
    int32_t matopt;       /* 1: higher-order Laplacian in MATRIX (functions.cu:901-914) */
    int32_t srcopt;       /* 1: higher-order source with error compensation
                             (functions.cu:928-988)                              */
    int32_t ref_gpu;      /* 1: the reference's CUDA path instead of its CPU path
                             (2D, FP64): neighbour cells and radius re+DELTA with
                             the particle itself in its list
                             (neighbour_parallel.cu:37-72, 145-199); MATRIX on
                             x* and n* (functions.cu:1007-1133, main.cu:1165);
                             gradient with 2 p^_i (382-385); BOUNDARIE with the
                             collision factor cc and Ymax - DL/2 (2023-2047);
                             PREPDATA copies u of every particle (2748-2782)   */
    double  wc_beta;      /* method 3: free-surface PND ratio of the equation of
                             state (BETA, main.cu:262)                           */
} vox_params;

/* PPE solve report */
typedef struct vox_solve_info {
    int32_t iterations;   /* CG iterations taken                                 */
    int32_t converged;    /* 1 if ||r||/||b|| <= tol                              */
    int32_t unknowns;     /* rows with bcon == 0                                 */
    int32_t nnz;          /* stored off-diagonal + diagonal entries              */
    double  relres;       /* final ||r||/||b|| (FP64)                            */
    double  bnorm;        /* ||b||                                               */
} vox_solve_info;

/* per-step status counters (PREPDATA checks, functions.cu:2718-2743) */
typedef struct vox_status {
    int32_t nan_count;        /* ERROR#1 particles                               */
    int32_t out_of_box;       /* ERROR#2 particles                               */
    int32_t max_count_s;      /* longest row of list S                           */
    int32_t max_count_l;      /* longest row of list L                           */
    int32_t collisions;       /* applied collision impulses in the last step     */
    int32_t collision_cand;   /* candidate pairs screened by the exact pass      */
    int32_t steps;            /* steps completed since vox_init                  */
    int32_t last_iters;       /* CG iterations of the last solve                 */
    double  last_relres;      /* residual of the last solve                      */
    double  n0_s, n0_l;       /* reference densities                             */
    double  time;             /* simulated time                                  */
    int32_t lpc_fallbacks;    /* learned solves that fell back to Jacobi-CG      */
    int32_t lpc_baseline;     /* Jacobi-CG iterations the learned solver is held to */
    int32_t product_col_bytes;/* bytes per column the last PPE product read: 2
                                 (16-bit steps or window ranks) or 4           */
    int32_t product_window_cols; /* columns the last product staged in its LDS
                                 windows, summed over its blocks (0: none)     */
} vox_status;

typedef struct vox_ctx vox_ctx;

/* ---- library / context --------------------------------------------------- */
const char* vox_last_error(void);
const char* vox_version(void);
/* sizeof of the three ABI structs, so a binding can check its mirror. */
int vox_abi_sizes(int32_t* params, int32_t* solve_info, int32_t* status);
int vox_device_count(int* count);

/* Reference defaults for a fully incompressible run (main.cu:211-494). */
int vox_params_default(vox_params* p, int dim, double dl);
/* Grid box from positions: min/max over particles expanded by DL (main.cu:692-715). */
int vox_params_box(vox_params* p, int tp, const double* x, const double* y, const double* z);

/* The run main.cu sets up from its 16 command-line arguments
 * (main.cu:65-209): test ("dam", "drop", "shear", "turb", "oil"), the
 * viscoplastic / turbulence / compressibility / matopt / srcopt switches,
 * code_opt ("sequential" or "openmp": the reference's CPU path; "gpu": its
 * CUDA path, ref_gpu = 1), 3D (the dam test only), the two fluids and the
 * time step (at most 0.005 s). Sets the constants those select
 * (main.cu:211-494: spacing, radius, kernel, lambda, relaxation, collision,
 * gravity, rheology, method, DTcalculation every step), the box of the tp
 * input positions (main.cu:692-715, widened for the drop test) and PMAX =
 * 4.5 tp (main.cu:548). The drop test's velocity override u = -x, v = y
 * (main.cu:645-651) is the caller's, on the input arrays. */
int vox_params_main(vox_params* p, const char* test, int visc, int turb, int comp, int matopt,
                    int srcopt, const char* code_opt, int threedim, double rho1, double nu1,
                    double rho2, double nu2, double dt, int tp, const double* x, const double* y,
                    const double* z);
/* Input file name main.cu opens for a test (main.cu:505-533). */
const char* vox_main_input(const char* test, int threedim);

/* Allocate device state for tp particles ordered ghost -> wall -> fluid
 * (gp, wp, fp counts; main.cu:635-682). */
int vox_create(vox_ctx** ctx, const vox_params* p, int tp, int gp, int wp, int fp);
int vox_destroy(vox_ctx* ctx);
int vox_get_params(vox_ctx* ctx, vox_params* p);
/* Change run-time parameters (dt, solver tolerances, radii unchanged). */
int vox_set_params(vox_ctx* ctx, const vox_params* p);

/* Host <-> device copies of one field, 0-based, tp elements (int32 for PTYPE/BCON). */
int vox_upload(vox_ctx* ctx, int field, const void* host);
int vox_download(vox_ctx* ctx, int field, void* host);
/* Upload the initial state read from a .grd file (x, y, z may be NULL in 2D for z). */
int vox_upload_state(vox_ctx* ctx, const double* x, const double* y, const double* z,
                     const double* u, const double* v, const double* w,
                     const double* p, const int32_t* ptype);

/* ---- stage entry points (one per reference host routine) ----------------- */
/* Density by particle type (main.cu:966-978). */
int vox_density(vox_ctx* ctx);
/* NEIGHBOR (functions.cuh:74, functions.cu:2788-2929): rebuild lists S and L. */
int vox_neighbor(vox_ctx* ctx);
/* PNUM (functions.cuh:56, functions.cu:94-123). loopstarted == 0 also copies
 * p -> pnew and sets n0 = max n (and n0_l on list L). */
int vox_pnum(vox_ctx* ctx, int loopstarted);
/* PNUMSTAR (functions.cuh:58, functions.cu:157-180). */
int vox_pnumstar(vox_ctx* ctx);
/* SPS (functions.cuh:80, functions.cu:1719-1841). */
int vox_sps(vox_ctx* ctx);
/* V_FRACTION, method 2 (functions.cuh:105, functions.cu:2580-2638). */
int vox_vfraction(vox_ctx* ctx);
/* VISCOSITY (functions.cuh:88, functions.cu:2053-2184). */
int vox_viscosity(vox_ctx* ctx);
/* PREDICTION (functions.cuh:91, functions.cu:2283-2369). */
int vox_prediction(vox_ctx* ctx);
/* BCON (functions.cuh:97, functions.cu:433-445). */
int vox_bcon(vox_ctx* ctx);
/* MATRIX, Method 1 (functions.cuh:62, functions.cu:869-1005), with the
 * higher-order Laplacian (matopt) and source (srcopt) when set: device matrix
 * over the bcon == 0 rows plus right-hand side. */
int vox_ppe_assemble(vox_ctx* ctx);
/* Solve the assembled PPE into pnew from a zero start (replaces INCDECOM + CGM,
 * functions.cu:1139-1200, 506-716). */
int vox_ppe_solve(vox_ctx* ctx, vox_solve_info* info);
/* PRESSURECALC (functions.cuh:101, functions.cu:2465-2549). Method 1: pnew =
 * 0, assemble, solve, clamp to [pmin, pmax] with NaN -> pmax. Method 3: the
 * equation of state on n* for non-dummy particles (2471-2500). */
int vox_pressurecalc(vox_ctx* ctx, vox_solve_info* info);
/* PHATCALC (functions.cuh:69, functions.cu:207-224). */
int vox_phat(vox_ctx* ctx);
/* PRESSGRAD (functions.cuh:71, functions.cu:245-355). */
int vox_pressgrad(vox_ctx* ctx);
/* EULERINTEGRATION + BOUNDARIE (functions.cuh:85, functions.cu:1963-2021). */
int vox_euler(vox_ctx* ctx);
/* COLLISION2 (functions.cuh:76, functions.cu:1523-1601), with the reference's
 * in-place sequential semantics reproduced exactly. */
int vox_collision(vox_ctx* ctx);
/* PREPDATA (functions.cuh:108, functions.cu:2688-2746). */
int vox_prepdata(vox_ctx* ctx);
/* DTcalculation (functions.cuh:78, functions.cu:1684-1698) when dt_mode == 1. */
int vox_dtcalc(vox_ctx* ctx);

/* Initialisation of main.cu:966-1012: density, NEIGHBOR, PNUM(loopstarted=0). */
int vox_init(vox_ctx* ctx);
/* nsteps iterations of the time loop body of main.cu:1027-1363 (no I/O). */
int vox_step(vox_ctx* ctx, int nsteps);
int vox_sync(vox_ctx* ctx);
int vox_get_status(vox_ctx* ctx, vox_status* st);
/* Batched scenarios (no counterpart in the reference; for ensembles of
 * independent cases laid out apart from each other in one context): region[i]
 * is particle i's scenario and boxes holds 6 bounds per scenario (xmin, xmax,
 * ymin, ymax, zmin, zmax). BOUNDARIE (functions.cu:1963-2002) and the PREPDATA
 * box check (functions.cu:2688-2746) then use the particle's own box instead of
 * the grid box. nregion = 0 clears it. */
int vox_set_regions(vox_ctx* ctx, int nregion, const int32_t* region, const double* boxes);

/* Tile-shape autotuner (no counterpart in the reference): times both
 * neighbour-search walks (1: one thread per particle, 2: one wave per 64
 * cell-sorted particles) over `reps` builds on the current positions and keeps the faster per
 * list (modes[2] receives the choice for lists S and L); then runs the step's
 * stages up to MATRIX once and times the predictor's variants (FP64: records
 * gathered per group 4, 8 or 2 with 256 threads per block, 4 or 8 with 128;
 * variants 5-8: 8 or 4 records with 128 or 256 threads and the gradient terms
 * deferred to a queue)
 * and the PPE product's (0: 8 entries per load batch with p gathered from a
 * per-block LDS window of the block's distinct columns; 1, 2, 3: 4, 16 and 8
 * entries on the 32-bit columns; 4: 8 entries on 16-bit column steps; 0 and 4
 * fall back to the next form where the assembly's rows do not allow theirs).
 * Every variant gives the same bits; the context keeps the fastest. Requires
 * vox_init. */
int vox_autotune(vox_ctx* ctx, int reps, int32_t* modes);
/* The context's tuning (no counterpart in the reference): choice[4] = walk of
 * list S, walk of list L (0 = heuristic), predictor variant (0..8), product
 * variant (0..4); vox_set_tuning applies a cached choice without timing. */
int vox_get_tuning(vox_ctx* ctx, int32_t* choice);
int vox_set_tuning(vox_ctx* ctx, const int32_t* choice);

/* Reference densities; vox_set_n0 lets a caller impose its own. */
int vox_get_n0(vox_ctx* ctx, double* n0_s, double* n0_l);
int vox_set_n0(vox_ctx* ctx, double n0_s, double n0_l);

/* ---- neighbour lists ----------------------------------------------------- */
/* Row lengths (tp ints) and row-major ids (0-based, row i at ids[i*stride]). */
int vox_get_neighbors(vox_ctx* ctx, int which, int32_t* counts, int32_t* ids, int stride);
/* The reference slot layout (functions.cu:2919-2920): a flat (tp+1)*stride
 * buffer, neighb[I*stride+1] = count+1, neighb[I*stride+2..] = 1-based ids. */
int vox_get_neighbors_slot(vox_ctx* ctx, int which, int32_t* neighb, int stride);
/* Cell key per particle (0-based cell number of functions.cu:2822-2834). */
int vox_get_cells(vox_ctx* ctx, int which, int32_t* cell);

/* ---- PPE export (dumps for training / checks) ---------------------------- */
/* First call with NULL arrays to get nu and nnz; then pass arrays of that size:
 * rowptr[nu+1], cols[nnz] (unknown indices), vals[nnz], rhs[nu], ids[nu]
 * (particle id of each unknown). */
int vox_ppe_export(vox_ctx* ctx, int32_t* nu, int32_t* nnz, int32_t* rowptr,
                   int32_t* cols, double* vals, double* rhs, int32_t* ids);
/* The inverse of vox_ppe_export (the "load" half of the PPE dump
 * for training data and checks; no counterpart in the reference): a CSR
 * system in the export's form (rowptr[nu+1], each row's diagonal first, cols
 * as unknown indices, ids the distinct particle of each unknown) replaces the
 * context's assembled PPE, so that vox_ppe_solve solves it. Requires vox_init. */
int vox_ppe_load(vox_ctx* ctx, int32_t nu, int32_t nnz, const int32_t* rowptr,
                 const int32_t* cols, const double* vals, const double* rhs, const int32_t* ids);

/* ---- PPE solve on the reference's device slot layout --------------------- */
/* Replaces CGM (functions.cuh:60, functions.cu:506-716) and the CUDA path's
 * INCDECOM + cgm1..cgm12 loop over d_poiss / d_neighb (main.cu:1159-1248,
 * functions.cuh:176-198) for a caller that keeps those arrays. All pointers
 * are device memory in the reference's 1-based layout: neighb and poiss hold
 * (tp+1)*stride elements (stride = NEIGHBORS), neighb[i*stride+1] = count+1,
 * neighb[i*stride+l] (l = 2..count+1) a 1-based neighbour id, poiss[i*stride+1]
 * the diagonal and poiss[i*stride+l] that neighbour's coefficient
 * (functions.cu:1007-1060); source, bcon and pnew hold tp+1 elements. Rows
 * with bcon == 0 are the unknowns and a neighbour with bcon != 0 is a
 * Dirichlet p = 0. Jacobi-PCG from pnew = 0 to ||r|| <= tol ||b|| (FP64, dot
 * products in a fixed order); pnew receives the solution and 0 on the other
 * rows. Runs on the null stream of the current device and returns when done.
 * The CG workspace (3 (tp+1) doubles) is allocated and freed inside the call,
 * as the reference's CGM does (functions.cu:508-512). info->unknowns and
 * info->nnz are -1 (not counted). */
int vox_cgm_slot(int tp, int stride, const int32_t* neighb, const double* poiss,
                 const double* source, const int32_t* bcon, double* pnew, double tol, int maxit,
                 vox_solve_info* info);
/* No counterpart in the reference: vox_cgm_slot for callers that cannot
 * allocate device memory (code under stream capture, serving loops, harnesses
 * that forbid hipMalloc). Same arguments, kernels and results (bit for bit);
 * the workspace is a static array of the library, one per GPU (24.1 MiB),
 * so tp is limited to 1,048,575 (VOX_ERR_ARG beyond). It calls no hipMalloc
 * or hipFree. Calls that overlap on one GPU share that array and race. */
int vox_cgm_slot_noalloc(int tp, int stride, const int32_t* neighb, const double* poiss,
                         const double* source, const int32_t* bcon, double* pnew, double tol,
                         int maxit, vox_solve_info* info);

/* ---- learned preconditioner (no counterpart in the reference) ------------ */
/* A message-passing network over the PPE matrix graph (hidden width 64,
 * messages of 16, 3 layers, bf16 weights on the matrix cores with FP32
 * accumulation) maps per-unknown features (PND ratio, wall flag, off-diagonal
 * dominance, first moments of the neighbour offsets) to a diagonal scaling g
 * and a coupling factor c of the preconditioner
 *   z_r = g_r w_r / A_rr + sqrt(c_r) sum_{k strong} O_rk / (A_rr A_kk) sqrt(c_k) w_k
 * (O = -offdiag(A); strong: O_rk at least a quarter of the coupling of two
 * particles one spacing apart, about the neighbours within 2.3 spacings),
 * used inside flexible CG. Training runs on the GPU over PPE systems dumped from
 * contexts, minimising the mean of ||A M w - w||^2 / ||w||^2 over probes w
 * (the right-hand side and Rademacher vectors) with Adam.
 * Parameters are VOX_LPC_NPARAM floats in the packed order W0[64][8], b0[64],
 * then per layer l = 1..3: P[16][64], U[64][64], V[64][16], c[64], then
 * Wo[2][64], bo[2]. */
#define VOX_LPC_NPARAM 19330
typedef struct vox_lpc vox_lpc;
/* New network on the current device; weights from SplitMix64(seed). */
int vox_lpc_create(vox_lpc** out, uint64_t seed);
int vox_lpc_destroy(vox_lpc* lpc);
int vox_lpc_get_weights(vox_lpc* lpc, float* w);
int vox_lpc_set_weights(vox_lpc* lpc, const float* w);
/* Append the context's assembled PPE (after vox_ppe_assemble) to the training
 * set; *index receives its number. */
int vox_lpc_dump(vox_lpc* lpc, vox_ctx* ctx, int32_t* index);
int vox_lpc_count(vox_lpc* lpc, int32_t* systems, int64_t* rows);
/* Forward pass on stored system `index`: features feat[8*nu], scaling g[nu]
 * and coupling c[nu] (any may be NULL). */
int vox_lpc_eval(vox_lpc* lpc, int32_t index, float* feat, float* g, float* c);
/* Loss and its gradient with respect to the parameters on stored system
 * `index`, probes drawn from `seed`. */
int vox_lpc_loss_grad(vox_lpc* lpc, int32_t index, uint64_t seed, double* loss, float* grad);
/* Adam over every stored system, `epochs` passes, one step per system; mean
 * loss of the first and of the last pass. */
int vox_lpc_train(vox_lpc* lpc, int32_t epochs, double lr, uint64_t seed, double* loss_first,
                  double* loss_last);
/* Use `lpc` (NULL detaches) when the context's solver is VOX_SOLVER_LEARNED_FCG.
 * The first learned solve runs Jacobi-CG to fix the baseline; a learned solve
 * that needs more than twice the baseline's iterations, or breaks down, is
 * redone by Jacobi-CG. */
int vox_attach_lpc(vox_ctx* ctx, vox_lpc* lpc);

/* ---- profiling ----------------------------------------------------------- */
/* When enabled, HIP events bracket every launch of the named kernel classes on
 * the context's stream; vox_prof_get returns total ms and launch count. Class
 * names: cell, sort, list (list S search), list_l (list L search), pnd, sps,
 * vfrac, visc, pred, bcon, assemble, cg_init, spmv, cg_update, cg_p,
 * ir_residual, phat, pgrad, euler, collision, prep, misc, lpc_net (features,
 * network and coupling setup of a learned solve), lpc_apply (the
 * preconditioner inside flexible CG). */
int vox_prof_enable(vox_ctx* ctx, int on);
/* Restrict the events to one class (NULL: every class), so that a timed run
 * pays for the events of the kernel it measures only. No counterpart in the
 * reference (it has no profiling). */
int vox_prof_only(vox_ctx* ctx, const char* kernel);
int vox_prof_get(vox_ctx* ctx, const char* kernel, double* total_ms, int64_t* launches);
int vox_prof_reset(vox_ctx* ctx);

/* ---- .grd input (main.cu:496-682) and VTU output (inOut.cpp:104-336) ------ */
/* Read the particle count of a .grd file. */
int vox_grd_count(const char* path, int* tp);
/* Read a .grd file into caller arrays of tp elements (z, w ignored in 2D). */
int vox_grd_read(const char* path, int dim, int tp, double* x, double* y, double* z,
                 double* u, double* v, double* w, double* p, int32_t* ptype,
                 int* gp, int* wp, int* fp);
/* Write a .grd file with full precision (%.17g). */
int vox_grd_write(const char* path, int dim, int tp, const double* x, const double* y,
                  const double* z, const double* u, const double* v, const double* w,
                  const double* p, const int32_t* ptype);
/* Write an ASCII UnstructuredGrid VTU of all particles (3D writes z and w,
 * which the reference omits: inOut.cpp:132-136). fluid_only writes PTYPE > 0.
 * Values are Float64 with 17 significant digits (restart-grade), in the line
 * layout the reference's getVTU reads (inOut.cpp:43-102). */
int vox_write_vtu(vox_ctx* ctx, const char* path, int fluid_only);
/* Point count of a VTU (getTP, inOut.cpp:25-41). */
int vox_vtu_count(const char* path, int* tp);
/* Restart state from a VTU (getVTU, inOut.cpp:43-102): positions,
 * velocities, pressure and type of tp points into caller arrays (z, w
 * ignored in 2D) and the ghost / wall / fluid counts. */
int vox_vtu_read(const char* path, int dim, int tp, double* x, double* y, double* z,
                 double* u, double* v, double* w, double* p, int32_t* ptype,
                 int* gp, int* wp, int* fp);

#ifdef __cplusplus
}
#endif
#endif /* VOXARMPS_H */
