/*
 * oracle.h — CPU restatement of the reference's per-time-step MPS pipeline.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it. The product (libvoxarmps.so) never links or calls it.
 *
 * Every routine restates the sequential ("sequential" codeOpt) semantics of
 * the reference's host twin in VoxarMPS/VoxarMPS/functions.cu, with 0-based
 * arrays of tp elements instead of the reference's 1-based arrays of tp+1.
 * The PPE is solved by Jacobi-preconditioned CG to a relative residual instead
 * of the reference's IC(0)-CG with its absolute stop rule; both converge to the
 * same A^-1 b, which is what is compared.
 */
#ifndef VOX_ORACLE_H
#define VOX_ORACLE_H

#include "../include/voxarmps.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

int  orc_create(orc_ctx** ctx, const vox_params* p, int tp, int gp, int wp, int fp);
void orc_destroy(orc_ctx* ctx);
int  orc_set_params(orc_ctx* ctx, const vox_params* p);
int  orc_get_params(orc_ctx* ctx, vox_params* p);
int  orc_upload(orc_ctx* ctx, int field, const void* host);
int  orc_download(orc_ctx* ctx, int field, void* host);

int  orc_density(orc_ctx* ctx);
int  orc_neighbor(orc_ctx* ctx);
int  orc_pnum(orc_ctx* ctx, int loopstarted);
int  orc_pnumstar(orc_ctx* ctx);
int  orc_sps(orc_ctx* ctx);
int  orc_vfraction(orc_ctx* ctx);
int  orc_viscosity(orc_ctx* ctx);
int  orc_prediction(orc_ctx* ctx);
int  orc_bcon(orc_ctx* ctx);
int  orc_ppe_assemble(orc_ctx* ctx);
int  orc_ppe_solve(orc_ctx* ctx, vox_solve_info* info);
/* The reference's own solver on the assembled system: INCDECOM (IC(0) in the
 * list-slot layout, functions.cu:1139-1200) and CGM (IC-preconditioned CG from
 * a zero start, stop when max |r_i| dt^2 <= eps after at least imin
 * iterations, at most imax; functions.cu:506-716). Writes pnew (no clamp). */
int  orc_ppe_solve_iccg(orc_ctx* ctx, int imin, int imax, double eps, vox_solve_info* info);
int  orc_pressurecalc(orc_ctx* ctx, vox_solve_info* info);
int  orc_phat(orc_ctx* ctx);
int  orc_pressgrad(orc_ctx* ctx);
int  orc_euler(orc_ctx* ctx);
int  orc_collision(orc_ctx* ctx);
int  orc_prepdata(orc_ctx* ctx);
int  orc_dtcalc(orc_ctx* ctx);
int  orc_threads(void);
int  orc_set_threads(int n);
int  orc_init(orc_ctx* ctx);
int  orc_step(orc_ctx* ctx, int nsteps);
int  orc_get_status(orc_ctx* ctx, vox_status* st);
int  orc_get_n0(orc_ctx* ctx, double* n0_s, double* n0_l);
int  orc_set_n0(orc_ctx* ctx, double n0_s, double n0_l);

int  orc_get_neighbors(orc_ctx* ctx, int which, int32_t* counts, int32_t* ids, int stride);
int  orc_max_count(orc_ctx* ctx, int which);
int  orc_get_cells(orc_ctx* ctx, int which, int32_t* cell);
int  orc_ppe_export(orc_ctx* ctx, int32_t* nu, int32_t* nnz, int32_t* rowptr,
                    int32_t* cols, double* vals, double* rhs, int32_t* ids);

/* Weight function W (functions.cu:32-70) for the closed-form checks. */
double orc_weight(double r, int ktype, int dim, double re);

/* O(N^2) neighbour search that applies the same cell/stencil/accept rules
 * without binning: for every i it scans all j, keeps those whose cell lies in
 * the clipped 3^d stencil and that pass the distance rule, then orders them by
 * stencil offset (x outer, y, z inner) and id. Independent of orc_neighbor. */
int  orc_neighbor_bruteforce(const vox_params* p, double re, int tp, const double* x,
                             const double* y, const double* z, int32_t* counts,
                             int32_t* ids, int stride);

#ifdef __cplusplus
}
#endif
#endif
