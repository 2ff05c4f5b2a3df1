// vox_internal.hpp — device context and helpers shared by the HIP sources of
// libvoxarmps. Everything here is gfx950 (wave64) code; no other target.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/voxarmps.h"

namespace vox {

constexpr int kBlock = 256;          // threads per workgroup (4 waves)
constexpr int kWave = 64;            // CDNA wavefront width
constexpr int kMaxPartials = 8192;   // reduction partial slots per scalar
constexpr int kPartialSlots = 10;    // scalars reduced through ctx->partials
constexpr int kPredVariants = 9;     // predictor: records per group x block size x deferred gradient
constexpr int kSpmvVariants = 5;     // PPE product: 8 entries per batch gathered from an LDS
                                     // window; 4, 16, 8 on 32-bit columns; 8 on 16-bit steps
constexpr int kWinMax = 7936;        // distinct columns of a 256-row PPE chunk in the window
constexpr int kWinSpanWords = 8192;  // bitmap words of the window build: column span <= 262,144
                                     // (bitmap and word prefixes: 48 KiB of LDS at most)


// Gather records: everything a neighbour loop reads about particle j in one
// 32- or 64-byte aligned record (one cache line) instead of one line per array.
struct __align__(32) PosRec {
    double x, y, z;
    int32_t id, aux;
};
struct __align__(64) PredRec {
    double x, y, z, u, v, w, pn, meu;
};

// ---- error plumbing ----------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define VOX_HIP(call)                                                              \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess)                                                      \
            return ::vox::fail(VOX_ERR_HIP, std::string(#call) + ": " +            \
                                                hipGetErrorString(e_));            \
    } while (0)

#define VOX_TRY(expr)                                                              \
    do {                                                                           \
        int rc_ = (expr);                                                          \
        if (rc_ != VOX_OK) return rc_;                                             \
    } while (0)

inline int grid_for(long n, int block = kBlock)
{
    long g = (n + block - 1) / block;
    return (int)(g < 1 ? 1 : g);
}

// ---- device flags (int32 slots of ctx->dflags) -------------------------------
enum DevFlag {
    FLAG_DOMAIN = 0,      // particle outside neighbour grid
    FLAG_OVERFLOW_S,      // list S row longer than capacity
    FLAG_OVERFLOW_L,
    FLAG_MAXCNT_S,        // longest row (atomicMax)
    FLAG_MAXCNT_L,
    FLAG_NAN,             // PREPDATA ERROR#1 count
    FLAG_OOB,             // PREPDATA ERROR#2 count
    FLAG_COLL_HITS,       // applied collision impulses
    FLAG_COLL_CAND,       // candidate pairs of the exact pass
    FLAG_COLL_VIOLATION,  // margin of the candidate screen exceeded
    FLAG_COLL_BIG,        // a candidate component exceeds the per-thread replay bound
    FLAG_COLL_OPEN,       // the replay moved a particle outside the active set
    FLAG_NU,              // number of PPE unknowns
    FLAG_CG_DONE,         // CG finished (converged, breakdown or cap)
    FLAG_CG_ITERS,
    FLAG_CG_CONV,
    FLAG_CG_BREAK,
    FLAG_DOMAIN_FIRST,    // smallest index of a particle outside the grid (atomicMin)
    FLAG_DCOL_BAD,        // a PPE row's column steps do not fit 16 bits (see dcol)
    FLAG_WIN_N,           // most distinct columns of a PPE chunk (atomicMax; see wcol)
    FLAG_WIN_SUM,         // distinct columns over all chunks (atomicAdd)
    FLAG_COL_BELOW,       // max over PPE rows of row - first column (atomicMax)
    FLAG_COL_ABOVE,       // max over PPE rows of last column - row (atomicMax)
    FLAG_COUNT
};

// device scalar slots of ctx->dscal (double); the solver names its own
constexpr int SC_COUNT = 12;

// ---- neighbour grid + list ------------------------------------------------------
struct Grid {
    double xmin, ymin, zmin;
    double hx, hz;          // cell width in x/y and in z (re and re+DELTA)
    int ncx, ncy, ncz, dim;
    int incl;               // accept d <= R with the particle itself (3D; the CUDA path)
    long tnc;
};

struct DevList {
    double re = 0.0;
    int cap = 0;             // entries per row (multiple of 8)
    Grid g{};
    int32_t* cnt = nullptr;        // tp
    int32_t* ids = nullptr;        // sliced rows, entry l of row i at sl_index(i, l, cap)
    int32_t* cell = nullptr;       // tp
    int32_t* cell_start = nullptr; // tnc + 1
    int32_t* cell_fill = nullptr;  // tnc
    int32_t* order_raw = nullptr;  // tp (atomic scatter result)
    PosRec* rec = nullptr;         // tp: positions + id sorted by (cell, id)
    bool sorted = false;           // rec holds a sort (a NEIGHBOR call has run)
    long tnc_alloc = 0;
    bool compact = false;          // rows hold only the entries of nonzero weight (see
                                   // list_compact_ok); the reference rows are rebuilt on demand
    uint64_t epoch = 0;            // vox_ctx::pos_epoch when the rows were searched
    int32_t* pos_of = nullptr;     // tp: particle id -> its position in rec
    bool pos_rows = false;         // row entries are positions in rec, not particle ids,
                                   // and each row (and count) sits at its particle's
                                   // position too (list L's compact FP64 rows)
};

// Sliced row storage in chunks of 8 (SELL-64x8). Rows are grouped in slices
// of 64 (one wavefront) and cut into chunks of 8 consecutive entries; chunk g
// of a slice's 64 rows is one contiguous block of 64 x 8 entries:
//   entry l of row i sits at (((i/64) * (cap/8) + l/8) * 64 + i%64) * 8 + l%8.
// A row's 8 consecutive entries are contiguous (one 32-byte sector of int32),
// so a lane that appends to its own row fills whole sectors instead of leaving
// one partly written sector per entry, and a wave reading chunk g of its 64
// rows reads one contiguous block (2 KB of int32). cap is a multiple of 8.
constexpr int kRowChunk = 8;
__host__ __device__ __forceinline__ long sl_index(long i, int l, int cap)
{
    return ((((i >> 6) * (long)(cap >> 3) + (l >> 3)) * 64 + (i & 63)) << 3) + (l & 7);
}
// Offset of entry l from entry 0 of the same row.
__host__ __device__ __forceinline__ long sl_step(int l) { return ((long)(l >> 3) << 9) + (l & 7); }
inline int round_cap(int cap) { return (cap + kRowChunk - 1) & ~(kRowChunk - 1); }

// Matrix rows (PPE) use plain SELL-64: entry l of row r at ((r/64)*cap + l)*64
// + r%64, so a wave reading entry l of its 64 rows (the product's access) reads
// one contiguous 256/512-byte block.
__host__ __device__ __forceinline__ long mx_index(long r, int l, int cap)
{
    return ((r >> 6) * (long)cap + l) * 64 + (r & 63);
}

typedef int32_t v4i __attribute__((ext_vector_type(4)));
typedef uint32_t v4u __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));
typedef float v4f __attribute__((ext_vector_type(4)));

// Chunk g (entries 8g..8g+7) of row r's 16-bit column steps, in 16-byte
// units: a wave's 64 rows read chunk g as one contiguous 1 KiB block.
__host__ __device__ inline long dc_index(long r, int g, int cap)
{
    return ((r >> 6) * (long)(cap >> 3) + g) * 64 + (r & 63);
}

// Non-temporal loads of 8 consecutive entries (p 32-byte aligned): streamed
// data that should not displace gathered vectors from the caches.
__device__ __forceinline__ void load8_nt(const int32_t* __restrict__ p, int32_t (&o)[8])
{
    const v4i a = __builtin_nontemporal_load(reinterpret_cast<const v4i*>(p));
    const v4i b = __builtin_nontemporal_load(reinterpret_cast<const v4i*>(p) + 1);
    o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
    o[4] = b.x; o[5] = b.y; o[6] = b.z; o[7] = b.w;
}
__device__ __forceinline__ void load8_nt(const double* __restrict__ p, double (&o)[8])
{
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const v2d a = __builtin_nontemporal_load(reinterpret_cast<const v2d*>(p) + k);
        o[2 * k] = a.x; o[2 * k + 1] = a.y;
    }
}
__device__ __forceinline__ void load8_nt(const float* __restrict__ p, float (&o)[8])
{
    const v4f a = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(p));
    const v4f b = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(p) + 1);
    o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
    o[4] = b.x; o[5] = b.y; o[6] = b.z; o[7] = b.w;
}

// The 8 entries of chunk g of a row (row = address of the row's entry 0,
// 32-byte aligned): two 16-byte loads.
__device__ __forceinline__ void load_chunk8(const int32_t* __restrict__ row, int g, int32_t (&o)[8])
{
    const v4i* p = reinterpret_cast<const v4i*>(row + ((long)g << 9));
    const v4i a = p[0], b = p[1];
    o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
    o[4] = b.x; o[5] = b.y; o[6] = b.z; o[7] = b.w;
}
inline long sl_size(long n, int cap) { return ((n + 63) / 64) * 64 * (long)cap; }

// The particle a per-particle kernel's thread t takes: position t of a list's
// cell-sorted order (a wave's particles then share most of their neighbours),
// or particle t before the first sort.
__device__ __forceinline__ int sorted_particle(const PosRec* order, int t)
{
    return order ? order[t].id : t;
}

// Lightweight by-value view of a list for kernels.
struct ListView {
    const int32_t* __restrict__ cnt;
    const int32_t* __restrict__ ids;
    int cap;
};

struct ProfRec { int kclass; hipEvent_t a, b; };

}  // namespace vox

// ---- the context -----------------------------------------------------------------
struct vox_ctx {
    vox_params p{};
    int tp = 0, gp = 0, wp = 0, fp = 0;
    int device = 0;
    hipStream_t stream = nullptr;

    double* f[VOX_F_COUNT] = {};
    int32_t* ptype = nullptr;
    long n_fluid2 = -1;            // PTYPE 2 particles in the uploaded types (-1: not known)
    uint64_t pos_epoch = 0;        // advanced by every write of x, y, z (EULER, COLLISION2,
                                   // PREPDATA, uploads)
    bool meu_uniform = false;      // MEU was last written by the Newtonian VISCOSITY with no
                                   // PTYPE 2 particle: nu1 rho1 everywhere
    int32_t* bcon = nullptr;

    vox::DevList list[2];
    int tuned_mode[2] = {0, 0};    // neighbour walk chosen by the autotuner (0: none)
    int tune_pred = 6;             // predictor variant (vox::kPredVariants): 8 records,
                                   // 256 threads, gradient deferred unless tuned
    int tune_spmv = 0;             // PPE product variant (vox::kSpmvVariants)
    bool two_lists = false;
    double n0s = 0.0, n0l = 0.0;
    bool initialised = false;

    // strain-rate scratch of SPS (s11 s12 s22 s13 s23 s33)
    double* sps_s[6] = {};
    double* scratch_d = nullptr;   // tp doubles
    vox::PredRec* pred_rec = nullptr;  // tp gather records of the predictor
    int32_t* pt_sorted = nullptr;  // tp: PTYPE in list L's cell order (predictor)
    vox::PosRec* pos_rec = nullptr;    // tp gather records (positions + two ints)

    // PPE (rows are unknowns r = uidx[i]; SELL-64 rows, mx_index)
    int32_t* uidx = nullptr;       // tp: particle -> unknown or -1
    int32_t* urow = nullptr;       // tp: unknown -> particle
    int32_t* uflag = nullptr;      // tp scan input
    int32_t* rowlen = nullptr;     // tp: stored off-diagonal entries of each row
    int32_t* uorder = nullptr;     // tp: particles in the numbering's cell order
    int32_t* uscan = nullptr;      // tp: unknown index at each position of uorder
    int32_t* zcell = nullptr;      // 2 (tnc + 1): cell counts and starts, z fastest
    long zcell_alloc = 0;
    bool uorder_cells = false;     // unknowns numbered in cell order (else particle order)
    bool uorder_particles = false; // force particle order (a dump for the learned network)
    int32_t* acol = nullptr;       // cap_l * tp, mx_index
    // The same columns as 16-bit steps: entry k of row r holds col_k - col_(k-1)
    // (entry 0: 0, col_0 in cbase[r]); a lane's 8 steps of entries 8g..8g+7
    // are one 16-byte chunk, chunk g of row r at ((r/64)(cap/8) + g) 64 + r%64.
    // Valid when every step of the assembly fit (dcol_ok): the rows' columns
    // ascend when the unknowns are numbered in the list's cell order.
    uint32_t* dcol = nullptr;      // cap_l * tp 16-bit steps
    int32_t* cbase = nullptr;      // tp: first column of each row
    bool dcol_ok = false;
    bool acol_ok = false;          // acol holds the current system's columns (the
                                   // assembly stores them only where the steps cannot
                                   // stand in; ensure_acol decodes them on demand)
    // Per chunk of 256 rows (one product block): the distinct columns its rows
    // read, ascending (at most kWinMax, at wuniq + chunk kWinMax), and each
    // entry's rank among them, 16-bit in dcol's layout (wcol). The product
    // stages p at those columns in LDS once per block and gathers from there.
    // Built on demand for the current assembly (win_asm == asm_count).
    uint32_t* wcol = nullptr;      // cap_l * tp 16-bit ranks
    int32_t* wuniq = nullptr;
    int32_t* wn = nullptr;         // distinct columns per chunk
    long wchunks_alloc = 0;
    int wcap_alloc = 0;            // a_cap wcol was allocated for
    int win_n = 0;                 // most distinct columns of a chunk
    int win_sum = 0;               // distinct columns over all chunks
    int col_below = 0, col_above = 0;   // every row's columns lie in [r - below, r + above]
    bool win_ok = false;
    long asm_count = 0;            // PPE assemblies and loads so far
    long win_asm = -1;             // the asm_count the windows were built for
    double* aval = nullptr;        // cap_l * tp, mx_index
    double* adiag = nullptr;       // tp
    double* rhs = nullptr;         // tp
    int a_cap = 0;
    int nu = 0;
    long nnz_total = 0;            // stored entries incl. diagonal (last solve)
    int cg_expect = 0;             // iterations of the previous FP64 solve
    int cg_expect_inner = 0;       // iterations of the previous FP32 inner solve
    int fcg_expect = 0;            // the same for flexible CG
    int fcg_expect_inner = 0;
    int ir_outer_last = 0;
    // CG vectors over unknowns
    double* cg_x = nullptr;
    double* cg_r = nullptr;
    double* cg_p = nullptr;
    double* cg_q = nullptr;
    double* cg_d = nullptr;        // FP64 inner solve of a warm-started solve: x
    double* cg_e = nullptr;        //   and residual
    // mixed precision: FP32 matrix values and inner-solve vectors
    float* aval32 = nullptr;       // same sliced layout as aval
    bool aval32_ok = false;        // aval32 holds the current assembly's values
    bool aval64_ok = false;        // aval holds them in FP64 (not after a mixed assembly)
    float* cg_x32 = nullptr;
    float* cg_r32 = nullptr;
    float* cg_p32 = nullptr;
    float* cg_q32 = nullptr;

    // learned preconditioner: attached network, per-unknown buffers of its
    // forward pass and the coupling operator of flexible CG
    vox_lpc* lpc = nullptr;
    int lpc_baseline = 0;          // Jacobi-CG iterations of the last plain solve
    int lpc_rows = 0;              // rows the buffers below hold (multiple of 64)
    float* lpc_feat = nullptr;     // 8 per row
    float* lpc_inv_osum = nullptr;
    float* lpc_invd = nullptr;     // 1 / A_rr
    float* lpc_t = nullptr;        // sqrt(c_r) / A_rr
    vox::v4f* lpc_upos = nullptr;  // FP32 position of each unknown
    uint16_t* lpc_h[2] = {};       // bf16 hidden states, 64 per row (ping-pong)
    uint16_t* lpc_s = nullptr;     // bf16 messages, 16 per row
    uint16_t* lpc_m = nullptr;     // bf16 aggregated messages, 16 per row
    float* lpc_o = nullptr;        // network output, 2 per row
    float* lpc_gd = nullptr;       // g_r / A_rr
    int32_t* lpc_slen = nullptr;   // coupling operator, SELL-64 with cap a_cap
    int32_t* lpc_scol = nullptr;
    float* lpc_sval = nullptr;
    int lpc_scap = 0;
    double* cg_z = nullptr;        // flexible CG: preconditioned residual
    float* cg_z32 = nullptr;

    // collision exact-emulation buffers
    int32_t* coll_cnt = nullptr;   // tp
    int32_t* coll_off = nullptr;   // tp + 1
    int32_t* coll_i = nullptr;     // cand capacity
    int32_t* coll_j = nullptr;
    int32_t* coll_stamp = nullptr; // tp
    int32_t* coll_root = nullptr;  // tp: smallest particle of each candidate component
    int32_t* coll_act = nullptr;   // tp: particle in the active set of the replay
    int32_t* coll_cnt2 = nullptr;  // tp: active candidates of each particle
    int32_t* coll_off2 = nullptr;  // tp + 1
    int32_t* coll_i2 = nullptr;    // active candidates, reference order
    int32_t* coll_j2 = nullptr;

    // batched scenarios: region of each particle and 6 box bounds per region
    int32_t* region = nullptr;
    double* rbox = nullptr;
    int nregion = 0;
    int32_t* coll_logidx = nullptr;// tp
    double* coll_log = nullptr;    // 8 doubles per logged particle
    int32_t* coll_logn = nullptr;  // 1
    long coll_cap = 0;
    int32_t coll_epoch = 0;
    double coll_margin = 0.1;      // fraction of coll added to the screen radius

    // scan scratch
    int32_t* scan_blocks = nullptr;
    long scan_blocks_cap = 0;
    int32_t* blockmax = nullptr;   // one slot per 256-particle block
    // side stream of the neighbour build (list L beside list S), with its own
    // scan scratch and block-maximum slots, and the events joining the streams
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    int32_t* scan_blocks2 = nullptr;
    long scan_blocks2_cap = 0;
    int32_t* blockmax2 = nullptr;

    // reductions
    double* partials = nullptr;    // 8 * kMaxPartials
    double* dscal = nullptr;       // SC_COUNT
    int32_t* dflags = nullptr;     // FLAG_COUNT
    int32_t* hflags = nullptr;     // pinned mirror
    double* hscal = nullptr;       // pinned mirror

    vox_status st{};

    // profiling
    bool prof_on = false;
    int prof_only = -1;            // profile this kernel class only (-1: every class)
    std::vector<vox::ProfRec> prof_pending;
    std::vector<hipEvent_t> prof_free;
    double prof_ms[32] = {};
    int64_t prof_n[32] = {};
};

namespace vox {

// kernel classes for event profiling
enum KClass {
    KC_CELL = 0, KC_SORT, KC_LIST, KC_PND, KC_SPS, KC_VFRAC, KC_VISC, KC_PRED,
    KC_BCON, KC_ASSEMBLE, KC_CG_INIT, KC_SPMV, KC_CG_UPDATE, KC_CG_P, KC_PHAT,
    KC_PGRAD, KC_EULER, KC_COLL, KC_PREP, KC_MISC, KC_IR, KC_LIST_L, KC_LPC_NET, KC_LPC_APPLY,
    KC_COUNT
};
const char* kclass_name(int k);
int kclass_from_name(const char* name);

void prof_begin(vox_ctx* c, int kclass, hipEvent_t* a);
void prof_end(vox_ctx* c, int kclass, hipEvent_t a);
int prof_drain(vox_ctx* c);

struct ProfScope {
    vox_ctx* c; int k; hipEvent_t a = nullptr;
    bool on;
    ProfScope(vox_ctx* cc, int kk)
        : c(cc), k(kk), on(cc->prof_on && (cc->prof_only < 0 || cc->prof_only == kk))
    {
        if (on) prof_begin(c, k, &a);
    }
    ~ProfScope() { if (on) prof_end(c, k, a); }
};

inline ListView view(const vox_ctx* c, int which)
{
    const DevList& L = (which == VOX_LIST_L && c->two_lists) ? c->list[1] : c->list[0];
    return ListView{L.cnt, L.ids, L.cap};
}
inline const DevList& list_of(const vox_ctx* c, int which)
{
    return (which == VOX_LIST_L && c->two_lists) ? c->list[1] : c->list[0];
}

// host-side helpers implemented in the .hip files
int scan_exclusive(vox_ctx* c, const int32_t* in, int32_t* out, long n, int32_t* total_dev);
int ensure_acol(vox_ctx* c);                                    // vox_ppe.hip
int fill_i32(vox_ctx* c, int32_t* p, int32_t v, long n);
int fill_f64(vox_ctx* c, double* p, double v, long n);
int read_flags(vox_ctx* c);   // D2H of dflags + dscal into pinned mirrors (syncs)

int autotune_lists(vox_ctx* c, int reps, int32_t* modes);
int build_neighbors(vox_ctx* c);
bool list_compact_ok(const vox_ctx* c);
int ensure_reference_list(vox_ctx* c, int which);
int list_l_for_consumer(vox_ctx* c);
// The cell-sorted records of list `which` as a per-particle kernels' thread
// order (sorted_particle), or null before the first NEIGHBOR.
const PosRec* list_order(const vox_ctx* c, int which);
int ensure_list_storage(vox_ctx* c, int which, int cap);

// stages (vox_stages.hip)
int st_density(vox_ctx* c);
int st_pnd(vox_ctx* c, int on_star, int loopstarted);
int st_sps(vox_ctx* c);
int st_vfraction(vox_ctx* c);
int st_viscosity(vox_ctx* c);
int st_prediction(vox_ctx* c);
const char* pred_variant_name(int v);
int autotune_kernels(vox_ctx* c, int reps);   // predictor and PPE product variants (vox_api.hip)
int time_spmv_variants(vox_ctx* c, int reps, float* ms);   // vox_cg.hip
bool uniform_meu(const vox_ctx* c);
int st_prediction32(vox_ctx* c);
int st_phat32(vox_ctx* c);          // the FP32 twins of the mixed precision mode
int st_sps32(vox_ctx* c);
int st_visc_hb32(vox_ctx* c);
int st_vfraction32(vox_ctx* c);
int st_bcon(vox_ctx* c);
int st_phat(vox_ctx* c);
int st_pressgrad(vox_ctx* c);
int st_pressgrad32(vox_ctx* c);
int st_euler(vox_ctx* c);
int st_collision(vox_ctx* c);
int st_prepdata(vox_ctx* c);
int st_clamp_pressure(vox_ctx* c);
int st_n0(vox_ctx* c);

// PPE (vox_ppe.hip)
int ppe_assemble(vox_ctx* c);
int st_pressure_wc(vox_ctx* c);   // PRESSURECALC method 3
int ppe_solve(vox_ctx* c, vox_solve_info* info);
int ppe_export(vox_ctx* c, int32_t* nu, int32_t* nnz, int32_t* rowptr, int32_t* cols,
               double* vals, double* rhs, int32_t* ids);
int ppe_load(vox_ctx* c, int32_t nu, int32_t nnz, const int32_t* rowptr, const int32_t* cols,
             const double* vals, const double* rhs, const int32_t* ids);

// learned preconditioner (vox_lpc.hip): features, network and coupling
// operator for the context's assembled PPE
int lpc_prepare(vox_ctx* c);
int ppe_make_aval32(vox_ctx* c);   // FP32 copy of the assembled values (vox_ppe.hip)
int ppe_make_aval64(vox_ctx* c);   // FP64 copy of FP32-assembled values
void lpc_free_ctx(vox_ctx* c);

}  // namespace vox

// ---- device helpers ------------------------------------------------------------------
namespace vox {

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// This remap hands every XCD one contiguous range of logical blocks, so the
// spatially adjacent particles of neighbouring blocks share an L2.
__device__ __forceinline__ int xcd_block()
{
    const int nb = gridDim.x, b = blockIdx.x;
    const int q = nb >> 3, r = nb & 7, x = b & 7;
    return x * q + min(x, r) + (b >> 3);
}
__device__ __forceinline__ int xcd_tid() { return xcd_block() * blockDim.x + threadIdx.x; }


// Weight function W of functions.cu:32-70 (ktype 2 and 6 on the hot path).
__device__ __forceinline__ double wfun(double r, int ktype, double re)
{
    double q = r / re;
    if (ktype == 2) return (q < 1.0) ? (1 / q) - 1 : 0.0;
    if (ktype == 6) return (q <= 1.0) ? (1 - q) * (1 - q) * (1 - q) : 0.0;
    if (ktype == 1) {
        if (q < 0.5) return 2.0 - 4.0 * (q * q);
        if (q <= 1.0) return (2 * q - 2) * (2 * q - 2);
        return 0.0;
    }
    if (ktype == 5) return (q <= 1.0) ? 1.5 * log(1 / (q + 0.000001)) : 0.0;
    return 0.0;
}

// a / b, correctly rounded, from y = 1.0 / b correctly rounded: q = a*y is
// within one ulp of a/b, the residual a - b*q is exact with an FMA, and one
// correction step q + (a - b*q)*y rounds to exactly the IEEE quotient
// (Markstein's theorem; binary64, no overflow or underflow on the way). It
// replaces a full division when several quotients share one divisor. b = 0
// gives y = inf and a non-finite result, as a / 0 does up to the sign of NaN;
// callers mask those entries.
__device__ __forceinline__ double div_by(double a, double b, double y)
{
    const double q = a * y;
    const double r = fma(-q, b, a);
    return fma(r, y, q);
}

// wfun with the division by re done through inv_re = 1.0 / re (identical
// result, see div_by).
__device__ __forceinline__ double wfun_r(double r, int ktype, double re, double inv_re)
{
    double q = div_by(r, re, inv_re);
    if (ktype == 2) return (q < 1.0) ? (1 / q) - 1 : 0.0;
    if (ktype == 6) return (q <= 1.0) ? (1 - q) * (1 - q) * (1 - q) : 0.0;
    if (ktype == 1) {
        if (q < 0.5) return 2.0 - 4.0 * (q * q);
        if (q <= 1.0) return (2 * q - 2) * (2 * q - 2);
        return 0.0;
    }
    if (ktype == 5) return (q <= 1.0) ? 1.5 * log(1 / (q + 0.000001)) : 0.0;
    return 0.0;
}

// wfun_r for a kernel type fixed at compile time (KT = 2 or 6; KT = 0 reads
// ktype at run time): the same expression, without the code of the other
// types (the logarithm of type 5 alone holds dozens of registers in a
// kernel that inlines it).
template <int KT>
__device__ __forceinline__ double wfun_k(double r, int ktype, double re, double inv_re)
{
    if constexpr (KT == 2) {
        double q = div_by(r, re, inv_re);
        return (q < 1.0) ? (1 / q) - 1 : 0.0;
    } else if constexpr (KT == 6) {
        double q = div_by(r, re, inv_re);
        return (q <= 1.0) ? (1 - q) * (1 - q) * (1 - q) : 0.0;
    } else {
        return wfun_r(r, ktype, re, inv_re);
    }
}

// The compile-time kernel type of a run (0: any other, read at run time).
inline int kt_of(int ktype) { return (ktype == 2 || ktype == 6) ? ktype : 0; }

// Fixed-order block sum of one double per thread (blockDim == kBlock).
__device__ __forceinline__ double block_sum(double v, double* lds)
{
    // wave-level tree through LDS keeps a fixed association order
    int t = threadIdx.x;
    lds[t] = v;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (t < s) lds[t] = lds[t] + lds[t + s];
        __syncthreads();
    }
    double r = lds[0];
    __syncthreads();
    return r;
}

}  // namespace vox
