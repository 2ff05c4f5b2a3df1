// vox_neighbor.hip — hash-sort neighbour grid for gfx950.
//
// Replaces NEIGHBOR (functions.cu:2788-2929) and the GPU hash search of
// neighbour_parallel.cu:344-444. The output reproduces the reference's CPU
// list bit-exactly:
//   cell  = int((x-Xmin)/re)+1, int((y-Ymin)/re)+1, int((z-Zmin)/(re+DELTA))+1
//           (IEEE divide, truncation; functions.cu:2822-2828),
//   order = stencil x-offset outer, y middle, z inner, clipped at the grid edge
//           (2865-2880), ascending particle id inside a cell (the serial
//           binning of 2820-2852 is a stable sort by cell),
//   rule  = 3D: sqrt(dx^2+dy^2+dz^2) <= re+DELTA, self included (2886);
//           2D: i != j and sqrt(dx^2+dy^2) < re (2907-2910).
// With ref_gpu the 2D list follows the reference's CUDA search instead
// (neighbour_parallel.cu:37-72, 145-199): cells re+DELTA wide in x and y,
// sqrt(dx^2+dy^2) <= re+DELTA with the particle itself, the same stencil order,
// ascending id inside a cell (its sort by cell key keeps the id order).
// The distance is evaluated exactly as the reference (no FMA: the library is
// built with -ffp-contract=off); a +-1e-10 relative band around the radius
// decides which candidates need the correctly rounded sqrt.
//
// Pipeline per list: key+count (atomic histogram) -> exclusive scan (cell
// start) -> atomic scatter -> per-cell rank sort by id (one wave per cell,
// LDS tile) with position records copied into cell order -> per-particle stencil
// walk (one thread per particle, or a wave per 64 sorted particles) writing sliced rows
// (sl_index). With two radii the two lists' pipelines run on two streams.
#include <utility>

#include "vox_internal.hpp"

namespace vox {

// ---- utility kernels ----------------------------------------------------------------
__global__ void k_fill_i32(int32_t* __restrict__ p, int32_t v, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long s = (long)gridDim.x * blockDim.x;
    for (; i < n; i += s) p[i] = v;
}

__global__ void k_fill_f64(double* __restrict__ p, double v, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long s = (long)gridDim.x * blockDim.x;
    for (; i < n; i += s) p[i] = v;
}

int fill_i32(vox_ctx* c, int32_t* p, int32_t v, long n)
{
    if (n <= 0) return VOX_OK;
    int g = grid_for(n);
    if (g > 4096) g = 4096;
    k_fill_i32<<<g, kBlock, 0, c->stream>>>(p, v, n);
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

int fill_f64(vox_ctx* c, double* p, double v, long n)
{
    if (n <= 0) return VOX_OK;
    int g = grid_for(n);
    if (g > 4096) g = 4096;
    k_fill_f64<<<g, kBlock, 0, c->stream>>>(p, v, n);
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

// ---- exclusive scan (int32), three phases ----------------------------------------------
constexpr int kScanItems = 4;                       // items per thread
constexpr int kScanTile = kBlock * kScanItems;      // 1024 per block

__device__ __forceinline__ int32_t block_excl_scan(int32_t v, int32_t* lds, int32_t* total)
{
    int t = threadIdx.x;
    lds[t] = v;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        int32_t add = (t >= off) ? lds[t - off] : 0;
        __syncthreads();
        lds[t] += add;
        __syncthreads();
    }
    int32_t incl = lds[t];
    *total = lds[kBlock - 1];
    __syncthreads();
    return incl - v;
}

// Exclusive scan in two launches: the sum of each tile of kScanTile items,
// then per tile its offset (the sum of the tiles before it, which each block
// adds up from the first pass's sums) plus the tile's own exclusive scan.
__global__ void k_scan_tile_sums(const int32_t* __restrict__ in, int32_t* __restrict__ block_sums,
                                 long n)
{
    __shared__ int32_t lds[kBlock];
    const long base = (long)blockIdx.x * kScanTile + (long)threadIdx.x * kScanItems;
    int32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) s += (base + k < n) ? in[base + k] : 0;
    int32_t tot;
    (void)block_excl_scan(s, lds, &tot);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

// total (if given) receives the sum of all n items; it may alias out + n.
__global__ void k_scan_out(const int32_t* in, int32_t* out, const int32_t* __restrict__ block_sums,
                           long n, int32_t* total)
{
    __shared__ int32_t lds[kBlock];
    int32_t before = 0;
    for (int k = threadIdx.x; k < (int)blockIdx.x; k += kBlock) before += block_sums[k];
    int32_t off;
    (void)block_excl_scan(before, lds, &off);
    const long base = (long)blockIdx.x * kScanTile + (long)threadIdx.x * kScanItems;
    int32_t v[kScanItems];
    int32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        s += v[k];
    }
    int32_t tot;
    int32_t ex = off + block_excl_scan(s, lds, &tot);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = ex;
        ex += v[k];
    }
    if (total && threadIdx.x == 0 && blockIdx.x == gridDim.x - 1) *total = off + tot;
}

// out[0..n) = exclusive scan of in[0..n); out[n] = total if out has room (caller
// passes total_dev to receive it; may alias out + n).
int scan_exclusive(vox_ctx* c, const int32_t* in, int32_t* out, long n, int32_t* total_dev)
{
    long nb = (n + kScanTile - 1) / kScanTile;
    if (nb < 1) nb = 1;
    if (nb > c->scan_blocks_cap) {
        if (c->scan_blocks) (void)hipFree(c->scan_blocks);
        c->scan_blocks = nullptr;
        VOX_HIP(hipMalloc(&c->scan_blocks, sizeof(int32_t) * (nb + 64)));
        c->scan_blocks_cap = nb + 64;
    }
    k_scan_tile_sums<<<(int)nb, kBlock, 0, c->stream>>>(in, c->scan_blocks, n);
    k_scan_out<<<(int)nb, kBlock, 0, c->stream>>>(in, out, c->scan_blocks, n, total_dev);
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

// ---- grid ------------------------------------------------------------------------------
static Grid make_grid(const vox_params& p, double re)
{
    Grid g;
    g.xmin = p.xmin; g.ymin = p.ymin; g.zmin = p.zmin;
    g.hx = re;
    g.hz = re + p.delta;
    g.dim = p.dim;
    g.incl = (p.dim == 3 || p.ref_gpu) ? 1 : 0;
    // the CUDA path's 2D cells are re+DELTA wide (neighbour_parallel.cu:46-47, 57-58)
    if (p.ref_gpu) g.hx = re + p.delta;
    g.ncx = (int)((p.xmax - p.xmin) / g.hx) + 1;                  // functions.cu:2790-2791
    g.ncy = (int)((p.ymax - p.ymin) / g.hx) + 1;
    g.ncz = (p.dim == 3) ? (int)((p.zmax - p.zmin) / g.hz) + 1 : 1; // 2796
    g.tnc = (long)g.ncx * g.ncy * g.ncz;
    return g;
}

// 1-based cell coordinates; false outside the grid (the reference has no check).
__device__ __forceinline__ bool cell_coords(const Grid& g, double x, double y, double z,
                                            int& ic, int& jc, int& kc)
{
    if (!isfinite(x) || !isfinite(y) || (g.dim == 3 && !isfinite(z))) return false;
    double ax = (x - g.xmin) / g.hx;
    double ay = (y - g.ymin) / g.hx;
    if (ax > 2e9 || ay > 2e9 || ax < -2e9 || ay < -2e9) return false;
    ic = (int)ax + 1;
    jc = (int)ay + 1;
    kc = 1;
    if (g.dim == 3) {
        double az = (z - g.zmin) / g.hz;
        if (az > 2e9 || az < -2e9) return false;
        kc = (int)az + 1;
    }
    return ic >= 1 && ic <= g.ncx && jc >= 1 && jc <= g.ncy && kc >= 1 && kc <= g.ncz;
}

// ctr[key] += 1 for every active lane of the wave, with one atomic per
// distinct key: consecutive particles mostly share a cell, so per-lane atomics
// on one counter serialise. Returns the lane's slot among the key's
// increments (the counter's old value plus its rank among the wave's lanes
// with that key). Must be reached by every lane of the wave; the loop runs
// once per distinct key among the active lanes.
__device__ __forceinline__ int32_t wave_add1(int32_t* __restrict__ ctr, int32_t key, bool active)
{
    const int lane = threadIdx.x & (kWave - 1);
    unsigned long long todo = __ballot(active);
    int32_t slot = 0;
    while (todo) {
        const int lead = __ffsll((unsigned long long)todo) - 1;
        const int32_t lk = __shfl(key, lead);
        const bool mine = active && key == lk;
        const unsigned long long m = __ballot(mine);
        int32_t base = 0;
        if (lane == lead) base = atomicAdd(&ctr[lk], (int32_t)__popcll(m));
        base = __shfl(base, lead);
        if (mine) slot = base + (int32_t)__popcll(m & ((1ull << lane) - 1ull));
        todo &= ~m;
    }
    return slot;
}

__global__ __launch_bounds__(kBlock) void k_cell_keys(Grid g, int n, const double* __restrict__ x,
                            const double* __restrict__ y, const double* __restrict__ z,
                            int32_t* __restrict__ cell, int32_t* __restrict__ count,
                            int32_t* __restrict__ flags)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = i < n;
    int32_t key = 0;
    if (act) {
        int ic, jc, kc;
        double zi = (g.dim == 3) ? z[i] : 0.0;
        if (cell_coords(g, x[i], y[i], zi, ic, jc, kc)) {
            key = (ic - 1) + (jc - 1) * g.ncx + (kc - 1) * g.ncx * g.ncy;   // 2826-2834, 0-based
        } else {
            atomicAdd(&flags[FLAG_DOMAIN], 1);
            atomicMin(&flags[FLAG_DOMAIN_FIRST], i);
        }
        cell[i] = key;
    }
    wave_add1(count, key, act);
}

// Places each particle in its cell's run of order_raw; the order inside a run
// is arbitrary here (the per-cell id sort or rank that follows makes it
// ascending by id).
__global__ __launch_bounds__(kBlock) void k_scatter(int n, const int32_t* __restrict__ cell,
                          const int32_t* __restrict__ start, int32_t* __restrict__ fill,
                          int32_t* __restrict__ order_raw)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = i < n;
    const int32_t k = act ? cell[i] : 0;
    const int32_t slot = wave_add1(fill, k, act);
    if (act) order_raw[start[k] + slot] = i;
}

// One wave per cell: rank of each id among the cell's ids (ids are unique), so
// the run becomes ascending by id; positions are copied into the same order,
// with the cell key in the record's aux slot.
constexpr int kSortTile = 512;   // ids per wave held in LDS

__global__ __launch_bounds__(kBlock) void k_cell_sort(
    long tnc, int dim, const int32_t* __restrict__ start, const int32_t* __restrict__ order_raw,
    PosRec* __restrict__ rec, int32_t* __restrict__ pos_of, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z)
{
    __shared__ int32_t tile[kBlock / kWave][kSortTile];
    int wv = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    long cellid = (long)blockIdx.x * (kBlock / kWave) + wv;
    if (cellid >= tnc) return;
    int s = start[cellid], e = start[cellid + 1], n = e - s;
    if (n <= 0) return;
    const int32_t* src = order_raw + s;
    bool in_lds = (n <= kSortTile);
    if (in_lds)
        for (int k = lane; k < n; k += kWave) tile[wv][k] = src[k];
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    for (int k = lane; k < n; k += kWave) {
        int32_t id = in_lds ? tile[wv][k] : src[k];
        int rank = 0;
        for (int m = 0; m < n; ++m) {
            int32_t o = in_lds ? tile[wv][m] : src[m];
            rank += (o < id);
        }
        int pos = s + rank;
        PosRec rr;
        rr.x = x[id];
        rr.y = y[id];
        rr.z = (dim == 3) ? z[id] : 0.0;
        rr.id = id;
        rr.aux = (int32_t)cellid;
        rec[pos] = rr;
        pos_of[id] = pos;
    }
}

// The same order with one thread per particle, for grids whose cells hold a
// few particles each (a wave per cell would leave most lanes idle and visit
// every empty cell): the particle at scattered slot t counts the ids of its
// cell's members below its own, which is its rank in the ascending run.
__global__ __launch_bounds__(kBlock) void k_cell_rank(
    int n, int dim, const int32_t* __restrict__ cell, const int32_t* __restrict__ start,
    const int32_t* __restrict__ order_raw, PosRec* __restrict__ rec, int32_t* __restrict__ pos_of,
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int32_t id = order_raw[t];
    const int32_t c = cell[id];
    const int s = start[c], e = start[c + 1];
    int rank = 0;
    for (int m = s; m < e; ++m) rank += (order_raw[m] < id);
    PosRec rr;
    rr.x = x[id];
    rr.y = y[id];
    rr.z = (dim == 3) ? z[id] : 0.0;
    rr.id = id;
    rr.aux = c;
    rec[s + rank] = rr;
    pos_of[id] = s + rank;
}

// Per-particle stencil walk. R is the accept radius (3D and the CUDA path:
// re+DELTA, inclusive; 2D: re, exclusive).
__device__ __forceinline__ int wave_max_i32(int v)
{
    for (int off = kWave / 2; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, kWave));
    return v;
}

// Row output through a per-thread ring of 16 ids in LDS: a completed group of
// 8 entries is stored at once (one 32-byte store into the lane's own chunk),
// instead of 8 separate 4-byte stores that would leave the sector partly
// written several times. Groups past the row capacity are not stored (the
// overflow is reported and the search repeats with wider rows).
constexpr int kIdRing = 17;   // stride: lanes' equal slots fall on distinct banks

__device__ __forceinline__ void flush_ids(const int32_t* ring, int c, int cap,
                                          int32_t* __restrict__ row)
{
    if ((c << 3) >= cap) return;
    const int s0 = (c & 1) * 8;
    v4i* o = reinterpret_cast<v4i*>(row + ((long)c << 9));
    o[0] = v4i{ring[s0 + 0], ring[s0 + 1], ring[s0 + 2], ring[s0 + 3]};
    o[1] = v4i{ring[s0 + 4], ring[s0 + 5], ring[s0 + 6], ring[s0 + 7]};
}

// Thread t takes the particle at position t of the cell-sorted order (its
// record holds position, id and cell), so a wave's lanes sit in a few
// neighbouring cells of one grid row: they walk the same or overlapping
// stencils, with equal trip counts over a shared cell and candidate loads
// that hit the same lines.
__global__ __launch_bounds__(kBlock) void k_build_list(
    Grid g, int n, double R, const int32_t* __restrict__ start, const PosRec* __restrict__ rec,
    int cap, int32_t* __restrict__ ids, int32_t* __restrict__ cnt, int32_t* __restrict__ flags,
    int flag_overflow, int flag_max, int32_t* __restrict__ blockmax)
{
    __shared__ int32_t ring_all[kBlock * kIdRing];
    int32_t* ring = ring_all + threadIdx.x * kIdRing;
    const int t = xcd_tid();
    int k = 0;
    if (t < n) {
        const PosRec me = rec[t];
        const int i = me.id;
        const int32_t key = me.aux;
        const int ncxy = g.ncx * g.ncy;
        const int ic = key % g.ncx + 1;
        const int jc = (key / g.ncx) % g.ncy + 1;
        const int kc = key / ncxy + 1;
        const int m1 = (ic == 1) ? 0 : -1, m2 = (ic == g.ncx) ? 0 : 1;
        const int m3 = (jc == 1) ? 0 : -1, m4 = (jc == g.ncy) ? 0 : 1;
        int m5 = 0, m6 = 0;
        if (g.dim == 3) { m5 = (kc == 1) ? 0 : -1; m6 = (kc == g.ncz) ? 0 : 1; }
        const double xi = me.x, yi = me.y, zi = me.z;
        const double R2 = R * R;
        const double lo = R2 * (1.0 - 1e-10), hi = R2 * (1.0 + 1e-10);
        const long rb = sl_index(i, 0, cap);
        for (int row = m1; row <= m2; ++row)
            for (int col = m3; col <= m4; ++col)
                for (int dep = m5; dep <= m6; ++dep) {
                    int cn = (ic - 1 + row) + (jc - 1 + col) * g.ncx + (kc - 1 + dep) * ncxy;
                    const int s = start[cn], e = start[cn + 1];
                    // chunks of 8 candidates: loads first, then the ordered
                    // accept/store pass (stores would serialise later loads)
                    constexpr int U = 8;
                    for (int j0 = s; j0 < e; j0 += U) {
                        PosRec cr[U];
#pragma unroll
                        for (int u = 0; u < U; ++u) cr[u] = rec[(j0 + u < e) ? j0 + u : s];
                        const int k0 = k;
#pragma unroll
                        for (int u = 0; u < U; ++u) {
                            // branch-free test (the loads above stay unconditional)
                            // z is 0 in 2D records: dx^2+dy^2+0 equals the 2D sum
                            const PosRec cu = cr[u];
                            const double dx = cu.x - xi, dy = cu.y - yi, dz = cu.z - zi;
                            const double d2 = dx * dx + dy * dy + dz * dz;
                            bool acc = d2 < lo;
                            if (!acc && !(d2 > hi))   // rare band: exact sqrt decides
                                acc = g.incl ? (sqrt(d2) <= R) : (sqrt(d2) < R);
                            acc = acc && (j0 + u < e) && (g.incl || cu.id != i);
                            if (acc) {
                                ring[k & 15] = cu.id;
                                ++k;
                            }
                        }
                        // at most 8 accepts per chunk: at most one group completed
                        if ((k0 >> 3) != (k >> 3)) flush_ids(ring, k0 >> 3, cap, ids + rb);
                    }
                }
        if (k & 7) flush_ids(ring, k >> 3, cap, ids + rb);
        cnt[i] = k;
        if (k > cap) atomicAdd(&flags[flag_overflow], 1);
    }
    // longest row: wave max -> block max -> one slot per block (no atomics)
    __shared__ int wmax[kBlock / kWave];
    int m = wave_max_i32(k);
    if ((threadIdx.x % kWave) == 0) wmax[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        int b = wmax[0];
        for (int w = 1; w < kBlock / kWave; ++w) b = max(b, wmax[w]);
        blockmax[blockIdx.x] = b;
    }
    (void)flag_max;
}

// Stencil walk for large cells (3D list L: 64-80 particles per cell). A wave
// takes 64 consecutive particles of the cell-sorted order - one to a few
// cells of one x-row of the grid - and streams the union of their stencils
// once: x-blocks from the first lane's cell - 1 to the last lane's + 1, in
// each the (y, z) cells of the row's stencil in reference order, each cell's
// records in chunks of 64 through LDS, every lane testing every candidate
// (broadcast reads). A lane takes the candidates of its own three x-blocks
// only; its candidate sequence is therefore exactly its stencil in reference
// order (x outer, then y, then z, ascending id inside a cell). Lanes of
// another grid row, or more than one cell further along x than the group's
// first lane, are served by the next pass of the same wave, so a pass never
// streams more than four x-blocks. Against one wave per cell, lanes stay
// filled where a cell holds more than 64 particles or fewer.
// The accept test screens with a fused distance (within a few ulps of the
// reference's unfused sum); only inside the +-1e-10 band around R, where the
// two could disagree, does the reference's own expression and sqrt decide.
// The order and the row storage are those of k_build_list.
// kPos: the rows hold each neighbour's position in rec instead of its id.
template <bool kPos>
__global__ __launch_bounds__(kWave) void k_build_list_xrow(
    Grid g, int n, double R, const int32_t* __restrict__ start, const PosRec* __restrict__ rec,
    int cap, int32_t* __restrict__ ids, int32_t* __restrict__ cnt, int32_t* __restrict__ flags,
    int flag_overflow, int flag_max)
{
    __shared__ PosRec tile[kWave];
    __shared__ int32_t ring_all[kWave * kIdRing];
    int32_t* ring = ring_all + threadIdx.x * kIdRing;
    const int lane = threadIdx.x;
    const long t = (long)xcd_block() * kWave + lane;
    const bool valid = t < n;
    const PosRec me = rec[valid ? t : 0];
    const int i = me.id;
    const int key = me.aux;                           // cell, 0-based
    const int ix = key % g.ncx;
    const int rowkey = key / g.ncx;                   // y + ncy z
    const long ncxy = (long)g.ncx * g.ncy;
    // rows of positions are stored at the particle's own position too
    const long rb = sl_index(kPos ? t : (long)i, 0, cap);
    const double R2 = R * R;
    const double lo = R2 * (1.0 - 1e-10), hi = R2 * (1.0 + 1e-10);
    int k = 0;
    const int tt = (int)t;
    unsigned long long todo = __ballot(valid);
    while (todo) {
        // the group's cell row and x-range are the same for every lane: kept
        // in scalar registers, so the walk's cell bounds and the candidate
        // index are scalar too
        const int lead = __ffsll(todo) - 1;
        const int grow = __builtin_amdgcn_readfirstlane(__shfl(rowkey, lead));
        const int gx = __builtin_amdgcn_readfirstlane(__shfl(ix, lead));
        const bool mine = valid && rowkey == grow && ix <= gx + 1 && ((todo >> lane) & 1ull);
        const unsigned long long m = __ballot(mine);
        todo &= ~m;
        const int last = 63 - __clzll(m);
        const int x_lo = max(gx - 1, 0);
        const int x_hi = min(__builtin_amdgcn_readfirstlane(__shfl(ix, last)) + 1, g.ncx - 1);
        const int gy = grow % g.ncy, gz = grow / g.ncy;
        const int y_lo = max(gy - 1, 0), y_hi = min(gy + 1, g.ncy - 1);
        const int z_lo = (g.dim == 3) ? max(gz - 1, 0) : 0;
        const int z_hi = (g.dim == 3) ? min(gz + 1, g.ncz - 1) : 0;
        for (int xb = x_lo; xb <= x_hi; ++xb) {
            const bool on = mine && xb >= ix - 1 && xb <= ix + 1;
            for (int yb = y_lo; yb <= y_hi; ++yb)
                for (int zb = z_lo; zb <= z_hi; ++zb) {
                    const long cn = xb + yb * (long)g.ncx + zb * ncxy;
                    const int cs = __builtin_amdgcn_readfirstlane(start[cn]);
                    const int ce = __builtin_amdgcn_readfirstlane(start[cn + 1]);
                    for (int j0 = cs; j0 < ce; j0 += kWave) {
                        const int nc = min(kWave, ce - j0);
                        __syncthreads();                   // last chunk's reads are done
                        if (lane < nc) tile[lane] = rec[j0 + lane];
                        __syncthreads();
                        // groups of 8 candidates: at most one group of 8
                        // accepts completes per group. Slots past nc hold
                        // stale records and are never accepted.
                        for (int u0 = 0; u0 < nc; u0 += 8) {
                            const int k0 = k;
#pragma unroll
                            for (int uu = 0; uu < 8; ++uu) {
                                const int u = u0 + uu;
                                // z is 0 in 2D records
                                const double dx = tile[u].x - me.x, dy = tile[u].y - me.y,
                                             dz = tile[u].z - me.z;
                                const double d2f = fma(dz, dz, fma(dy, dy, dx * dx));
                                const bool sure = d2f < lo;
                                const bool near = d2f <= hi;
                                bool acc = sure;
                                if (near != sure) {          // rare band: the reference's expression
                                    const double d2 = dx * dx + dy * dy + dz * dz;
                                    acc = g.incl ? (sqrt(d2) <= R) : (sqrt(d2) < R);
                                }
                                // the particle itself: by its position in rec (rows
                                // of positions), else by id
                                const bool other = kPos ? (j0 + u != tt) : (tile[u].id != i);
                                acc = acc && on && (u < nc) && (g.incl || other);
                                // the next free slot takes every candidate; only
                                // an accepted one advances past it
                                ring[k & 15] = kPos ? j0 + u : tile[u].id;
                                k += acc ? 1 : 0;
                            }
                            if ((k0 >> 3) != (k >> 3)) flush_ids(ring, k0 >> 3, cap, ids + rb);
                        }
                    }
                }
        }
    }
    if (valid && (k & 7)) flush_ids(ring, k >> 3, cap, ids + rb);
    int kmax = 0;
    if (valid) {
        cnt[kPos ? t : (long)i] = k;
        if (k > cap) atomicAdd(&flags[flag_overflow], 1);
        kmax = k;
    }
    kmax = wave_max_i32(kmax);
    if (lane == 0) atomicMax(&flags[flag_max], kmax);
}

__global__ void k_reduce_max(const int32_t* __restrict__ v, int n, int32_t* __restrict__ out)
{
    __shared__ int lds[kBlock];
    int m = 0;
    for (int k = threadIdx.x; k < n; k += kBlock) m = max(m, (int)v[k]);
    lds[threadIdx.x] = m;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) lds[threadIdx.x] = max(lds[threadIdx.x], lds[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = lds[0];
}

// ---- host side --------------------------------------------------------------------------
int ensure_list_storage(vox_ctx* c, int which, int cap)
{
    DevList& L = c->list[which];
    cap = round_cap(cap);
    if (cap <= L.cap && L.ids) return VOX_OK;
    if (L.ids) (void)hipFree(L.ids);
    L.ids = nullptr;
    VOX_HIP(hipMalloc(&L.ids, sizeof(int32_t) * (size_t)sl_size(c->tp, cap)));
    L.cap = cap;
    return VOX_OK;
}

static int search_one(vox_ctx* c, int which, double re);

// List L in 3D can hold compact rows: the reference accepts d <= re + DELTA
// with the particle itself (functions.cu:2886), but its consumers, PREDICTION
// and MATRIX, weigh each entry by W(d) with W = 0 at d >= re (kernel types 2
// and 6) and skip the particle itself, so the entries they actually use are
// those at d < re other than the particle itself - the 2D accept rule at the
// same radius. The compact rows keep the reference order (a subsequence of it,
// about 18% shorter), and a sum over them adds exactly the nonzero terms of
// the reference's sum in the same order (the skipped terms are +0 added to a
// sum that is never -0), so every result is the same bits for finite fields
// (a non-finite neighbour value times W = 0 is NaN in the reference's sum).
// MATRIX on the CPU path assembles at the positions the list was searched at
// (x, not x*), so its distances are the search's own. matopt and srcopt weigh
// by re/d^3 at any d and need the reference rows, as does an export:
// ensure_reference_list rebuilds them from the cell-sorted records of the same
// NEIGHBOR call.
const PosRec* list_order(const vox_ctx* c, int which)
{
    const DevList& L = list_of(c, which);
    return L.sorted ? L.rec : nullptr;
}

bool list_compact_ok(const vox_ctx* c)
{
    const vox_params& p = c->p;
    return c->two_lists && p.dim == 3 && !p.matopt && !p.srcopt && !p.ref_gpu &&
           (p.ktype == 2 || p.ktype == 6);
}

// Expected particles per cell (cell volume / DL^dim) below 24: the cells are
// sorted one thread per particle rather than one wave per cell.
static bool few_per_cell(const vox_ctx* c, const DevList& L)
{
    const double per_cell = (c->p.dim == 3)
        ? L.g.hx * L.g.hx * L.g.hz / (c->p.dl * c->p.dl * c->p.dl)
        : L.g.hx * L.g.hx / (c->p.dl * c->p.dl);
    return per_cell < 24.0;
}

// Grid, cell sort and one search pass of list `which` on c->stream, without
// waiting for it. Reports, once the stream has run, go to the list's overflow
// and longest-row flags and to the shared domain flags.
static int launch_one(vox_ctx* c, int which, double re)
{
    DevList& L = c->list[which];
    const int n = c->tp;
    L.re = re;
    L.g = make_grid(c->p, re);
    if (L.g.tnc <= 0 || L.g.tnc > (1L << 30))
        return fail(VOX_ERR_ARG, "neighbour grid has an invalid number of cells");
    if (L.g.tnc > L.tnc_alloc) {
        if (L.cell_start) (void)hipFree(L.cell_start);
        if (L.cell_fill) (void)hipFree(L.cell_fill);
        L.cell_start = L.cell_fill = nullptr;
        VOX_HIP(hipMalloc(&L.cell_start, sizeof(int32_t) * (L.g.tnc + 1)));
        VOX_HIP(hipMalloc(&L.cell_fill, sizeof(int32_t) * (L.g.tnc + 1)));
        L.tnc_alloc = L.g.tnc;
    }
    {
        ProfScope ps(c, KC_CELL);
        VOX_TRY(fill_i32(c, L.cell_fill, 0, L.g.tnc));
        k_cell_keys<<<grid_for(n), kBlock, 0, c->stream>>>(L.g, n, c->f[VOX_F_X], c->f[VOX_F_Y],
                                                             c->f[VOX_F_Z], L.cell, L.cell_fill,
                                                             c->dflags);
        VOX_HIP(hipGetLastError());
    }
    {
        ProfScope ps(c, KC_SORT);
        VOX_TRY(scan_exclusive(c, L.cell_fill, L.cell_start, L.g.tnc, L.cell_start + L.g.tnc));
        VOX_TRY(fill_i32(c, L.cell_fill, 0, L.g.tnc));
        k_scatter<<<grid_for(n), kBlock, 0, c->stream>>>(n, L.cell, L.cell_start, L.cell_fill,
                                                           L.order_raw);
        // cells of fewer than ~24 expected particles: one thread per particle
        if (few_per_cell(c, L)) {
            k_cell_rank<<<grid_for(n), kBlock, 0, c->stream>>>(n, c->p.dim, L.cell, L.cell_start,
                                                                L.order_raw, L.rec, L.pos_of,
                                                                c->f[VOX_F_X],
                                                                c->f[VOX_F_Y], c->f[VOX_F_Z]);
        } else {
            long nb = (L.g.tnc + (kBlock / kWave) - 1) / (kBlock / kWave);
            k_cell_sort<<<(int)nb, kBlock, 0, c->stream>>>(L.g.tnc, c->p.dim, L.cell_start,
                                                            L.order_raw, L.rec, L.pos_of,
                                                            c->f[VOX_F_X],
                                                            c->f[VOX_F_Y], c->f[VOX_F_Z]);
        }
        VOX_HIP(hipGetLastError());
    }
    L.sorted = true;
    return search_one(c, which, re);
}

// One search pass over the sorted cells of list `which` (rows at the list's
// current capacity), on c->stream.
static int search_one(vox_ctx* c, int which, double re)
{
    DevList& L = c->list[which];
    const int n = c->tp;
    const int fo = which == 0 ? FLAG_OVERFLOW_S : FLAG_OVERFLOW_L;
    const int fm = which == 0 ? FLAG_MAXCNT_S : FLAG_MAXCNT_L;
    // compact rows (list L): d < re, the particle itself excluded
    const bool compact = (which == 1) && list_compact_ok(c);
    Grid gg = L.g;
    if (compact) gg.incl = 0;
    const double R = gg.incl ? re + c->p.delta : re;
    L.compact = compact;
    L.epoch = c->pos_epoch;
    // Walk per cell when cells hold enough particles to fill most of a wave
    // (expected count = cell volume / DL^dim), else per particle.
    const double per_cell = (c->p.dim == 3)
        ? L.g.hx * L.g.hx * L.g.hz / (c->p.dl * c->p.dl * c->p.dl)
        : L.g.hx * L.g.hx / (c->p.dl * c->p.dl);
    const int mode = c->tuned_mode[which] ? c->tuned_mode[which] : c->p.list_mode;
    const bool cell_walk = (mode == 2 || (mode == 0 && per_cell >= 24.0));
    ProfScope ps(c, which == 0 ? KC_LIST : KC_LIST_L);
    VOX_TRY(fill_i32(c, c->dflags + fo, 0, 1));
    // Compact rows of the cell walk hold positions in the cell-sorted records
    // (and sit at the particle's position): their consumers gather from
    // records in that order, where a row's consecutive entries (one cell's
    // particles) sit side by side.
    L.pos_rows = compact && cell_walk;
    if (cell_walk) {
        VOX_TRY(fill_i32(c, c->dflags + fm, 0, 1));
        if (L.pos_rows)
            k_build_list_xrow<true><<<(n + kWave - 1) / kWave, kWave, 0, c->stream>>>(
                gg, n, R, L.cell_start, L.rec, L.cap, L.ids, L.cnt, c->dflags, fo, fm);
        else
            k_build_list_xrow<false><<<(n + kWave - 1) / kWave, kWave, 0, c->stream>>>(
                gg, n, R, L.cell_start, L.rec, L.cap, L.ids, L.cnt, c->dflags, fo, fm);
    } else {
        const int nb = grid_for(n);
        k_build_list<<<nb, kBlock, 0, c->stream>>>(gg, n, R, L.cell_start, L.rec, L.cap, L.ids,
                                                    L.cnt, c->dflags, fo, fm, c->blockmax);
        k_reduce_max<<<1, kBlock, 0, c->stream>>>(c->blockmax, nb, c->dflags + fm);
    }
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

// After read_flags: the domain error, naming the first particle outside the
// grid (its type, position and velocity).
static int domain_error(vox_ctx* c)
{
    const int i = c->hflags[FLAG_DOMAIN_FIRST];
    double v[5] = {0, 0, 0, 0, 0};
    int32_t t = 0;
    const int fs[5] = {VOX_F_X, VOX_F_Y, VOX_F_Z, VOX_F_U, VOX_F_V};
    if (i >= 0 && i < c->tp) {
        for (int k = 0; k < 5; ++k)
            VOX_HIP(hipMemcpy(&v[k], c->f[fs[k]] + i, sizeof(double), hipMemcpyDeviceToHost));
        VOX_HIP(hipMemcpy(&t, c->ptype + i, sizeof(int32_t), hipMemcpyDeviceToHost));
    }
    char buf[256];
    std::snprintf(buf, sizeof buf, "; first: particle %d (PTYPE %d) at (%.6g, %.6g, %.6g), velocity (%.6g, %.6g)",
                  i, t, v[0], v[1], v[2], v[3], v[4]);
    return fail(VOX_ERR_DOMAIN, "particle outside the neighbour grid box (" +
                                    std::to_string(c->hflags[FLAG_DOMAIN]) + " particles)" + buf);
}

// After read_flags: rows of list `which` that did not fit are searched again
// with wider rows (up to 3 more passes).
static int regrow_one(vox_ctx* c, int which, double re)
{
    const int fo = which == 0 ? FLAG_OVERFLOW_S : FLAG_OVERFLOW_L;
    const int fm = which == 0 ? FLAG_MAXCNT_S : FLAG_MAXCNT_L;
    for (int attempt = 0; attempt < 4; ++attempt) {
        if (c->hflags[fo] == 0) return VOX_OK;
        if (attempt == 3) break;
        int need = c->hflags[fm];
        int cap = ((need + need / 4) + 7) & ~7;
        VOX_TRY(ensure_list_storage(c, which, cap));
        VOX_TRY(search_one(c, which, re));
        VOX_TRY(read_flags(c));
    }
    return fail(VOX_ERR_OVERFLOW, "neighbour row capacity could not be satisfied");
}

// The reference rows of list `which` in place of compact ones, searched again
// over the cell-sorted records of the NEIGHBOR call that built the list (the
// positions of that call, whatever the particles did since), by the cell-row
// walk, with wider rows while any overflows.
int ensure_reference_list(vox_ctx* c, int which)
{
    DevList& L = c->list[which];
    if (!L.compact) return VOX_OK;
    const int fo = which == 0 ? FLAG_OVERFLOW_S : FLAG_OVERFLOW_L;
    const int fm = which == 0 ? FLAG_MAXCNT_S : FLAG_MAXCNT_L;
    const double R = L.g.incl ? L.re + c->p.delta : L.re;
    const int n = c->tp;
    for (int attempt = 0; attempt < 4; ++attempt) {
        VOX_TRY(fill_i32(c, c->dflags + fo, 0, 1));
        VOX_TRY(fill_i32(c, c->dflags + fm, 0, 1));
        k_build_list_xrow<false><<<(n + kWave - 1) / kWave, kWave, 0, c->stream>>>(
            L.g, n, R, L.cell_start, L.rec, L.cap, L.ids, L.cnt, c->dflags, fo, fm);
        VOX_HIP(hipGetLastError());
        L.pos_rows = false;
        VOX_TRY(read_flags(c));
        if (c->hflags[fo] == 0) {
            L.compact = false;
            return VOX_OK;
        }
        const int need = c->hflags[fm];
        VOX_TRY(ensure_list_storage(c, which, ((need + need / 4) + 7) & ~7));
    }
    return fail(VOX_ERR_OVERFLOW, "neighbour row capacity could not be satisfied");
}

// Called by every consumer of list L before it reads the rows: compact rows
// that no longer stand for the reference rows - the particles have moved since
// the search (the reference's rows, searched with the DELTA margin, are then
// read at other positions), or the settings have changed (a weight function
// that is not zero at re, matopt, srcopt) - are replaced by the reference rows.
int list_l_for_consumer(vox_ctx* c)
{
    const DevList& L = c->list[1];
    if (c->two_lists && L.compact && (L.epoch != c->pos_epoch || !list_compact_ok(c)))
        return ensure_reference_list(c, 1);
    return VOX_OK;
}

static int build_one(vox_ctx* c, int which, double re)
{
    VOX_TRY(launch_one(c, which, re));
    VOX_TRY(read_flags(c));
    if (c->hflags[FLAG_DOMAIN] != 0) return domain_error(c);
    return regrow_one(c, which, re);
}

// Tile-shape autotuner of the neighbour search: times both walks (one thread
// per particle, one wave per 64 cell-sorted particles) on the current positions, `reps` builds
// each, and keeps the faster one per list. The lists are identical either way,
// so the choice only changes speed. modes[w] receives 1 or 2 per list.
int autotune_lists(vox_ctx* c, int reps, int32_t* modes)
{
    if (reps < 1) reps = 1;
    hipEvent_t e0, e1;
    VOX_HIP(hipEventCreate(&e0));
    VOX_HIP(hipEventCreate(&e1));
    int rc = VOX_OK;
    for (int w = 0; w < (c->two_lists ? 2 : 1) && rc == VOX_OK; ++w) {
        const double re = (w == 0) ? c->p.re_s : c->p.re_l;
        float best = 0.f;
        int pick = 1;
        for (int mode = 1; mode <= 2 && rc == VOX_OK; ++mode) {
            c->tuned_mode[w] = mode;
            if ((rc = build_one(c, w, re)) != VOX_OK) break;       // warm (and grows rows)
            (void)hipEventRecord(e0, c->stream);
            for (int r = 0; r < reps && rc == VOX_OK; ++r) rc = build_one(c, w, re);
            (void)hipEventRecord(e1, c->stream);
            if (rc != VOX_OK) break;
            if (hipEventSynchronize(e1) != hipSuccess) { rc = fail(VOX_ERR_HIP, "autotune sync"); break; }
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (mode == 1 || ms < best) { best = ms; pick = mode; }
        }
        c->tuned_mode[w] = pick;
        if (modes) modes[w] = pick;
    }
    if (!c->two_lists) {
        c->tuned_mode[1] = c->tuned_mode[0];
        if (modes) modes[1] = modes[0];
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

// List L on the side stream: the context's stream, scan scratch and
// block-maximum slots are swapped for the side ones while its work is queued.
static int launch_side(vox_ctx* c, int which, double re)
{
    std::swap(c->stream, c->stream2);
    std::swap(c->scan_blocks, c->scan_blocks2);
    std::swap(c->scan_blocks_cap, c->scan_blocks2_cap);
    std::swap(c->blockmax, c->blockmax2);
    const int rc = launch_one(c, which, re);
    std::swap(c->stream, c->stream2);
    std::swap(c->scan_blocks, c->scan_blocks2);
    std::swap(c->scan_blocks_cap, c->scan_blocks2_cap);
    std::swap(c->blockmax, c->blockmax2);
    return rc;
}

int build_neighbors(vox_ctx* c)
{
    VOX_TRY(fill_i32(c, c->dflags + FLAG_DOMAIN, 0, 1));
    VOX_TRY(fill_i32(c, c->dflags + FLAG_DOMAIN_FIRST, 0x7fffffff, 1));
    if (c->two_lists && c->stream2) {
        // The two lists are independent: list L is built on a second stream
        // beside list S (the side stream first waits for everything queued so
        // far), and the main stream waits for it before the one flag read.
        VOX_HIP(hipEventRecord(c->ev_fork, c->stream));
        VOX_HIP(hipStreamWaitEvent(c->stream2, c->ev_fork, 0));
        VOX_TRY(launch_one(c, 0, c->p.re_s));
        VOX_TRY(launch_side(c, 1, c->p.re_l));
        VOX_HIP(hipEventRecord(c->ev_join, c->stream2));
        VOX_HIP(hipStreamWaitEvent(c->stream, c->ev_join, 0));
        VOX_TRY(read_flags(c));
        if (c->hflags[FLAG_DOMAIN] != 0) return domain_error(c);
        VOX_TRY(regrow_one(c, 0, c->p.re_s));
        VOX_TRY(regrow_one(c, 1, c->p.re_l));
    } else {
        VOX_TRY(build_one(c, 0, c->p.re_s));
        if (c->two_lists) VOX_TRY(build_one(c, 1, c->p.re_l));
    }
    c->st.max_count_s = c->hflags[FLAG_MAXCNT_S];
    c->st.max_count_l = c->two_lists ? c->hflags[FLAG_MAXCNT_L] : c->hflags[FLAG_MAXCNT_S];
    return VOX_OK;
}

}  // namespace vox
