// Counting-sort binning of particles into PP sub-cells.
//
// Replaces the serial head/linked-list build of yukawa_pp.py:44-52 (and
// yukawa_p3m.c:265-291).  Sub-cells are half a reference cell wide (the
// reference cell is L/floor(L/rc), yukawa_pp.py:17-23), so a 2x2x2 block of
// sub-cells is exactly one reference cell.  Three kernels:
//   1. count: sub-cell of each particle, slot = atomicAdd(count[cell])
//   2. scan : exclusive prefix sum of the counts (and reset of the counts)
//   3. place: sorted index = start[cell] + slot; sorted (x, y, z, Z) copy.
// With Verlet pair lists (vlist.hip) the sorted order is frozen between list
// builds: on a list step the three kernels return at once (the step's mode is
// on the device) and the sorted copy holds the positions in the frozen order.
// Positions are wrapped into [0, L) for the sorted copy (a position at
// exactly L, which the reference's cell index leaves out of range, lands in
// the periodic image at 0); in-range positions are copied bit for bit.
// Every kernel that turns a position into a cell or mesh index reads the
// sorted copy, so the wrap must hold for any input: the Verlet drift wraps
// once per step (velocity_verlet_p3m_ord6.py:12-19), and an ion faster than
// L/dt leaves the box by more than one period.  Such positions are reduced
// with floor(x / L); a non-finite one is placed at 0.
#include "mdf_internal.h"

namespace mdf {

__device__ __forceinline__ double wrap_box(double x, double L) {
    if (x < 0.0) x += L;
    if (x >= L) x -= L;
    if (!(x >= 0.0 && x < L)) {
        x = __builtin_isfinite(x) ? x - L * floor(x / L) : 0.0;
        if (!(x >= 0.0 && x < L)) x = 0.0;  // x - L floor(x/L) may round to L
    }
    return x;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

struct BinArgs {
    const double *pos;
    const double *Z;
    int n;
    Geom g;
    int *cell_count;
    int *part_cell;
    int *part_slot;
    int ncell;
    int *tile;         // scan tile sums (+ total at tile[ntile])
    int ntile;
    int *cell_start;
    double4 *spos;
    int *perm;
    const int *mode;   // pair-list mode of the step (2: lists, binning skipped), or null
    double4 *vl_ref;   // pair lists' build positions, written beside spos on a build, or null
};

__device__ __forceinline__ void bin_count_body(const BinArgs &A, const Blk &b) {
    const double *__restrict__ pos = A.pos;
    const Geom &g = A.g;
    const int n = A.n;
    int *__restrict__ cell_count = A.cell_count;
    int *__restrict__ part_cell = A.part_cell;
    int *__restrict__ part_slot = A.part_slot;
    const int i = b.x * blockDim.x + threadIdx.x;
    if (i >= n || (A.mode && *A.mode == 2)) return;
    int c[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double x = wrap_box(pos[3 * i + d], g.L[d]);
        c[d] = clampi((int)floor(x / g.w[d]), 0, g.nsub[d] - 1);
    }
    const int cell = c[0] + g.nsub[0] * (c[1] + g.nsub[1] * c[2]);
    part_cell[i] = cell;
    part_slot[i] = atomicAdd(&cell_count[cell], 1);
}

__global__ void k_bin_count(BinArgs A) { bin_count_body(A, blk_plain()); }
__global__ void k_bin_count_seg(const BinArgs *__restrict__ tab, SegGrid S) {
    Blk b;
    const int r = seg_locate(S, b);
    bin_count_body(tab[r], b);
}

// Exclusive scan of the cell counts in two passes over tiles of kScanTile
// cells: per-tile sums, then a per-tile scan that adds the sum of the earlier
// tiles' sums (each block adds those up itself: a few hundred tiles at most
// in practice, 8192 at the cell limit), writes cell_start and resets the
// counts for the next binning.
constexpr int kScanTile = 1024;  // 256 threads x 4 cells

__device__ __forceinline__ int block_incl_scan_256(int v, int *wsum) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    if (lane == 63) wsum[wv] = v;
    __syncthreads();
    int add = 0;
    for (int w = 0; w < wv; ++w) add += wsum[w];
    __syncthreads();
    return v + add;
}

__device__ __forceinline__ void scan_tile_sums_body(const BinArgs &A, const Blk &b, int *wsum) {
    if (A.mode && *A.mode == 2) return;  // uniform over the launch
    const int *__restrict__ cell_count = A.cell_count;
    const int ncell = A.ncell;
    const int base = b.x * kScanTile + threadIdx.x * 4;
    int s = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) s += (base + q < ncell) ? cell_count[base + q] : 0;
    const int incl = block_incl_scan_256(s, wsum);
    if (threadIdx.x == 255) A.tile[b.x] = incl;
}

__global__ __launch_bounds__(256) void k_scan_tile_sums(BinArgs A) {
    __shared__ int wsum[4];
    scan_tile_sums_body(A, blk_plain(), wsum);
}
__global__ __launch_bounds__(256) void k_scan_tile_sums_seg(const BinArgs *__restrict__ tab, SegGrid S) {
    __shared__ int wsum[4];
    Blk b;
    const int r = seg_locate(S, b);
    scan_tile_sums_body(tab[r], b, wsum);
}

__device__ __forceinline__ void scan_apply_body(const BinArgs &A, const Blk &b, int *wsum) {
    if (A.mode && *A.mode == 2) return;  // uniform over the launch
    int *__restrict__ cell_count = A.cell_count;
    const int ncell = A.ncell;
    const int *__restrict__ tile_sum = A.tile;
    int *__restrict__ cell_start = A.cell_start;
    // offset of this tile: the sum of the earlier tiles' sums
    int part = 0;
    for (int k = threadIdx.x; k < b.x; k += blockDim.x) part += tile_sum[k];
    const int part_incl = block_incl_scan_256(part, wsum);
    if (threadIdx.x == 255) wsum[4] = part_incl;
    __syncthreads();
    const int tile_off = wsum[4];
    const int base = b.x * kScanTile + threadIdx.x * 4;
    int v[4];
    int s = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        v[q] = (base + q < ncell) ? cell_count[base + q] : 0;
        s += v[q];
    }
    const int incl = block_incl_scan_256(s, wsum);
    int run = tile_off + incl - s;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (base + q < ncell) {
            cell_start[base + q] = run;
            cell_count[base + q] = 0;  // ready for the next binning
        }
        run += v[q];
    }
    // the last tile's last thread holds the total (cells past ncell count 0)
    if (b.x == b.nx - 1 && threadIdx.x == 255) cell_start[ncell] = run;
}

__global__ __launch_bounds__(256) void k_scan_apply(BinArgs A) {
    __shared__ int wsum[5];
    scan_apply_body(A, blk_plain(), wsum);
}
__global__ __launch_bounds__(256) void k_scan_apply_seg(const BinArgs *__restrict__ tab, SegGrid S) {
    __shared__ int wsum[5];
    Blk b;
    const int r = seg_locate(S, b);
    scan_apply_body(tab[r], b, wsum);
}

__device__ __forceinline__ void bin_place_body(const BinArgs &A, const Blk &b) {
    const double *__restrict__ pos = A.pos;
    const double *__restrict__ Z = A.Z;
    const Geom &g = A.g;
    const int n = A.n;
    const int *__restrict__ cell_start = A.cell_start;
    const int *__restrict__ part_cell = A.part_cell;
    const int *__restrict__ part_slot = A.part_slot;
    double4 *__restrict__ spos = A.spos;
    int *__restrict__ perm = A.perm;
    const int i = b.x * blockDim.x + threadIdx.x;
    if (i >= n || (A.mode && *A.mode == 2)) return;
    const int s = cell_start[part_cell[i]] + part_slot[i];
    const double x = wrap_box(pos[3 * i + 0], g.L[0]);
    const double y = wrap_box(pos[3 * i + 1], g.L[1]);
    const double z = wrap_box(pos[3 * i + 2], g.L[2]);
    const double4 v = make_double4(x, y, z, Z ? Z[i] : 1.0);
    spos[s] = v;
    perm[s] = i;
    if (A.vl_ref) A.vl_ref[s] = v;
}

__global__ void k_bin_place(BinArgs A) { bin_place_body(A, blk_plain()); }
__global__ void k_bin_place_seg(const BinArgs *__restrict__ tab, SegGrid S) {
    Blk b;
    const int r = seg_locate(S, b);
    bin_place_body(tab[r], b);
}

__global__ void k_fill_zero(double *__restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

__global__ void k_fill_zero_i(int *__restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

static unsigned grid_for(size_t n, unsigned bs) {
    size_t b = (n + bs - 1) / bs;
    if (b > 65536) b = 65536;
    return (unsigned)(b ? b : 1);
}

int fill_zero(double *p, size_t n, hipStream_t s) {
    if (!n) return MDF_OK;
    hipLaunchKernelGGL(k_fill_zero, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n);
    MDF_LAUNCHED();
    return MDF_OK;
}

int fill_zero_i(int *p, size_t n, hipStream_t s) {
    if (!n) return MDF_OK;
    hipLaunchKernelGGL(k_fill_zero_i, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n);
    MDF_LAUNCHED();
    return MDF_OK;
}

static BinArgs bin_args(mdf_p3m_ctx *c, const double *pos, const double *Z, int n) {
    BinArgs A;
    A.pos = pos;
    A.Z = Z;
    A.n = n;
    A.g = c->g;
    A.cell_count = c->cell_count;
    A.part_cell = c->part_cell;
    A.part_slot = c->part_slot;
    A.ncell = c->ncell;
    A.tile = c->scan_tmp;
    A.ntile = (c->ncell + kScanTile - 1) / kScanTile;
    A.cell_start = c->cell_start;
    A.spos = c->spos;
    A.perm = c->perm;
    A.mode = nullptr;
    A.vl_ref = nullptr;
    return A;
}

int bin_particles_mode(mdf_p3m_ctx *c, const double *pos, const double *Z, int n, const int *mode, hipStream_t s,
                       double4 *vl_ref) {
    if (n <= 0) return MDF_OK;
    BinArgs A = bin_args(c, pos, Z, n);
    A.mode = mode;
    A.vl_ref = vl_ref;
    const unsigned nb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_bin_count, dim3(nb), dim3(256), 0, s, A);
    MDF_LAUNCHED();
    hipLaunchKernelGGL(k_scan_tile_sums, dim3(A.ntile), dim3(256), 0, s, A);
    MDF_LAUNCHED();
    hipLaunchKernelGGL(k_scan_apply, dim3(A.ntile), dim3(256), 0, s, A);
    MDF_LAUNCHED();
    hipLaunchKernelGGL(k_bin_place, dim3(nb), dim3(256), 0, s, A);
    MDF_LAUNCHED();
    return MDF_OK;
}

int bin_particles(mdf_p3m_ctx *c, const double *pos, const double *Z, int n, hipStream_t s) {
    // with pair lists the sorted arrays are the lists' frozen order: binning
    // outside a list-aware step replaces it, so the next step builds
    if (c->vl_skin > 0.0) c->vl_force_build = 1;
    return bin_particles_mode(c, pos, Z, n, nullptr, s);
}

// ---- segmented stages (replica batches) ----
// With pair lists the binning returns at once on a list step (the mode the
// decision stage wrote), as in bin_step.
static BinArgs seg_bin_args(mdf_p3m_ctx *c, const StepIO &io) {
    BinArgs A = bin_args(c, io.pos, nullptr, io.n);
    if (c->vl_skin > 0.0 && io.n > 0) A.mode = c->vl_flags + 3;
    return A;
}
static int build_bin_particles(mdf_p3m_ctx *c, const StepIO &io, void *args, int4 &dims, size_t &lds) {
    *(BinArgs *)args = seg_bin_args(c, io);
    dims = make_int4((io.n + 255) / 256 > 0 ? (io.n + 255) / 256 : 1, 1, 1, 0);
    lds = 0;
    return MDF_OK;
}
static int build_bin_tiles(mdf_p3m_ctx *c, const StepIO &io, void *args, int4 &dims, size_t &lds) {
    const BinArgs A = seg_bin_args(c, io);
    *(BinArgs *)args = A;
    dims = make_int4(A.ntile, 1, 1, 0);
    lds = 0;
    return MDF_OK;
}
#define MDF_SEG_LAUNCH(fname, kern, threads, Args)                                                        \
    static int fname(const void *tab, const SegGrid &S, int total, size_t lds, int k, hipStream_t s) {   \
        (void)lds;                                                                                          \
        (void)k;                                                                                            \
        hipLaunchKernelGGL(kern, dim3(total), dim3(threads), 0, s, (const Args *)tab, S);                   \
        MDF_LAUNCHED();                                                                                     \
        return MDF_OK;                                                                                      \
    }
MDF_SEG_LAUNCH(launch_bin_count, k_bin_count_seg, 256, BinArgs)
MDF_SEG_LAUNCH(launch_scan_tile_sums, k_scan_tile_sums_seg, 256, BinArgs)
MDF_SEG_LAUNCH(launch_scan_apply, k_scan_apply_seg, 256, BinArgs)
MDF_SEG_LAUNCH(launch_bin_place, k_bin_place_seg, 256, BinArgs)
#undef MDF_SEG_LAUNCH

void seg_stages_bin(std::vector<SegStage> &out) {
    out.push_back({"bin_count", sizeof(BinArgs), build_bin_particles, launch_bin_count});
    out.push_back({"scan_tile_sums", sizeof(BinArgs), build_bin_tiles, launch_scan_tile_sums});
    out.push_back({"scan_apply", sizeof(BinArgs), build_bin_tiles, launch_scan_apply});
    out.push_back({"bin_place", sizeof(BinArgs), build_bin_particles, launch_bin_place});
}

}  // namespace mdf
