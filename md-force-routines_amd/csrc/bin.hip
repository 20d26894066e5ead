// Counting-sort binning of particles into PP sub-cells.
//
// Replaces the serial head/linked-list build of yukawa_pp.py:72-80 (and
// yukawa_p3m.c:265-291).  Sub-cells are half a reference cell wide (the
// reference cell is L/floor(L/rc), yukawa_pp.py:45-51), so a 2x2x2 block of
// sub-cells is exactly one reference cell.  Three kernels:
//   1. count: sub-cell of each particle, slot = atomicAdd(count[cell])
//   2. scan : exclusive prefix sum of the counts (and reset of the counts)
//   3. place: sorted index = start[cell] + slot; sorted (x, y, z, Z) copy.
// Positions are wrapped once into [0, L] for the sorted copy (a position at
// exactly L, which the reference's cell index leaves out of range, lands in
// the periodic image at 0); in-range positions are copied bit for bit.
#include "mdf_internal.h"

namespace mdf {

__device__ __forceinline__ double wrap_box(double x, double L) {
    if (x < 0.0) x += L;
    if (x >= L) x -= L;
    return x;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__global__ void k_bin_count(const double *__restrict__ pos, int n, Geom g, int *__restrict__ cell_count,
                            int *__restrict__ part_cell, int *__restrict__ part_slot) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
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

// Exclusive scan of the cell counts in three passes over tiles of
// kScanTile cells: per-tile sums, a scan of the tile sums (one block), and a
// per-tile scan that adds the tile offset, writes cell_start and resets the
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

__global__ __launch_bounds__(256) void k_scan_tile_sums(const int *__restrict__ cell_count, int ncell,
                                                        int *__restrict__ tile_sum) {
    __shared__ int wsum[4];
    const int base = blockIdx.x * kScanTile + threadIdx.x * 4;
    int s = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) s += (base + q < ncell) ? cell_count[base + q] : 0;
    const int incl = block_incl_scan_256(s, wsum);
    if (threadIdx.x == 255) tile_sum[blockIdx.x] = incl;
}

// One block of 1024 threads scans up to 1024 * 8 tile sums in place (exclusive).
__global__ __launch_bounds__(1024) void k_scan_tiles(int *__restrict__ tile_sum, int ntile,
                                                     int *__restrict__ total) {
    __shared__ int wsum[16];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    constexpr int per = 8;
    int v[per];
    int s = 0;
#pragma unroll
    for (int q = 0; q < per; ++q) {
        const int k = t * per + q;
        v[q] = k < ntile ? tile_sum[k] : 0;
        s += v[q];
    }
    int incl = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(incl, o, 64);
        if (lane >= o) incl += u;
    }
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    int add = 0;
    for (int w = 0; w < wv; ++w) add += wsum[w];
    int run = incl - s + add;
#pragma unroll
    for (int q = 0; q < per; ++q) {
        const int k = t * per + q;
        if (k < ntile) tile_sum[k] = run;
        run += v[q];
    }
    if (t == 1023) *total = run;
}

__global__ __launch_bounds__(256) void k_scan_apply(int *__restrict__ cell_count, int ncell,
                                                    const int *__restrict__ tile_off, const int *__restrict__ total,
                                                    int *__restrict__ cell_start) {
    __shared__ int wsum[4];
    const int base = blockIdx.x * kScanTile + threadIdx.x * 4;
    int v[4];
    int s = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        v[q] = (base + q < ncell) ? cell_count[base + q] : 0;
        s += v[q];
    }
    const int incl = block_incl_scan_256(s, wsum);
    int run = tile_off[blockIdx.x] + incl - s;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (base + q < ncell) {
            cell_start[base + q] = run;
            cell_count[base + q] = 0;  // ready for the next binning
        }
        run += v[q];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) cell_start[ncell] = *total;
}

__global__ void k_bin_place(const double *__restrict__ pos, const double *__restrict__ Z, int n, Geom g,
                            const int *__restrict__ cell_start, const int *__restrict__ part_cell,
                            const int *__restrict__ part_slot, double4 *__restrict__ spos,
                            int *__restrict__ perm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = cell_start[part_cell[i]] + part_slot[i];
    const double x = wrap_box(pos[3 * i + 0], g.L[0]);
    const double y = wrap_box(pos[3 * i + 1], g.L[1]);
    const double z = wrap_box(pos[3 * i + 2], g.L[2]);
    spos[s] = make_double4(x, y, z, Z ? Z[i] : 1.0);
    perm[s] = i;
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

int bin_particles(mdf_p3m_ctx *c, const double *pos, const double *Z, int n, hipStream_t s) {
    if (n <= 0) return MDF_OK;
    const unsigned nb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_bin_count, dim3(nb), dim3(256), 0, s, pos, n, c->g, c->cell_count, c->part_cell,
                       c->part_slot);
    MDF_LAUNCHED();
    const int ntile = (c->ncell + kScanTile - 1) / kScanTile;
    int *tile = c->scan_tmp;
    hipLaunchKernelGGL(k_scan_tile_sums, dim3(ntile), dim3(256), 0, s, c->cell_count, c->ncell, tile);
    MDF_LAUNCHED();
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, s, tile, ntile, tile + ntile);
    MDF_LAUNCHED();
    hipLaunchKernelGGL(k_scan_apply, dim3(ntile), dim3(256), 0, s, c->cell_count, c->ncell, tile, tile + ntile,
                       c->cell_start);
    MDF_LAUNCHED();
    hipLaunchKernelGGL(k_bin_place, dim3(nb), dim3(256), 0, s, pos, Z, n, c->g, c->cell_start, c->part_cell,
                       c->part_slot, c->spos, c->perm);
    MDF_LAUNCHED();
    return MDF_OK;
}

}  // namespace mdf
