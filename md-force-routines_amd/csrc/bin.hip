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

// One block of 1024 threads; each thread scans a contiguous run of cells.
__global__ __launch_bounds__(1024) void k_scan_cells(int *__restrict__ cell_count, int *__restrict__ cell_start,
                                                     int ncell) {
    __shared__ int wsum[16];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int per = (ncell + 1023) / 1024;
    const int lo = min(t * per, ncell), hi = min(lo + per, ncell);
    int s = 0;
    for (int i = lo; i < hi; ++i) s += cell_count[i];
    // inclusive scan inside the wave
    int v = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    if (lane == 63) wsum[wv] = v;
    __syncthreads();
    if (t < 64) {
        int x = t < 16 ? wsum[t] : 0;
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
            const int u = __shfl_up(x, o, 64);
            if (t >= o) x += u;
        }
        if (t < 16) wsum[t] = x;
    }
    __syncthreads();
    int run = v - s + (wv > 0 ? wsum[wv - 1] : 0);
    for (int i = lo; i < hi; ++i) {
        const int cnt = cell_count[i];
        cell_start[i] = run;
        run += cnt;
        cell_count[i] = 0;  // ready for the next binning
    }
    if (t == 1023) cell_start[ncell] = wsum[15];
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
    hipLaunchKernelGGL(k_scan_cells, dim3(1), dim3(1024), 0, s, c->cell_count, c->cell_start, c->ncell);
    MDF_LAUNCHED();
    hipLaunchKernelGGL(k_bin_place, dim3(nb), dim3(256), 0, s, pos, Z, n, c->g, c->cell_start, c->part_cell,
                       c->part_slot, c->spos, c->perm);
    MDF_LAUNCHED();
    return MDF_OK;
}

}  // namespace mdf
