// Short-range Yukawa-Ewald PP forces and energy on a sub-cell list.
//
// Reference: particle_particle, yukawa_pp.py:6-167 (C twin yukawa_p3m.c:182-416).
// Same pair set (every pair with minimum-image distance r < rc, rc <= L/3),
// same pair terms (yukawa_pp.py:143-152), same displacement convention
// d = x_i - (x_j + shift) (:133-135).  Differences are only in how the work is
// organised and in summation order:
//
// * One workgroup (8 waves) per reference cell = a 2x2x2 block of half-width
//   sub-cells.  The 6x6x6 sub-cells around the block (the block plus two
//   sub-cells on every side) are staged in LDS with their periodic image
//   shift already applied, as 36 rows of 6 sub-cells; a row is one or two
//   contiguous runs of the sorted particle array.
// * Each wave owns one i at a time and sweeps the 5x5 rows of i's stencil,
//   culling rows (and the outer sub-cells of a row) that lie farther than rc
//   from i.  Lanes test 64 candidates at once; in-cutoff j are compacted
//   into a per-wave LDS queue with ballot + popcount, and the expensive pair
//   terms run only on full 64-entry queue batches, so lanes stay busy
//   despite the ~20% in-cutoff rate of a cell stencil.
// * Each pair is evaluated from both ends (no Newton-3 scatter, no atomics);
//   i's force and energy are summed across the wave with shuffles.  The
//   energy is halved at the end: U_s_r = 1/2 sum_i sum_{j != i} phi.
// * Pair terms use e^{+kr} erfc(ar + b) = erfcx(ar + b) e^{-(ar)^2 - b^2} and
//   e^{-kr} erfc(ar - b) = erfcx(ar - b) e^{-(ar)^2 - b^2} (b = kappa/2alpha),
//   and e^{-(ar+b)^2} e^{kr} = e^{-(ar-b)^2} e^{-kr} = e^{-(ar)^2 - b^2}: one exp
//   and two erfcx per pair instead of eight exp and four erfc.
// * If a tile exceeds the LDS capacity the rows are processed in several
//   passes (i's force accumulates in global memory across passes).
#include "mdf_internal.h"

namespace mdf {

struct PPArgs {
    Geom g;
    const double4 *__restrict__ spos;
    const int *__restrict__ cell_start;
    double *__restrict__ acc;     // sorted (n, 3)
    double *__restrict__ part;    // per-block energy partial
    int *__restrict__ stats;
    int tile_max;
    double b;                     // kappa / (2 alpha)
    double b2;
    double two_alpha_sqrtpi;      // 2 alpha / sqrt(pi)
    double rc2m;                  // candidate test radius^2 (rc slightly widened)
};

__device__ __forceinline__ void pp_pair_fp64(const PPArgs &P, double dx, double dy, double dz, double &fx,
                                             double &fy, double &fz, double &u) {
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    if (!(r < P.g.rc)) return;  // the reference's exact test, yukawa_pp.py:140
    const double ir = 1.0 / r;
    const double ar = P.g.alpha * r;
    const double ge = exp(-(ar * ar) - P.b2);
    const double ep = erfcx(ar + P.b) * ge;  // e^{kr} erfc(ar + b)
    const double em = erfcx(ar - P.b) * ge;  // e^{-kr} erfc(ar - b)
    const double kr = P.g.kappa * r;
    u += 0.5 * ir * (ep + em);
    const double f = 0.5 * ir * ir * (ep * (1.0 - kr) + em * (1.0 + kr)) + P.two_alpha_sqrtpi * ir * ge;
    const double fr = f * ir;
    fx += fr * dx;
    fy += fr * dy;
    fz += fr * dz;
}

// Broadcast lane 0's value (the caller guarantees it is wave-uniform) so the
// compiler keeps it in scalar registers.
__device__ __forceinline__ double uniform(double v) {
    const unsigned long long b = __double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)(b & 0xffffffffull));
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ void wrap_cell(int c, int n, double L, int &cw, double &shift) {
    if (c < 0) {
        cw = c + n;
        shift = -L;
    } else if (c >= n) {
        cw = c - n;
        shift = L;
    } else {
        cw = c;
        shift = 0.0;
    }
}

// Distance from offset o (in [0, w)) of i inside its sub-cell to the slab of
// the sub-cell d steps away along one axis.
__device__ __forceinline__ double slab_dist(int d, double o, double w) {
    double r = 0.0;
    if (d > 0) r = d * w - o;
    else if (d < 0) r = o - (d + 1) * w;
    return r > 0.0 ? r : 0.0;
}

__global__ __launch_bounds__(kPPThreads) void k_pp_fp64(PPArgs P) {
    extern __shared__ double lds[];
    const int T = P.tile_max;
    double *tx = lds;
    double *ty = tx + T;
    double *tz = ty + T;
    int *cum = (int *)(tz + T);          // 217 prefix offsets of the tile sub-cells
    int *queue = cum + 224;              // kPPWaves * kPPQueue
    __shared__ int tcell[216];
    __shared__ double red[kPPWaves];

    const Geom &g = P.g;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int bid = blockIdx.x;
    const int bx = bid % g.nblk[0];
    const int by = (bid / g.nblk[0]) % g.nblk[1];
    const int bz = bid / (g.nblk[0] * g.nblk[1]);

    // 1. counts of the 216 tile sub-cells, tile order x fastest
    if (tid < 216) {
        const int txs = tid % 6, tys = (tid / 6) % 6, tzs = tid / 36;
        int cx, cy, cz;
        double sdummy;
        wrap_cell(2 * bx - 2 + txs, g.nsub[0], g.L[0], cx, sdummy);
        wrap_cell(2 * by - 2 + tys, g.nsub[1], g.L[1], cy, sdummy);
        wrap_cell(2 * bz - 2 + tzs, g.nsub[2], g.L[2], cz, sdummy);
        const int cell = cx + g.nsub[0] * (cy + g.nsub[1] * cz);
        tcell[tid] = cell;
        cum[tid + 1] = P.cell_start[cell + 1] - P.cell_start[cell];
    }
    __syncthreads();
    // 2. exclusive prefix over the 216 counts (wave 0: 4 entries per lane)
    if (wave == 0) {
        int v[4];
        int s = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int k = lane * 4 + q;
            v[q] = k < 216 ? cum[k + 1] : 0;
            s += v[q];
        }
        int incl = s;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int u = __shfl_up(incl, o, 64);
            if (lane >= o) incl += u;
        }
        int run = incl - s;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int k = lane * 4 + q;
            if (k < 216) cum[k] = run;
            run += v[q];
        }
        if (lane == 63) cum[216] = incl;
    }
    __syncthreads();

    // home block: tile sub-cells (2+hx, 2+hy, 2+hz), hx,hy,hz in {0,1}
    int hoff[9];
    hoff[0] = 0;
#pragma unroll
    for (int h = 0; h < 8; ++h) {
        const int idx = (2 + (h & 1)) + 6 * ((2 + ((h >> 1) & 1)) + 6 * (2 + (h >> 2)));
        hoff[h + 1] = hoff[h] + (cum[idx + 1] - cum[idx]);
    }
    const int n_home = hoff[8];
    int *q = queue + wave * kPPQueue;
    const double rc2m = P.rc2m;
    double u_acc = 0.0;

    int r0 = 0, pass = 0;
    while (r0 < 36) {
        const int base = cum[6 * r0];
        int r1 = r0;
        while (r1 < 36 && cum[6 * (r1 + 1)] - base <= T) ++r1;
        if (r1 == r0) {  // a single row larger than the LDS tile: unsupported density
            if (tid == 0) atomicOr(&P.stats[2], 1);
            r1 = r0 + 1;
            r0 = r1;
            continue;
        }
        if (pass == 1 && tid == 0) atomicAdd(&P.stats[0], 1);
        // 3. stage rows [r0, r1) with their image shifts
        for (int r = r0 + wave; r < r1; r += kPPWaves) {
            const int tys = r % 6, tzs = r / 6;
            int cy, cz;
            double sy, sz;
            wrap_cell(2 * by - 2 + tys, g.nsub[1], g.L[1], cy, sy);
            wrap_cell(2 * bz - 2 + tzs, g.nsub[2], g.L[2], cz, sz);
            int txs = 0;
            while (txs < 6) {
                int cx0;
                double sx;
                wrap_cell(2 * bx - 2 + txs, g.nsub[0], g.L[0], cx0, sx);
                int txe = txs + 1;
                while (txe < 6) {
                    int cxe;
                    double sxe;
                    wrap_cell(2 * bx - 2 + txe, g.nsub[0], g.L[0], cxe, sxe);
                    if (sxe != sx) break;
                    ++txe;
                }
                const int c0 = cx0 + g.nsub[0] * (cy + g.nsub[1] * cz);
                const int gs = P.cell_start[c0];
                const int ge = P.cell_start[c0 + (txe - txs)];
                const int dst = cum[6 * r + txs] - base;
                for (int t = lane; t < ge - gs; t += 64) {
                    const double4 p = P.spos[gs + t];
                    tx[dst + t] = p.x + sx;
                    ty[dst + t] = p.y + sy;
                    tz[dst + t] = p.z + sz;
                }
                txs = txe;
            }
        }
        __syncthreads();

        // 4. one i per wave
        for (int k = wave; k < n_home; k += kPPWaves) {
            int h = 0, hb = 0;
#pragma unroll
            for (int hh = 1; hh < 8; ++hh)
                if (k >= hoff[hh]) {
                    h = hh;
                    hb = hoff[hh];
                }
            const int hx = h & 1, hy = (h >> 1) & 1, hz = h >> 2;
            const int hidx = (2 + hx) + 6 * ((2 + hy) + 6 * (2 + hz));
            const int gi = __builtin_amdgcn_readfirstlane(P.cell_start[tcell[hidx]] + (k - hb));
            const double4 pi = P.spos[gi];
            const double xi = uniform(pi.x), yi = uniform(pi.y), zi = uniform(pi.z);
            const double ox = xi - (2 * bx + hx) * g.w[0];
            const double oy = yi - (2 * by + hy) * g.w[1];
            const double oz = zi - (2 * bz + hz) * g.w[2];
            double fx = 0.0, fy = 0.0, fz = 0.0, u = 0.0;
            int qn = 0;
            for (int dz = -2; dz <= 2; ++dz) {
                const double sdz = slab_dist(dz, oz, g.w[2]);
                const double d2z = sdz * sdz;
                if (d2z > rc2m) continue;
                for (int dy = -2; dy <= 2; ++dy) {
                    const double sdy = slab_dist(dy, oy, g.w[1]);
                    const double d2yz = sdy * sdy + d2z;
                    if (d2yz > rc2m) continue;
                    const int r = (2 + hy + dy) + 6 * (2 + hz + dz);
                    if (r < r0 || r >= r1) continue;
                    const double sxl = ox + g.w[0];
                    const double sxr = 2.0 * g.w[0] - ox;
                    const int xa = hx + ((sxl * sxl + d2yz > rc2m) ? 1 : 0);
                    const int xb = hx + 4 - ((sxr * sxr + d2yz > rc2m) ? 1 : 0);
                    const int lo = cum[6 * r + xa] - base;
                    const int hi = cum[6 * r + xb + 1] - base;
                    for (int j0 = lo; j0 < hi; j0 += 64) {
                        const int j = j0 + lane;
                        bool hit = false;
                        if (j < hi) {
                            const double ddx = xi - tx[j], ddy = yi - ty[j], ddz = zi - tz[j];
                            const double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
                            hit = (r2 < rc2m) && (r2 > 0.0);
                        }
                        const unsigned long long m = __ballot(hit);
                        if (hit) q[qn + __popcll(m & ((1ull << lane) - 1ull))] = j;
                        qn += __popcll(m);
                        if (qn >= 64) {
                            const int jj = q[lane];
                            pp_pair_fp64(P, xi - tx[jj], yi - ty[jj], zi - tz[jj], fx, fy, fz, u);
                            const int rest = qn - 64;
                            const int moved = lane < rest ? q[64 + lane] : 0;
                            if (lane < rest) q[lane] = moved;
                            qn = rest;
                        }
                    }
                }
            }
            if (lane < qn) {
                const int jj = q[lane];
                pp_pair_fp64(P, xi - tx[jj], yi - ty[jj], zi - tz[jj], fx, fy, fz, u);
            }
            fx = wave_sum(fx);
            fy = wave_sum(fy);
            fz = wave_sum(fz);
            u = wave_sum(u);
            if (lane == 0) {
                double *a = P.acc + 3 * (size_t)gi;
                if (pass == 0) {
                    a[0] = fx;
                    a[1] = fy;
                    a[2] = fz;
                } else {
                    a[0] += fx;
                    a[1] += fy;
                    a[2] += fz;
                }
                u_acc += u;
            }
        }
        __syncthreads();
        r0 = r1;
        ++pass;
    }
    if (lane == 0) red[wave] = u_acc;
    __syncthreads();
    if (tid == 0) {
        double s = 0.0;
        for (int w = 0; w < kPPWaves; ++w) s += red[w];
        P.part[bid] = 0.5 * s;
    }
}

int pp_forces(mdf_p3m_ctx *c, int n, hipStream_t s) {
    (void)n;
    PPArgs P;
    P.g = c->g;
    P.spos = c->spos;
    P.cell_start = c->cell_start;
    P.acc = c->acc_pp_s;
    P.part = c->part_pp;
    P.stats = c->stats;
    P.tile_max = c->tile_max;
    P.b = 0.5 * c->g.kappa / c->g.alpha;
    P.b2 = P.b * P.b;
    P.two_alpha_sqrtpi = 2.0 * c->g.alpha / sqrt(M_PI);
    const double rcm = c->g.rc * (1.0 + 1e-12);
    P.rc2m = rcm * rcm;
    MDF_TRY(allow_lds(k_pp_fp64, c->pp_lds));
    hipLaunchKernelGGL(k_pp_fp64, dim3(c->nblock), dim3(kPPThreads), c->pp_lds, s, P);
    MDF_LAUNCHED();
    return MDF_OK;
}

}  // namespace mdf
