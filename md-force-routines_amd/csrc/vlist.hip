// Verlet pair lists for the PP stage.
//
// The reference builds its cell list and tests every candidate pair on every
// call (yukawa_pp.py:44-112).  With a skin s > 0 the PP kernel instead tests
// candidates only on a build step, against the list radius rc + s, and writes
// every candidate within it to a per-particle list of tile slots; the steps
// after a build evaluate the listed pairs with the exact cutoff r < rc.  The
// pair set is the reference's as long as no particle has moved more than s / 2
// since the build: a pair not listed was at least rc + s apart then and is
// still at least rc apart.  The forces are the reference's up to summation
// order.
//
// A list is indexed by the build's sorted order and its tile slots by the
// build's cell starts.  The binning keeps both until the next build ("frozen"
// order): on a list step it is skipped, and the positions of the step are
// gathered into the frozen order instead, each at the periodic image nearest
// its build position (so that a wrap by the drift does not move a particle by
// L in its tile), while the largest displacement since the build is recorded.
// Every stage of a list step (PP, assignment, gather) then works in the frozen
// order; the PM stages take the few ions that left their block's brick through
// global memory.
//
// Per step (bin_step):
//   k_vl_prep    positions into the frozen order (spos), max displacement^2
//   k_vl_decide  one block: build (requested, last build incomplete, or a
//                displacement over s / 2) or lists; the binning and the PP
//                kernel read the mode
//   binning      skipped on a list step
//   k_vl_refresh on a build step: the build positions := the new sorted copy
#include "mdf_internal.h"

namespace mdf {

constexpr int kPrepThreads = 256;
constexpr int kPrepPerThread = 4;  // particles a thread: the block maxima stay few

__global__ __launch_bounds__(kPrepThreads) void k_vl_prep(const double *__restrict__ pos, int n,
                                                          const int *__restrict__ perm,
                                                          const double4 *__restrict__ vl_ref,
                                                          double4 *__restrict__ spos, double Lx, double Ly, double Lz,
                                                          const int *__restrict__ flags, float *__restrict__ bmax) {
    __shared__ float wmax[kPrepThreads / 64];
    if (flags[0]) return;  // a build is requested: the frozen order is replaced anyway
    float d2f = 0.0f;
    for (int q = 0; q < kPrepPerThread; ++q) {
        const int s = (blockIdx.x * kPrepPerThread + q) * kPrepThreads + threadIdx.x;
        if (s >= n) break;
        const int i = perm[s];
        const double4 r = vl_ref[s];
        double dx = pos[3 * (size_t)i + 0] - r.x;
        double dy = pos[3 * (size_t)i + 1] - r.y;
        double dz = pos[3 * (size_t)i + 2] - r.z;
        dx -= Lx * rint(dx / Lx);
        dy -= Ly * rint(dy / Ly);
        dz -= Lz * rint(dz / Lz);
        spos[s] = make_double4(r.x + dx, r.y + dy, r.z + dz, r.w);
        const double d2 = dx * dx + dy * dy + dz * dz;
        d2f = fmaxf(d2f, d2 <= 1e30 ? (float)d2 : 1e30f);  // a non-finite position forces a build
    }
    // block maximum to its slot (one atomic on one address per wave serialised
    // ~0.15 ms at N = 1e6)
    for (int o = 32; o > 0; o >>= 1) d2f = fmaxf(d2f, __shfl_xor(d2f, o, 64));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = d2f;
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = wmax[0];
        for (int w = 1; w < kPrepThreads / 64; ++w) m = fmaxf(m, wmax[w]);
        bmax[blockIdx.x] = m;
    }
}

// The decision over the block maxima of k_vl_prep (nb of them; skipped when a
// build is already requested): one block, its thread 0 writes.
__global__ __launch_bounds__(1024) void k_vl_decide(int *__restrict__ flags, const float *__restrict__ bmax, int nb,
                                                    float thr) {
    __shared__ float wmax[16];
    __shared__ int forced;
    if (threadIdx.x == 0) forced = flags[0];
    __syncthreads();
    float m = 0.0f;
    if (!forced)
        for (int b = threadIdx.x; b < nb; b += blockDim.x) m = fmaxf(m, bmax[b]);
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fmaxf(m, wmax[w]);
    m = fmaxf(m, wmax[0]);
    const bool build = forced != 0 || flags[1] != 0 || m > thr;
    flags[3] = build ? 1 : 2;
    flags[0] = 0;
    if (build) {
        flags[1] = 0;
        flags[4] += 1;
    } else {
        flags[5] += 1;
    }
}

__global__ __launch_bounds__(256) void k_vl_refresh(const double4 *__restrict__ spos, int n,
                                                    double4 *__restrict__ vl_ref, const int *__restrict__ flags) {
    if (flags[3] != 1) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) vl_ref[t] = spos[t];
}

__global__ void k_vl_set(int *__restrict__ p, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

int bin_step(mdf_p3m_ctx *c, const double *pos, const double *Z, int n, hipStream_t s) {
    if (!(c->vl_skin > 0.0) || n <= 0) return bin_particles_mode(c, pos, Z, n, nullptr, s);
    if (n != c->vl_n) {  // the frozen order holds the last build's particles only
        c->vl_force_build = 1;
        c->vl_n = n;
    }
    if (c->vl_force_build) {
        hipLaunchKernelGGL(k_vl_set, dim3(1), dim3(64), 0, s, c->vl_flags, 1);
        MDF_LAUNCHED();
        c->vl_force_build = 0;
    }
    const int per_block = kPrepThreads * kPrepPerThread;
    const int nb = (n + per_block - 1) / per_block;
    if (nb > kMaxPartials) {
        set_error("pair lists: %d particles exceed the displacement partials", n);
        return MDF_EINVAL;
    }
    float *bmax = (float *)c->part_red;  // free until the n(q) stage of the step
    hipLaunchKernelGGL(k_vl_prep, dim3(nb), dim3(kPrepThreads), 0, s, pos, n, c->perm, c->vl_ref, c->spos,
                       c->g.L[0], c->g.L[1], c->g.L[2], c->vl_flags, bmax);
    MDF_LAUNCHED();
    // build when some particle moved more than skin / 2 (the float threshold
    // rounded down, so that an equal distance builds)
    const double h = 0.5 * c->vl_skin;
    float thr = (float)(h * h);
    if ((double)thr > h * h) thr = nextafterf(thr, 0.0f);
    hipLaunchKernelGGL(k_vl_decide, dim3(1), dim3(1024), 0, s, c->vl_flags, bmax, nb, thr);
    MDF_LAUNCHED();
    MDF_TRY(bin_particles_mode(c, pos, Z, n, c->vl_flags + 3, s));
    hipLaunchKernelGGL(k_vl_refresh, dim3((n + 255) / 256), dim3(256), 0, s, c->spos, n, c->vl_ref, c->vl_flags);
    MDF_LAUNCHED();
    return MDF_OK;
}

void vlist_free(mdf_p3m_ctx *c) {
    void *ptrs[] = {c->vl_list, c->vl_cnt, c->vl_ref, c->vl_flags};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    c->vl_list = nullptr;
    c->vl_cnt = nullptr;
    c->vl_ref = nullptr;
    c->vl_flags = nullptr;
    c->vl_skin = 0.0;
    c->vl_cap = 0;
}

int vlist_alloc(mdf_p3m_ctx *c, double skin) {
    vlist_free(c);
    if (!(skin > 0.0)) return MDF_OK;
    // entries a particle: its neighbours within rc + skin at the mean density
    // of n_max particles, with room for 40% more and 64, rounded to the list
    // chunk of 64 (the lists are padded to whole chunks)
    const double V = c->g.L[0] * c->g.L[1] * c->g.L[2];
    const double rl = c->g.rc + skin;
    const double mean = (double)c->prm.n_max / V * 4.18879020478639 * rl * rl * rl;
    int cap = (int)(1.4 * mean + 64.0);
    cap = (cap + 63) & ~63;
    if (cap > 65535) cap = 65535 & ~63;
    const size_t n = (size_t)c->prm.n_max;
    hipError_t e = hipSuccess;
    auto alloc = [&](void **p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, bytes ? bytes : 1);
    };
    alloc((void **)&c->vl_list, n * cap * sizeof(unsigned short));
    alloc((void **)&c->vl_cnt, n * sizeof(int));
    alloc((void **)&c->vl_ref, n * sizeof(double4));
    alloc((void **)&c->vl_flags, 8 * sizeof(int));
    if (e != hipSuccess) {
        set_error("pair lists: hipMalloc failed (%s)", hipGetErrorString(e));
        vlist_free(c);
        return MDF_ENOMEM;
    }
    MDF_TRY(fill_zero_i(c->vl_flags, 8, nullptr));
    c->vl_skin = skin;
    c->vl_cap = cap;
    c->vl_force_build = 1;
    return MDF_OK;
}

}  // namespace mdf
