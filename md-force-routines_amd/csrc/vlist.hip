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
//   binning      skipped on a list step; on a build step its placing kernel
//                also writes the build positions (the new sorted copy)
// (replica batches run k_vl_refresh as a stage of its own for the last)
#include "mdf_internal.h"

namespace mdf {

constexpr int kPrepThreads = 256;
constexpr int kPrepPerThread = 4;  // particles a thread: the block maxima stay few

struct VlPrepArgs {
    const double *pos;
    int n;
    const int *perm;
    const double4 *vl_ref;
    double4 *spos;
    double L[3];
    const int *flags;
    float *bmax;
};

__device__ __forceinline__ void vl_prep_body(const VlPrepArgs &A, int bx, float *wmax) {
    if (A.flags[0]) return;  // a build is requested: the frozen order is replaced anyway
    const double Lx = A.L[0], Ly = A.L[1], Lz = A.L[2];
    float d2f = 0.0f;
    for (int q = 0; q < kPrepPerThread; ++q) {
        const int s = (bx * kPrepPerThread + q) * kPrepThreads + threadIdx.x;
        if (s >= A.n) break;
        const int i = A.perm[s];
        const double4 r = A.vl_ref[s];
        double dx = A.pos[3 * (size_t)i + 0] - r.x;
        double dy = A.pos[3 * (size_t)i + 1] - r.y;
        double dz = A.pos[3 * (size_t)i + 2] - r.z;
        dx -= Lx * rint(dx / Lx);
        dy -= Ly * rint(dy / Ly);
        dz -= Lz * rint(dz / Lz);
        A.spos[s] = make_double4(r.x + dx, r.y + dy, r.z + dz, r.w);
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
        A.bmax[bx] = m;
    }
}

__global__ __launch_bounds__(kPrepThreads) void k_vl_prep(VlPrepArgs A) {
    __shared__ float wmax[kPrepThreads / 64];
    vl_prep_body(A, blockIdx.x, wmax);
}
__global__ __launch_bounds__(kPrepThreads) void k_vl_prep_seg(const VlPrepArgs *__restrict__ tab, SegGrid S) {
    __shared__ float wmax[kPrepThreads / 64];
    Blk b;
    const int r = seg_locate(S, b);
    vl_prep_body(tab[r], b.x, wmax);
}

// The decision over the block maxima of k_vl_prep (nb of them; skipped when a
// build is already requested): one block of 1024 threads, its thread 0 writes.
struct VlDecideArgs {
    int *flags;
    const float *bmax;
    int nb;
    float thr;
};

__device__ __forceinline__ void vl_decide_body(const VlDecideArgs &A, float *wmax, int &forced) {
    int *flags = A.flags;
    if (threadIdx.x == 0) forced = flags[0];
    __syncthreads();
    float m = 0.0f;
    if (!forced)
        for (int b = threadIdx.x; b < A.nb; b += blockDim.x) m = fmaxf(m, A.bmax[b]);
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fmaxf(m, wmax[w]);
    m = fmaxf(m, wmax[0]);
    const bool build = forced != 0 || flags[1] != 0 || m > A.thr;
    flags[3] = build ? 1 : 2;
    flags[0] = 0;
    if (build) {
        flags[1] = 0;
        flags[4] += 1;
    } else {
        flags[5] += 1;
    }
}

__global__ __launch_bounds__(1024) void k_vl_decide(VlDecideArgs A) {
    __shared__ float wmax[16];
    __shared__ int forced;
    vl_decide_body(A, wmax, forced);
}
__global__ __launch_bounds__(1024) void k_vl_decide_seg(const VlDecideArgs *__restrict__ tab, SegGrid S) {
    __shared__ float wmax[16];
    __shared__ int forced;
    Blk b;
    const int r = seg_locate(S, b);
    vl_decide_body(tab[r], wmax, forced);
}

struct VlRefreshArgs {
    const double4 *spos;
    int n;
    double4 *vl_ref;
    const int *flags;
};

__device__ __forceinline__ void vl_refresh_body(const VlRefreshArgs &A, int bx) {
    if (A.flags[3] != 1) return;
    const int t = bx * blockDim.x + threadIdx.x;
    if (t < A.n) A.vl_ref[t] = A.spos[t];
}
__global__ __launch_bounds__(256) void k_vl_refresh(VlRefreshArgs A) { vl_refresh_body(A, blockIdx.x); }
__global__ __launch_bounds__(256) void k_vl_refresh_seg(const VlRefreshArgs *__restrict__ tab, SegGrid S) {
    Blk b;
    const int r = seg_locate(S, b);
    vl_refresh_body(tab[r], b.x);
}

__global__ void k_vl_set(int *__restrict__ p, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

static int prep_blocks(int n) { return (n + kPrepThreads * kPrepPerThread - 1) / (kPrepThreads * kPrepPerThread); }

// Build when some particle moved more than skin / 2 (the float threshold
// rounded down, so that an equal distance builds).
static float vl_threshold(const mdf_p3m_ctx *c) {
    const double h = 0.5 * c->vl_skin;
    float thr = (float)(h * h);
    if ((double)thr > h * h) thr = nextafterf(thr, 0.0f);
    return thr;
}

static VlPrepArgs prep_args(mdf_p3m_ctx *c, const double *pos, int n) {
    return VlPrepArgs{pos,       n, c->perm, c->vl_ref, c->spos, {c->g.L[0], c->g.L[1], c->g.L[2]}, c->vl_flags,
                      (float *)c->part_red};  // part_red is free until the n(q) stage of the step
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
    const int nb = prep_blocks(n);
    if (nb > kMaxPartials) {
        set_error("pair lists: %d particles exceed the displacement partials", n);
        return MDF_EINVAL;
    }
    const VlPrepArgs A = prep_args(c, pos, n);
    hipLaunchKernelGGL(k_vl_prep, dim3(nb), dim3(kPrepThreads), 0, s, A);
    MDF_LAUNCHED();
    hipLaunchKernelGGL(k_vl_decide, dim3(1), dim3(1024), 0, s, VlDecideArgs{c->vl_flags, A.bmax, nb, vl_threshold(c)});
    MDF_LAUNCHED();
    // a build's binning also records the build positions (the list steps'
    // reference), in its placing kernel: no launch of its own every step
    MDF_TRY(bin_particles_mode(c, pos, Z, n, c->vl_flags + 3, s, c->vl_ref));
    return MDF_OK;
}

int vlist_request_build(mdf_p3m_ctx *c, int n, hipStream_t s) {
    if (!(c->vl_skin > 0.0)) return MDF_OK;
    hipLaunchKernelGGL(k_vl_set, dim3(1), dim3(64), 0, s, c->vl_flags, 1);
    MDF_LAUNCHED();
    c->vl_force_build = 0;
    c->vl_n = n;
    return MDF_OK;
}

// ---- segmented stages (replica batches) ----
// A replica without pair lists has no blocks in these stages.
static bool seg_lists(const mdf_p3m_ctx *c, const StepIO &io) { return c->vl_skin > 0.0 && io.n > 0; }

static int build_vl_prep(mdf_p3m_ctx *c, const StepIO &io, void *args, int4 &dims, size_t &lds) {
    const bool on = seg_lists(c, io);
    const int nb = on ? prep_blocks(io.n) : 0;
    if (nb > kMaxPartials) {
        set_error("pair lists: %d particles exceed the displacement partials", io.n);
        return MDF_EINVAL;
    }
    *(VlPrepArgs *)args = prep_args(c, io.pos, io.n);
    dims = make_int4(nb, 1, 1, 0);
    lds = 0;
    return MDF_OK;
}
static int build_vl_decide(mdf_p3m_ctx *c, const StepIO &io, void *args, int4 &dims, size_t &lds) {
    const bool on = seg_lists(c, io);
    *(VlDecideArgs *)args = VlDecideArgs{c->vl_flags, (const float *)c->part_red, on ? prep_blocks(io.n) : 0,
                                         on ? vl_threshold(c) : 0.0f};
    dims = make_int4(on ? 1 : 0, 1, 1, 0);
    lds = 0;
    return MDF_OK;
}
static int build_vl_refresh(mdf_p3m_ctx *c, const StepIO &io, void *args, int4 &dims, size_t &lds) {
    const bool on = seg_lists(c, io);
    *(VlRefreshArgs *)args = VlRefreshArgs{c->spos, io.n, c->vl_ref, c->vl_flags};
    dims = make_int4(on ? (io.n + 255) / 256 : 0, 1, 1, 0);
    lds = 0;
    return MDF_OK;
}
#define MDF_SEG_LAUNCH(fname, kern, threads, Args)                                                        \
    static int fname(const void *tab, const SegGrid &S, int total, size_t, int, hipStream_t s) {        \
        if (total == 0) return MDF_OK; /* no replica with pair lists */                                   \
        hipLaunchKernelGGL(kern, dim3(total), dim3(threads), 0, s, (const Args *)tab, S);                   \
        MDF_LAUNCHED();                                                                                     \
        return MDF_OK;                                                                                      \
    }
MDF_SEG_LAUNCH(launch_vl_prep, k_vl_prep_seg, kPrepThreads, VlPrepArgs)
MDF_SEG_LAUNCH(launch_vl_decide, k_vl_decide_seg, 1024, VlDecideArgs)
MDF_SEG_LAUNCH(launch_vl_refresh, k_vl_refresh_seg, 256, VlRefreshArgs)
#undef MDF_SEG_LAUNCH

void seg_stages_vl(std::vector<SegStage> &out, int part) {
    if (part == 0) {
        out.push_back({"vl_prep", sizeof(VlPrepArgs), build_vl_prep, launch_vl_prep});
        out.push_back({"vl_decide", sizeof(VlDecideArgs), build_vl_decide, launch_vl_decide});
    } else {
        out.push_back({"vl_refresh", sizeof(VlRefreshArgs), build_vl_refresh, launch_vl_refresh});
    }
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
