// Integrator, thermostat, FP64 reductions and density modes.
//
// References:
//   kick / drift / wrap   velocity_verlet_p3m_ord6.py:5-26
//   velocity rescale      thermostat_p3m_ord6.py:5-25
//   K, E, T               yukawa_MD_p3m_ord6.py:199-201
//   n(q)                  yukawa_MD_p3m_ord6.py:214-219
//   U assembly            yukawa_p3m_ord6.py:21-24
// Arithmetic follows the reference expression by expression (v + 0.5*a*dt,
// x + v*dt, the strict > / < single wrap), so positions and velocities agree
// bit for bit given the same accelerations.  All reductions are FP64 over
// per-block partials in a fixed order (deterministic).
#include "mdf_internal.h"

namespace mdf {

__device__ __forceinline__ double block_sum_256(double v, double *red4) {
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red4[threadIdx.x >> 6] = v;
    __syncthreads();
    const double r = red4[0] + red4[1] + red4[2] + red4[3];
    __syncthreads();
    return r;
}

// Sum over a block of blockDim.x (<= 1024) threads; red holds 16 doubles.
__device__ __forceinline__ double block_sum(double v, double *red) {
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r += red[w];
    __syncthreads();
    return r;
}

struct WrapBox {
    double L[3], lo[3], hi[3];
};

struct KickDriftArgs {
    double *pos, *vel;
    const double *acc;
    int n3;
    double dt;
    int pbc;
    WrapBox wb;
    const double *scale;
};

__device__ __forceinline__ void kick_drift_body(const KickDriftArgs &A, const Blk &b) {
#pragma clang fp contract(off)
    double *__restrict__ pos = A.pos;
    double *__restrict__ vel = A.vel;
    const double *__restrict__ acc = A.acc;
    const double *__restrict__ scale = A.scale;
    const int n3 = A.n3, pbc = A.pbc;
    const double dt = A.dt;
    const WrapBox &wb = A.wb;
    const int i = b.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    const int d = i % 3;
    // separate roundings (no FMA contraction), as the reference's NumPy expressions
    double v = vel[i];
    if (scale) v = (v * *scale);
    v = v + (0.5 * acc[i]) * dt;
    double x = (pos[i] + (v * dt));
    if (pbc) {
        if (x > wb.hi[d]) x = x - wb.L[d];
        if (x < wb.lo[d]) x = x + wb.L[d];
    }
    vel[i] = v;
    pos[i] = x;
}

__global__ __launch_bounds__(256) void k_kick_drift(KickDriftArgs A) { kick_drift_body(A, blk_plain()); }
__global__ __launch_bounds__(256) void k_kick_drift_seg(const KickDriftArgs *__restrict__ tab, SegGrid S) {
    Blk b;
    const int r = seg_locate(S, b);
    kick_drift_body(tab[r], b);
}

__global__ __launch_bounds__(256) void k_kick(double *__restrict__ vel, const double *__restrict__ acc, int n3,
                                              double dt, double *__restrict__ part) {
#pragma clang fp contract(off)
    __shared__ double red[4];
    double s = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += gridDim.x * blockDim.x) {
        const double v = vel[i] + (0.5 * acc[i]) * dt;
        vel[i] = v;
        s += v * v;
    }
    s = block_sum_256(s, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_reduce(const double *__restrict__ part, int n, double *__restrict__ out) {
    __shared__ double red[4];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += part[i];
    s = block_sum_256(s, red);
    if (threadIdx.x == 0) *out = s;
}

// thermostat_p3m_ord6.py:9-23 with K = kf*sum v^2 and T = K/kf/N = sum v^2 / N
__device__ __forceinline__ double rescale_factor(double vsq, int n, double T_desired, int it, double *T_out) {
    const double T = vsq / (double)n;
    if (T_out) *T_out = T;
    return it <= 1999 ? sqrt(T_desired / T) : sqrt((20.0 * T_desired / T - 1.0) / 20.0);
}

__global__ __launch_bounds__(256) void k_rescale(double *__restrict__ vel, int n3, int n, const double *__restrict__ k_sum,
                                                 double T_desired, int it, double *__restrict__ T_out,
                                                 double *__restrict__ fact_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const double f = rescale_factor(*k_sum, n, T_desired, it, nullptr);
    if (i == 0) {
        rescale_factor(*k_sum, n, T_desired, it, T_out);
        if (fact_out) *fact_out = f;
    }
    if (i < n3) vel[i] = (vel[i] * f);
}

// Density modes: block (b, iq) sums cos and sin of q x_a over its particles.
__global__ __launch_bounds__(256) void k_modes(const double *__restrict__ pos, int n, const double *__restrict__ q,
                                               double *__restrict__ part) {
    __shared__ double red[4];
    const double qq = q[blockIdx.y];
    double c[3] = {0, 0, 0}, sn[3] = {0, 0, 0};
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            double sv, cv;
            sincos(-qq * pos[3 * j + a], &sv, &cv);
            c[a] += cv;
            sn[a] += sv;
        }
    }
    double *dst = part + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 6;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double cs = block_sum_256(c[a], red);
        const double ss = block_sum_256(sn[a], red);
        if (threadIdx.x == 0) {
            dst[2 * a] = cs;
            dst[2 * a + 1] = ss;
        }
    }
}

// Density modes on a grid q_m = m dq, m = 1..nq (nq <= 64).  Per particle
// and axis, one sincos gives e^{-i dq x} and the next modes of a chunk of
// kModeChunk follow by complex multiplication (rounding error grows like m
// ulps, ~1e-15 at m = 20); the first mode of each later chunk is a fresh
// sincos.  Threads keep the chunk's sums in registers; each wave reduces
// them once at the end and writes part[((m-1) * nwaves + wave) * 6 + c],
// the layout k_modes_final sums.
constexpr int kModeChunk = 24;

struct ModesArgs {
    const double *pos;
    int n;
    double dq;
    int nq;
    double *part;
};

// Sum over the wave with DPP row moves (no LDS): quad permutes, the half-row
// and row mirrors, then row_bcast:15 and row_bcast:31; the total ends in lane
// 63 (the other lanes hold partial sums).  All lanes must be active.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_add_d(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROW_MASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xF, false);
    return v + __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double wave_sum_last(double v) {
    v = dpp_add_d<0xB1, 0xF>(v);   // quad_perm [1, 0, 3, 2]
    v = dpp_add_d<0x4E, 0xF>(v);   // quad_perm [2, 3, 0, 1]
    v = dpp_add_d<0x141, 0xF>(v);  // row_half_mirror
    v = dpp_add_d<0x140, 0xF>(v);  // row_mirror
    v = dpp_add_d<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
    v = dpp_add_d<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

__device__ __forceinline__ void modes_grid_body(const ModesArgs &A, const Blk &b) {
    // b.y: axis, b.z: chunk of kModeChunk modes starting at m0
    const double *__restrict__ pos = A.pos;
    double *__restrict__ part = A.part;
    const int n = A.n, nq = A.nq;
    const double dq = A.dq;
    const int lane = threadIdx.x & 63;
    const int gw = (b.x * blockDim.x + threadIdx.x) >> 6;
    const int nw = (b.nx * blockDim.x) >> 6;
    const int stride = b.nx * blockDim.x;
    const int a = b.y;
    const int m0 = b.z * kModeChunk;
    double ar[kModeChunk], ai[kModeChunk];
#pragma unroll
    for (int k = 0; k < kModeChunk; ++k) ar[k] = ai[k] = 0.0;
    const double dqpi = dq * (1.0 / M_PI);  // angles in units of pi: sincospi reduces exactly
    for (int j = b.x * blockDim.x + threadIdx.x; j < n; j += stride) {
        const double th = dqpi * pos[3 * j + a];
        double s1, c1, zs, zc;
        sincospi(-th, &s1, &c1);
        if (m0 == 0) {
            zs = s1;
            zc = c1;
        } else {
            sincospi(-(double)(m0 + 1) * th, &zs, &zc);
        }
        ar[0] += zc;
        ai[0] += zs;
#pragma unroll
        for (int k = 1; k < kModeChunk; ++k) {
            const double nc = zc * c1 - zs * s1;
            const double ns = zc * s1 + zs * c1;
            zc = nc;
            zs = ns;
            ar[k] += zc;
            ai[k] += zs;
        }
    }
#pragma unroll
    for (int k = 0; k < kModeChunk; ++k) {
        const double sr = wave_sum_last(ar[k]);
        const double si = wave_sum_last(ai[k]);
        const int m = m0 + k;
        if (lane == 63 && m < nq) {
            part[((size_t)m * nw + gw) * 6 + 2 * a] = sr;
            part[((size_t)m * nw + gw) * 6 + 2 * a + 1] = si;
        }
    }
}

__global__ __launch_bounds__(256) void k_modes_grid(ModesArgs A) { modes_grid_body(A, blk_plain()); }
__global__ __launch_bounds__(256) void k_modes_grid_seg(const ModesArgs *__restrict__ tab, SegGrid S) {
    Blk b;
    const int r = seg_locate(S, b);
    modes_grid_body(tab[r], b);
}

// Sum the per-wave partials of k_modes_grid: block m, out[k * out_stride + m * 6 + c].
struct ModesFinalArgs {
    const double *part;
    int nw;
    double *out;
    int out_stride;
};

__device__ __forceinline__ void modes_final_body(const ModesFinalArgs &A, const Blk &b, int k, double *red) {
    const int m = b.x;
    double *out = A.out + (size_t)k * A.out_stride;
    for (int c = 0; c < 6; ++c) {
        double v = 0.0;
        for (int w = threadIdx.x; w < A.nw; w += blockDim.x) v += A.part[((size_t)m * A.nw + w) * 6 + c];
        v = block_sum_256(v, red);
        if (threadIdx.x == 0) out[(size_t)m * 6 + c] = v;
    }
}

__global__ __launch_bounds__(256) void k_modes_grid_final(ModesFinalArgs A) {
    __shared__ double red[4];
    modes_final_body(A, blk_plain(), 0, red);
}
__global__ __launch_bounds__(256) void k_modes_grid_final_seg(const ModesFinalArgs *__restrict__ tab, SegGrid S,
                                                              int k) {
    __shared__ double red[4];
    Blk b;
    const int r = seg_locate(S, b);
    modes_final_body(tab[r], b, k, red);
}

__global__ __launch_bounds__(64) void k_modes_final(const double *__restrict__ part, int nblk, int nq,
                                                    double *__restrict__ out) {
    const int iq = blockIdx.x;
    const int c = threadIdx.x;  // 0..5
    if (c >= 6) return;
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += part[((size_t)iq * nblk + b) * 6 + c];
    out[(size_t)iq * 6 + c] = s;  // (re, im) pairs for the three axes
}

// U = uf (U_pp + U_pm + U_self) with U_pm = 1/2 sum |rho_k|^2 G / V; optional
// K = kf sum v^2 and a per-step record (E, K, U, T).
struct FinalArgs {
    const double *pp;
    int npp;
    const double *pm;
    int npm;
    const double *kp;
    int nk;
    double half_inv_V, u_self, uf, kf;
    int n;
    double *scal;
    double *energies;  // 4 or null
    double *rec;       // 4 or null (segmented launches: step k writes rec + k * rec_stride)
    int rec_stride;
};

// Sum of x[0..n) over the block's threads, four independent chains a thread so
// that its loads are in flight together (the order is fixed: deterministic).
__device__ __forceinline__ double thread_sum4(const double *__restrict__ x, int n) {
    const int st = blockDim.x;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int i = threadIdx.x;
    for (; i + 3 * st < n; i += 4 * st) {
        s0 += x[i];
        s1 += x[i + st];
        s2 += x[i + 2 * st];
        s3 += x[i + 3 * st];
    }
    for (; i < n; i += st) s0 += x[i];
    return (s0 + s1) + (s2 + s3);
}

__device__ __forceinline__ void finalize_body(const FinalArgs &P, int step, double *red) {
    double a = thread_sum4(P.pp, P.npp);
    double b = thread_sum4(P.pm, P.npm);
    double k = thread_sum4(P.kp, P.nk);
    a = block_sum(a, red);
    b = block_sum(b, red);
    k = block_sum(k, red);
    if (threadIdx.x == 0) {
        const double Upp = a;
        const double Upm = b * P.half_inv_V;
        const double U = P.uf * (Upp + Upm + P.u_self);
        P.scal[kScalU] = U;
        P.scal[kScalUpp] = Upp;
        P.scal[kScalUpm] = Upm;
        P.scal[kScalUself] = P.u_self;
        if (P.energies) {
            P.energies[0] = U;
            P.energies[1] = Upp;
            P.energies[2] = Upm;
            P.energies[3] = P.u_self;
        }
        if (P.nk) {
            const double K = P.kf * k;
            P.scal[kScalK] = K;
            P.scal[kScalT] = K / P.kf / (double)P.n;
            if (P.rec) {
                double *rec = P.rec + (size_t)step * P.rec_stride;
                rec[0] = K + U;
                rec[1] = K;
                rec[2] = U;
                rec[3] = K / P.kf / (double)P.n;
            }
        }
    }
}

__global__ __launch_bounds__(1024) void k_finalize(FinalArgs P) {
    __shared__ double red[16];
    finalize_body(P, 0, red);
}
__global__ __launch_bounds__(1024) void k_finalize_seg(const FinalArgs *__restrict__ tab, SegGrid S, int k) {
    __shared__ double red[16];
    Blk b;
    const int r = seg_locate(S, b);
    finalize_body(tab[r], k, red);
}

int finalize_energy(mdf_p3m_ctx *c, int n, double uf, int have_pp, int have_pm, double *energies, hipStream_t s) {
    FinalArgs P{};
    P.pp = c->part_pp;
    P.npp = have_pp ? c->nblock : 0;
    P.pm = c->part_pm;
    P.npm = have_pm ? c->n_part_pm : 0;
    P.kp = nullptr;
    P.nk = 0;
    P.half_inv_V = 0.5 / (c->g.L[0] * c->g.L[1] * c->g.L[2]);
    P.u_self = (have_pp && have_pm) ? -(double)n * c->g.alpha / sqrt(M_PI) : 0.0;
    P.uf = uf;
    P.kf = 1.5;
    P.n = n;
    P.scal = c->scal;
    P.energies = energies;
    P.rec = nullptr;
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, s, P);
    MDF_LAUNCHED();
    return MDF_OK;
}

static FinalArgs finalize_step_args(mdf_p3m_ctx *c, int n, double uf, double kf, double *rec) {
    FinalArgs P{};
    P.pp = c->part_pp;
    P.npp = c->nblock;
    P.pm = c->part_pm;
    P.npm = c->n_part_pm;
    P.kp = c->part_red;
    P.nk = c->brick.nblock_gather;
    P.half_inv_V = 0.5 / (c->g.L[0] * c->g.L[1] * c->g.L[2]);
    P.u_self = -(double)n * c->g.alpha / sqrt(M_PI);
    P.uf = uf;
    P.kf = kf;
    P.n = n;
    P.scal = c->scal;
    P.energies = nullptr;
    P.rec = rec;
    P.rec_stride = 0;
    return P;
}

static int finalize_step(mdf_p3m_ctx *c, int n, double uf, double kf, double *rec, hipStream_t s) {
    const FinalArgs P = finalize_step_args(c, n, uf, kf, rec);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, s, P);
    MDF_LAUNCHED();
    return MDF_OK;
}

int kick_drift(mdf_p3m_ctx *c, double *pos, double *vel, const double *acc, int n, double dt, int pbc,
               const double lmin[3], const double lmax[3], const double *scale, hipStream_t s) {
    (void)c;
    if (n <= 0) return MDF_OK;
    WrapBox wb;
    for (int d = 0; d < 3; ++d) {
        wb.L[d] = c->g.L[d];
        wb.lo[d] = lmin ? lmin[d] : 0.0;
        wb.hi[d] = lmax ? lmax[d] : c->g.L[d];
    }
    KickDriftArgs A{pos, vel, acc, 3 * n, dt, pbc, wb, scale};
    hipLaunchKernelGGL(k_kick_drift, dim3((A.n3 + 255) / 256), dim3(256), 0, s, A);
    MDF_LAUNCHED();
    return MDF_OK;
}

int kick(mdf_p3m_ctx *c, double *vel, const double *acc, int n, double dt, double *k_sum, hipStream_t s) {
    if (n <= 0) return MDF_OK;
    const int n3 = 3 * n;
    int nb = (n3 + 255) / 256;
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(k_kick, dim3(nb), dim3(256), 0, s, vel, acc, n3, dt, c->part_red);
    MDF_LAUNCHED();
    if (k_sum) {
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, s, c->part_red, nb, k_sum);
        MDF_LAUNCHED();
    }
    return MDF_OK;
}

int density_modes(mdf_p3m_ctx *c, const double *pos, int n, const double *q, int nq, double *out, hipStream_t s) {
    if (nq <= 0) return MDF_OK;
    int nb = (n + 255) / 256;
    if (nb > 64) nb = 64;
    if (nb < 1) nb = 1;
    if ((size_t)nb * nq * 6 > (size_t)kMaxPartials) {
        set_error("density_modes: too many wave numbers (%d)", nq);
        return MDF_EINVAL;
    }
    hipLaunchKernelGGL(k_modes, dim3(nb, nq), dim3(256), 0, s, pos, n, q, c->part_red);
    MDF_LAUNCHED();
    hipLaunchKernelGGL(k_modes_final, dim3(nq), dim3(64), 0, s, c->part_red, nb, nq, out);
    MDF_LAUNCHED();
    return MDF_OK;
}

// Blocks per axis of k_modes_grid: about 16 particles per thread (at most 256
// blocks), so that the 2 x kModeChunk wave sums at the end of each wave cost
// little beside its particles' work.  With one particle in four threads
// (1024 blocks at N = 1e6) the sums took as long as the sincos and recurrences.
// A multiple of 8: workgroups are dealt to the 8 XCDs by their linear index
// x + nb (axis + 3 chunk), so with nb = 0 (mod 8) the three axes' blocks of
// one particle range run on one XCD and can share its L2 for the positions
// (each reads one coordinate of the same lines).
static int modes_blocks(int n) {
    int nb = (n + 4095) / 4096;
    if (nb > 256) nb = 256;
    nb = (nb + 7) & ~7;
    return nb;
}

int density_modes_grid(mdf_p3m_ctx *c, const double *pos, int n, double dq, int nq, double *out, hipStream_t s) {
    if (nq <= 0) return MDF_OK;
    if (nq > 64) {
        set_error("density modes on a q grid: nq = %d > 64", nq);
        return MDF_EINVAL;
    }
    const int nb = modes_blocks(n);
    const int nw = nb * 4;
    const int nchunk = (nq + kModeChunk - 1) / kModeChunk;
    const ModesArgs A{pos, n, dq, nq, c->part_red};
    hipLaunchKernelGGL(k_modes_grid, dim3(nb, 3, nchunk), dim3(256), 0, s, A);
    MDF_LAUNCHED();
    const ModesFinalArgs F{c->part_red, nw, out, 0};
    hipLaunchKernelGGL(k_modes_grid_final, dim3(nq), dim3(256), 0, s, F);
    MDF_LAUNCHED();
    return MDF_OK;
}

int thermostat_scale(mdf_p3m_ctx *c, double *vel, int n, const double *k_sum, double T_desired, int it,
                     double *t_out, hipStream_t s) {
    if (n <= 0) return MDF_OK;
    const int n3 = 3 * n;
    hipLaunchKernelGGL(k_rescale, dim3((n3 + 255) / 256), dim3(256), 0, s, vel, n3, n, k_sum, T_desired, it, t_out,
                       c->scal + kScalFact);
    MDF_LAUNCHED();
    return MDF_OK;
}

// One NVE step on device data (no host synchronisation).
int md_step(mdf_p3m_ctx *c, double *pos, double *vel, double *acc, const double *Z, int n, double dt, double af,
            double uf, double kf, double *rec, hipStream_t s) {
    prof_mark(c, kStageDrift, s);
    MDF_TRY(kick_drift(c, pos, vel, acc, n, dt, 1, nullptr, nullptr, nullptr, s));
    prof_mark(c, kStageBin, s);
    MDF_TRY(bin_step(c, pos, Z, n, s));
    prof_mark(c, kStagePP, s);
    MDF_TRY(pp_forces(c, n, s));
    if (c->mlp) MDF_TRY(mlp_correct(c->mlp, c, n, s));
    prof_mark(c, kStageAssign, s);
    MDF_TRY(pm_assign(c, n, nullptr, s));
    prof_mark(c, kStageFFT, s);
    MDF_TRY(pm_solve(c, s));
    GatherOut o{};
    o.acc = acc;
    o.af = af;
    o.include_pp = 1;
    o.vel = vel;
    o.kick_dt = dt;
    prof_mark(c, kStageGather, s);
    MDF_TRY(pm_gather(c, n, o, s));
    prof_mark(c, kStageFinal, s);
    MDF_TRY(finalize_step(c, n, uf, kf, rec, s));
    return MDF_OK;
}

// ---- segmented stages (replica batches) ----
static int build_drift(mdf_p3m_ctx *c, const StepIO &io, void *args, int4 &dims, size_t &lds) {
    WrapBox wb;
    for (int d = 0; d < 3; ++d) {
        wb.L[d] = c->g.L[d];
        wb.lo[d] = 0.0;
        wb.hi[d] = c->g.L[d];
    }
    *(KickDriftArgs *)args = KickDriftArgs{io.pos, io.vel, io.acc, 3 * io.n, io.dt, 1, wb, nullptr};
    const int nb = (3 * io.n + 255) / 256;
    dims = make_int4(nb > 0 ? nb : 1, 1, 1, 0);
    lds = 0;
    return MDF_OK;
}
static int launch_drift(const void *tab, const SegGrid &S, int total, size_t, int, hipStream_t s) {
    hipLaunchKernelGGL(k_kick_drift_seg, dim3(total), dim3(256), 0, s, (const KickDriftArgs *)tab, S);
    MDF_LAUNCHED();
    return MDF_OK;
}
static int build_final(mdf_p3m_ctx *c, const StepIO &io, void *args, int4 &dims, size_t &lds) {
    FinalArgs P = finalize_step_args(c, io.n, io.uf, io.kf, io.rec0);
    P.rec_stride = io.rec_stride;
    *(FinalArgs *)args = P;
    dims = make_int4(1, 1, 1, 0);
    lds = 0;
    return MDF_OK;
}
static int launch_final(const void *tab, const SegGrid &S, int total, size_t, int k, hipStream_t s) {
    hipLaunchKernelGGL(k_finalize_seg, dim3(total), dim3(1024), 0, s, (const FinalArgs *)tab, S, k);
    MDF_LAUNCHED();
    return MDF_OK;
}
static int build_modes(mdf_p3m_ctx *c, const StepIO &io, void *args, int4 &dims, size_t &lds) {
    if (io.nq > 64) {
        set_error("density modes on a q grid: nq = %d > 64", io.nq);
        return MDF_EINVAL;
    }
    *(ModesArgs *)args = ModesArgs{io.pos, io.n, 2.0 * M_PI / c->g.L[0], io.nq, c->part_red};
    dims = make_int4(modes_blocks(io.n), 3, (io.nq + kModeChunk - 1) / kModeChunk, 0);
    lds = 0;
    return MDF_OK;
}
static int launch_modes(const void *tab, const SegGrid &S, int total, size_t, int, hipStream_t s) {
    hipLaunchKernelGGL(k_modes_grid_seg, dim3(total), dim3(256), 0, s, (const ModesArgs *)tab, S);
    MDF_LAUNCHED();
    return MDF_OK;
}
static int build_modes_final(mdf_p3m_ctx *c, const StepIO &io, void *args, int4 &dims, size_t &lds) {
    *(ModesFinalArgs *)args = ModesFinalArgs{c->part_red, modes_blocks(io.n) * 4, io.modes0, io.modes_stride};
    dims = make_int4(io.nq, 1, 1, 0);
    lds = 0;
    return MDF_OK;
}
static int launch_modes_final(const void *tab, const SegGrid &S, int total, size_t, int k, hipStream_t s) {
    hipLaunchKernelGGL(k_modes_grid_final_seg, dim3(total), dim3(256), 0, s, (const ModesFinalArgs *)tab, S, k);
    MDF_LAUNCHED();
    return MDF_OK;
}

void seg_stages_md(std::vector<SegStage> &out, int part) {
    if (part == 0) out.push_back({"drift", sizeof(KickDriftArgs), build_drift, launch_drift});
    if (part == 1) out.push_back({"finalize", sizeof(FinalArgs), build_final, launch_final});
    if (part == 2) {
        out.push_back({"modes", sizeof(ModesArgs), build_modes, launch_modes});
        out.push_back({"modes_final", sizeof(ModesFinalArgs), build_modes_final, launch_modes_final});
    }
}

int seg_step_stages(int precision, std::vector<SegStage> &out) {
    out.clear();
    seg_stages_md(out, 0);
    seg_stages_vl(out, 0);
    seg_stages_bin(out);
    seg_stages_vl(out, 1);
    seg_stages_pp(out, precision);
    seg_stages_pm(out, precision, 0);
    seg_stages_fft(out, precision);
    seg_stages_pm(out, precision, 1);
    seg_stages_md(out, 1);
    seg_stages_md(out, 2);
    return (int)out.size();
}

}  // namespace mdf
