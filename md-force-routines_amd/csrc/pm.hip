// Order-6 charge assignment (scatter) and field back-interpolation (gather).
//
// Reference: charge_assgn_6_numba, yukawa_pm_ord6.py:5-95, and
// field_assgn_6_numba, yukawa_pm_ord6.py:212-312.  Same base index
// ix = floor(x/h), same offset x/h - (ix + 1/2), same six weight polynomials
// (:29-48, evaluated in Horner form), same nodes ix..ix+5 wrapped once into
// the mesh (:52-93), same acc = Z * E_p (:308-310).
//
// Work is organised per PM block of CB^3 reference cells (CB chosen per
// context so the block's mesh brick fits LDS; the block's particles are
// contiguous runs of the sorted array).  Each particle gets a group of
// lanes: six for the scatter, each taking one z-plane (36 nodes) of
// its 6x6x6 stencil each; four for the gather, each taking a quarter (FP64:
// three z-planes by three y-rows of six x nodes; FP32: three z-planes by six
// y-rows of three x nodes):
// * scatter: the block's particles deposit into a private LDS brick of the
//   mesh with LDS atomics, then the brick is added to the global mesh with
//   FP64 atomics (one per non-empty brick node instead of 216 per particle);
// * gather: the block's brick of the three field meshes is staged in LDS
//   (when it fits), every particle reads its 216 nodes from there, and the
//   epilogue writes acc = af * (acc_pp + Z E_p) straight to the caller's
//   particle order, plus the optional mirrors of the reference's scratch
//   buffers and an optional fused half kick.
// Summation order differs from the serial reference, so agreement is to
// rounding (~1e-16 relative), not bitwise.
#include "mdf_internal.h"

namespace mdf {

// Weights of yukawa_pm_ord6.py:29-48 in Horner form, in arithmetic T.  The
// six polynomials pair up as w[5-k](x) = w[k](-x) (the coefficients of each
// pair differ in the sign of the odd powers), so w0, w1, w2 give all six.
//
// Normalisation: FP64 divides each weight by 3840 (w0, w1) or 1920 (w2) as
// the reference does.  FP32 evaluates the numerators 3840 w, whose
// coefficients are exact, and applies 1/3840^3 once in FP64, to each deposit
// and to each gathered field (wscale).  The FP32 constant 1/3840 lies 5.2e-8
// above 1/3840; taken once per axis it raised every deposit and every
// gathered field by 1.6e-7 (sum of rho = N (1 + 1.6e-7)).
template <class T>
struct WNorm {
    static constexpr T c01 = (T)(1.0 / 3840.0), c2 = (T)(1.0 / 1920.0);
};
template <>
struct WNorm<float> {
    static constexpr float c01 = 1.0f, c2 = 2.0f;
};
template <class T>
__device__ __forceinline__ double wscale() {
    return sizeof(T) == 4 ? 1.0 / (3840.0 * 3840.0 * 3840.0) : 1.0;
}
template <class T>
__device__ __forceinline__ T w6_0(T x) {
    return ((T)1 + x * ((T)-10 + x * ((T)40 + x * ((T)-80 + x * ((T)80 + x * (T)-32))))) * WNorm<T>::c01;
}
template <class T>
__device__ __forceinline__ T w6_1(T x) {
    return ((T)237 + x * ((T)-750 + x * ((T)840 + x * ((T)-240 + x * ((T)-240 + x * (T)160))))) * WNorm<T>::c01;
}
template <class T>
__device__ __forceinline__ T w6_2(T x) {
    return ((T)841 + x * ((T)-770 + x * ((T)-440 + x * ((T)560 + x * ((T)80 + x * (T)-160))))) * WNorm<T>::c2;
}
// w0, w1 or w2 at x by k (0, 1, 2) with the coefficients selected, for a k
// that differs between lanes: one polynomial instead of three
template <class T>
__device__ __forceinline__ T w6_sel(int k, T x) {
    auto c = [k](T a, T b, T d) { return k == 0 ? a : (k == 1 ? b : d); };
    const T p = c((T)1, (T)237, (T)841) +
                x * (c((T)-10, (T)-750, (T)-770) +
                     x * (c((T)40, (T)840, (T)-440) +
                          x * (c((T)-80, (T)-240, (T)560) + x * (c((T)80, (T)-240, (T)80) + x * c((T)-32, (T)160, (T)-160)))));
    return p * (k == 2 ? WNorm<T>::c2 : WNorm<T>::c01);
}
template <class T>
__device__ __forceinline__ void weights6(T x, T w[6]) {
    w[0] = w6_0<T>(x);
    w[1] = w6_1<T>(x);
    w[2] = w6_2<T>(x);
    w[3] = w6_2<T>(-x);
    w[4] = w6_1<T>(-x);
    w[5] = w6_0<T>(-x);
}

// Base index floor(x/h) and the offset x/h - (ix + 1/2) (yukawa_pm_ord6.py:17-27
// takes x/h once and subtracts the same way); the index and offset are formed
// in FP64 from the FP64 position, with x/h as x * (1/h): one multiply instead
// of an FP64 division.  The two differ by an ulp of x/h; at a node boundary
// that can move ix by one with the offset at +-1/2, where the order-6 weights
// of the outermost node are zero, so the deposit agrees to rounding either way.
template <class T>
__device__ __forceinline__ int base_and_offset(double x, double ih, T &off) {
    const double s = x * ih;
    const int i0 = (int)floor(s);
    off = (T)(s - (i0 + 0.5));
    return i0;
}

// Brick node index t = x + B0 (y + B1 z) split into (x, y, z) by two 32-bit
// multiply-highs: m = ceil(2^32 / B) gives floor(t / B) exactly while t B <
// 2^32 (bricks hold < 2^20 nodes of extent < 2^8); the integer divisions by
// runtime extents cost ~15 instructions each.
struct Split3 {
    unsigned m0, m1;
    int B0, B1;
};
__device__ __forceinline__ Split3 split3_init(int B0, int B1) {
    return Split3{0xffffffffu / (unsigned)B0 + 1u, 0xffffffffu / (unsigned)B1 + 1u, B0, B1};
}
__device__ __forceinline__ void split3(const Split3 &d, int t, int &x, int &y, int &z) {
    const int q = (int)__umulhi((unsigned)t, d.m0);
    x = t - q * d.B0;
    z = (int)__umulhi((unsigned)q, d.m1);
    y = q - z * d.B1;
}

// Wrap for indices in [-M, 3M) (brick nodes and stencils, since M >= 8 and a
// brick spans at most M/3 + 9 nodes).
__device__ __forceinline__ int wrap3(int i, int M) {
    if (i < 0) i += M;
    if (i >= M) i -= M;
    if (i >= M) i -= M;
    return i;
}

// FP64 gather: a particle's stencil is split over four consecutive lanes: lane
// bit 0 picks the y rows (0, 1, 2 or 5, 4, 3), bit 1 the z planes (0, 2, 4 or
// 5, 3, 1).  By the mirror identity above the weights of a lane's rows are
// w0, w1, w2 at +-y and of its planes w0, w2 at +-z and w1 at -+z: every lane
// evaluates the same polynomials, with no lane-dependent selects.
constexpr int kGatherLanes = 4;   // lanes a particle in the gather
constexpr int kScatterLanes = 6;  // in the assignment (one z-plane each)

template <class T>
struct Quarter {
    int ix, iy, iz;     // base node per axis
    T wx[6], wy[3], wz[3];
    int yo[3], zo[3];   // this lane's rows and planes within the stencil
};

template <class T>
__device__ __forceinline__ void quarter(const Geom &g, const double4 &p, int sub, Quarter<T> &q) {
    T fx, fy, fz;
    q.ix = base_and_offset<T>(p.x, g.ih[0], fx);
    q.iy = base_and_offset<T>(p.y, g.ih[1], fy);
    q.iz = base_and_offset<T>(p.z, g.ih[2], fz);
    const int yh = sub & 1, zh = sub >> 1;
    weights6<T>(fx, q.wx);
    const T sy = yh ? -fy : fy, sz = zh ? -fz : fz;
    q.wy[0] = w6_0<T>(sy);
    q.wy[1] = w6_1<T>(sy);
    q.wy[2] = w6_2<T>(sy);
    q.wz[0] = w6_0<T>(sz);
    q.wz[1] = w6_2<T>(sz);
    q.wz[2] = w6_1<T>(-sz);
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        q.yo[b] = yh ? 5 - b : b;
        q.zo[b] = zh ? 5 - 2 * b : 2 * b;
    }
}

// Sum over each group of four lanes (all lanes active), by DPP quad permutes.
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <class T>
__device__ __forceinline__ T quad_sum(T v) {
    v += dpp_mov<0xB1>(v);  // quad_perm [1, 0, 3, 2]
    v += dpp_mov<0x4E>(v);  // quad_perm [2, 3, 0, 1]
    return v;
}

// A PM block is at most sb[0] x sb[1] x sb[2] sub-cells (the blocks of an
// axis share its sub-cells evenly): whole reference cells (sb = 2 cb), or for
// the gather, whose FP64 brick of one cell can exceed the LDS, half a cell
// along x (and y).  Its
// particles are sb[1] sb[2] (or fewer) contiguous runs of the sorted array:
// one per row of sub-cells along y and z, spanning the block in x.
struct PMBlock {
    int s0[3], ns[3];  // first sub-cell and sub-cell count per axis
    int o[3];          // brick origin (mesh node) per axis
};

__device__ __forceinline__ PMBlock pm_block(const Geom &g, const int *sb, int bid) {
    PMBlock b;
    const int nbx = (g.nsub[0] + sb[0] - 1) / sb[0], nby = (g.nsub[1] + sb[1] - 1) / sb[1];
    const int id[3] = {bid % nbx, (bid / nbx) % nby, bid / (nbx * nby)};
    const int nbz = (g.nsub[2] + sb[2] - 1) / sb[2];
    const int nbd[3] = {nbx, nby, nbz};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        // the nb = ceil(n / sb) blocks of an axis split its n sub-cells as
        // evenly as whole sub-cells allow (at most sb each): a trailing block
        // of the remainder alone made the workgroups' work uneven
        b.s0[d] = id[d] * g.nsub[d] / nbd[d];
        b.ns[d] = (id[d] + 1) * g.nsub[d] / nbd[d] - b.s0[d];
        // the lowest base index of the block's particles; one whose x/h
        // rounds below it (within an ulp of the block face) is outside the
        // brick and takes the kernels' global-memory path
        b.o[d] = (int)floor(b.s0[d] * g.w[d] / g.h[d]);
    }
    return b;
}

// Fill the run table (start in the sorted array, prefix offsets) in LDS;
// returns the number of runs.  Call with all threads; ends with a barrier.
__device__ int pm_runs(const Geom &g, const int *__restrict__ cell_start, const PMBlock &b, int *run_start,
                       int *run_off) {
    const int ny = b.ns[1], nz = b.ns[2];
    const int nrun = ny * nz;
    const int t = threadIdx.x;
    int len = 0;
    if (t < nrun) {
        const int sy = b.s0[1] + t % ny, sz = b.s0[2] + t / ny;
        const int cfirst = b.s0[0] + g.nsub[0] * (sy + g.nsub[1] * sz);
        const int s0 = cell_start[cfirst];
        run_start[t] = s0;
        len = cell_start[cfirst + b.ns[0]] - s0;
    }
    if (t < 64) {  // nrun <= 64: one wave scans the lengths
        int incl = len;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int u = __shfl_up(incl, o, 64);
            if (t >= o) incl += u;
        }
        run_off[t + 1] = incl;
        if (t == 0) run_off[0] = 0;
    }
    __syncthreads();
    return nrun;
}

// Sorted index of the block's k-th particle (binary search of the runs).
__device__ __forceinline__ int pm_particle(const int *run_start, const int *run_off, int nrun, int k) {
    int lo = 0, hi = nrun;  // run_off[lo] <= k < run_off[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (run_off[mid] <= k) lo = mid;
        else hi = mid;
    }
    return run_start[lo] + (k - run_off[lo]);
}

// The same for a k that never decreases between calls of one thread: the run
// index advances from the last call's (empty runs are stepped over).
__device__ __forceinline__ int pm_particle_seq(const int *run_start, const int *run_off, int nrun, int k, int &run) {
    while (run + 1 < nrun && run_off[run + 1] <= k) ++run;
    return run_start[run] + (k - run_off[run]);
}

struct AssignArgs {
    Geom g;
    Brick br;
    const double4 *__restrict__ spos;
    const int *__restrict__ cell_start;
    double *__restrict__ rho;
    int *__restrict__ stats;
};

template <bool LDS, class T>
__device__ __forceinline__ void assign_body(const AssignArgs &P, const Blk &bk, unsigned char *braw, int *run_start,
                                            int *run_off) {
    // FP64: the brick in FP64 (ds_add_f64).  FP32: the brick in 32-bit fixed
    // point, each deposit rounded to a multiple of 1/S (ds_add_u32: integer LDS
    // atomics run ~3x the rate of FP64 ones and ~25x that of FP32 ones here);
    // S = 2^k is the largest power of two for which no node of the block can
    // overflow (a deposit is at most 0.167 |Z|, so a node holds at most 0.167
    // sum|Z| over the block's particles), capped at 2^24.  Each lane carries
    // the rounding remainder of its deposits into the next one (error
    // diffusion), so its 36 deposits sum to its share of Z to within half a
    // unit; the carry starts at a hashed value in [-1/2, 1/2), which makes the
    // lane's total exactly unbiased (plain rounding drops the many corner
    // deposits below half a unit and biased the charge low by 2.5e-7 of sum Z
    // at N = 8e6; a zero start carries the trailing small deposits out of the
    // sum, 1e-9 low).  The flush adds the brick to the FP64 mesh with FP64
    // atomics.
    constexpr bool kFix = LDS && sizeof(T) == 4;
    double *brick = (double *)braw;
    int *ibrick = (int *)braw;
    const Geom &g = P.g;
    const int B0 = LDS ? P.br.B[0] : 0, B1 = LDS ? P.br.B[1] : 0, B2 = LDS ? P.br.B[2] : 0;
    const int nb = B0 * B1 * B2;
    const Split3 sp3 = split3_init(max(B0, 1), max(B1, 1));
    const PMBlock blk = pm_block(g, P.br.sb_assign, bk.x);
    const int o0 = blk.o[0], o1 = blk.o[1], o2 = blk.o[2];
    for (int t = threadIdx.x; t < nb; t += blockDim.x) {
        if (kFix) ibrick[t] = 0;
        else brick[t] = 0.0;
    }
    const int nrun = pm_runs(g, P.cell_start, blk, run_start, run_off);
    const int npart = run_off[nrun];
    float kq = 0.0f;     // S / 3840^3: FP32 weight numerators to fixed-point units
    double inv_s = 0.0;  // charge of one unit (1 / S up to kq's rounding)
    if constexpr (kFix) {
        __shared__ float zsum_w[kPMThreads / 64];
        float zs = 0.0f;
        int run = 0;
        for (int k = threadIdx.x; k < npart; k += blockDim.x)
            zs += fabsf((float)P.spos[pm_particle_seq(run_start, run_off, nrun, k, run)].w);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) zs += __shfl_xor(zs, o, 64);
        if ((threadIdx.x & 63) == 0) zsum_w[threadIdx.x >> 6] = zs;
        __syncthreads();
        float zt = 0.0f;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) zt += zsum_w[w];
        // largest k <= 24 with 0.17 zt 2^k <= 2^30 (margin for the rounding of each deposit)
        int e = 24;
        const float cap = 0.17f * zt * 1.01f + 1e-30f;
        while (e > 0 && cap * ldexpf(1.0f, e) > 1073741824.0f) --e;
        kq = (float)(ldexp(1.0, e) * wscale<T>());
        // a unit is wscale / kq of charge: kq's FP32 rounding (up to 6e-8)
        // stays out of the charge sum
        inv_s = wscale<T>() / (double)kq;
    }
    // six lanes a particle, one stencil z-plane each, ten particles a wave
    // (lanes 60-63 idle).  Eight lanes a particle (two idle) measured ~6%
    // faster than four when the deposits took six instructions, from fewer
    // LDS atomics of one instruction on one node; with four-instruction
    // deposits the busy lanes count for more
    const int lane = threadIdx.x & 63;
    const int sub = lane % kScatterLanes, pslot = lane / kScatterLanes;
    int run = 0;
    // each particle's position is loaded one trip ahead, so its memory
    // latency overlaps the deposits of the particle before
    // (unconditionally: a trip past the end reads the block's last particle)
    constexpr int kPerWave = 64 / kScatterLanes;
    const int kstep = (int)(blockDim.x >> 6) * kPerWave;
    int k = (int)(threadIdx.x >> 6) * kPerWave + min(pslot, kPerWave - 1);
    int si_n = 0;
    double4 p_n = make_double4(0.0, 0.0, 0.0, 0.0);
    if (k < npart) {
        si_n = pm_particle_seq(run_start, run_off, nrun, k, run);
        p_n = P.spos[si_n];
    }
    for (; k < npart; k += kstep) {
        const int si = si_n;
        const double4 p = p_n;
        si_n = pm_particle_seq(run_start, run_off, nrun, min(k + kstep, npart - 1), run);
        p_n = P.spos[si_n];
        if (pslot >= kPerWave) continue;
        T fx, fy, fz, wx[6], wy[6];
        const int ix = base_and_offset<T>(p.x, g.ih[0], fx);
        const int iy = base_and_offset<T>(p.y, g.ih[1], fy);
        const int iz = base_and_offset<T>(p.z, g.ih[2], fz);
        weights6<T>(fx, wx);
        weights6<T>(fy, wy);
        const int rx = ix - o0, ry = iy - o1, rz = iz - o2;
        const bool inside = rx >= 0 && ry >= 0 && rz >= 0 && rx + 5 < B0 && ry + 5 < B1 && rz + 5 < B2;
        // this lane's plane weight w_sub(z) = w_{5-sub}(-z): one polynomial
        // with lane-selected coefficients
        const T zw = (T)p.w * w6_sel<T>(sub < 3 ? sub : 5 - sub, sub < 3 ? fz : -fz);
        if (kFix && inside) {
            T wxk[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) wxk[c] = wx[c] * kq;
            unsigned h = (unsigned)si * 2654435761u + (unsigned)sub * 40503u;
            h = (h ^ (h >> 15)) * 2246822519u;
            float carry = (float)(h >> 8) * 5.9604644775390625e-8f - 0.5f;  // 2^-24
#pragma unroll
            for (int b = 0; b < 6; ++b) {
                const T zwy = zw * wy[b];
                int *row = ibrick + ((rz + sub) * B1 + (ry + b)) * B0 + rx;
#pragma unroll
                for (int c = 0; c < 6; ++c) {
                    const float y = fmaf(zwy, wxk[c], carry);
                    const float f = rintf(y);  // nearest (ties to even): one instruction
                    carry = y - f;
                    atomicAdd(row + c, (int)f);
                }
            }
        } else if (LDS && inside) {
#pragma unroll
            for (int b = 0; b < 6; ++b) {
                const T zwy = zw * wy[b];
                double *row = brick + ((rz + sub) * B1 + (ry + b)) * B0 + rx;
#pragma unroll
                for (int c = 0; c < 6; ++c) atomicAdd(row + c, zwy * wx[c]);
            }
        } else {
            if (LDS && sub == 0) atomicAdd(&P.stats[1], 1);
            const int gz = wrap3(iz + sub, g.M[2]);
            for (int b = 0; b < 6; ++b) {
                const int gy = wrap3(iy + b, g.M[1]);
                const T zwy = zw * wy[b];
                for (int c = 0; c < 6; ++c) {
                    const int gx = wrap3(ix + c, g.M[0]);
                    unsafeAtomicAdd(&P.rho[((size_t)gz * g.M[1] + gy) * g.M[0] + gx],
                                    (double)(zwy * wx[c]) * wscale<T>());
                }
            }
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nb; t += blockDim.x) {
        const double v = kFix ? (double)ibrick[t] * inv_s : brick[t] * wscale<T>();
        if (v != 0.0) {
            int tx, ty, tz;
            split3(sp3, t, tx, ty, tz);
            const int gx = wrap3(o0 + tx, g.M[0]);
            const int gy = wrap3(o1 + ty, g.M[1]);
            const int gz = wrap3(o2 + tz, g.M[2]);
            unsafeAtomicAdd(&P.rho[((size_t)gz * g.M[1] + gy) * g.M[0] + gx], v);
        }
    }
}

template <bool LDS, class T>
__global__ __launch_bounds__(kPMThreads) void k_assign(AssignArgs P) {
    extern __shared__ __align__(16) unsigned char braw[];
    __shared__ int run_start[64], run_off[65];
    assign_body<LDS, T>(P, blk_plain(), braw, run_start, run_off);
}
template <class T>
__global__ __launch_bounds__(kPMThreads) void k_assign_seg(const AssignArgs *__restrict__ tab, SegGrid S) {
    extern __shared__ __align__(16) unsigned char braw[];
    __shared__ int run_start[64], run_off[65];
    Blk b;
    const int r = seg_locate(S, b);
    if (tab[r].br.lds_assign) assign_body<true, T>(tab[r], b, braw, run_start, run_off);
    else assign_body<false, T>(tab[r], b, braw, run_start, run_off);
}

static AssignArgs assign_args(mdf_p3m_ctx *c) {
    AssignArgs P;
    P.g = c->g;
    P.br = c->brick;
    P.spos = c->spos;
    P.cell_start = c->cell_start;
    P.rho = c->rho;
    P.stats = c->stats;
    return P;
}

int pm_assign(mdf_p3m_ctx *c, int n, double *rho_out, hipStream_t s) {
    (void)n;
    MDF_TRY(fill_zero(c->rho, c->nmesh, s));
    const AssignArgs P = assign_args(c);
    const bool mixed = c->prm.precision == MDF_PREC_MIXED;
    const dim3 grid(c->brick.nblock_assign);
    if (c->brick.lds_assign) {
        // FP32: the fixed-point brick, 4 B a node (FP64: 8)
        const size_t lds = (mixed ? 4 : 8) * (size_t)c->brick.B[0] * c->brick.B[1] * c->brick.B[2];
        if (mixed) {
            MDF_TRY(allow_lds(k_assign<true, float>, lds));
            hipLaunchKernelGGL((k_assign<true, float>), grid, dim3(kPMThreads), lds, s, P);
        } else {
            MDF_TRY(allow_lds(k_assign<true, double>, lds));
            hipLaunchKernelGGL((k_assign<true, double>), grid, dim3(kPMThreads), lds, s, P);
        }
    } else if (mixed) {
        hipLaunchKernelGGL((k_assign<false, float>), grid, dim3(kPMThreads), 0, s, P);
    } else {
        hipLaunchKernelGGL((k_assign<false, double>), grid, dim3(kPMThreads), 0, s, P);
    }
    MDF_LAUNCHED();
    if (rho_out) MDF_HIP(hipMemcpyAsync(rho_out, c->rho, c->nmesh * sizeof(double), hipMemcpyDeviceToDevice, s));
    return MDF_OK;
}

// ------------------------------------------------------------------ gather --

struct GatherArgs {
    Geom g;
    Brick br;
    const double4 *__restrict__ spos;
    const int *__restrict__ cell_start;
    const int *__restrict__ perm;
    const void *__restrict__ emesh;     // interleaved (Ex, Ey, Ez), element type T of k_gather
    const double *__restrict__ acc_pp_s;
    double *__restrict__ part;          // K partials when kicking
    GatherOut o;
};

// The gather's epilogue, split over the lanes of a particle: lane c < 3 of its
// four (each holding the particle's summed field) writes component c of E_p and
// of what the outputs ask for, in the caller's order i, with the fused second
// half-kick; pp is component c of the particle's PP acceleration (read when
// the outputs ask for it).  Returns that component's v^2 when kicking (else 0).
template <class T>
__device__ __forceinline__ double gather_out_c(const GatherArgs &P, int c, int i, double zq, double pp, T ex, T ey,
                                               T ez) {
    const T e = c == 0 ? ex : (c == 1 ? ey : ez);
    const double ed = (double)e * wscale<T>();
    const double ap = zq * ed;
    const GatherOut &o = P.o;
    double *eo = c == 0 ? o.ex : (c == 1 ? o.ey : o.ez);
    if (eo) eo[i] = ed;
    const size_t ic = 3 * (size_t)i + c;
    if (o.acc_pm) o.acc_pm[ic] = ap;
    if (o.acc_pp) o.acc_pp[ic] = pp;
    double k2 = 0.0;
    if (o.acc) {
        const double a = o.include_pp ? (o.af * (pp + ap)) : ap;
        o.acc[ic] = a;
        if (o.vel) {
#pragma clang fp contract(off)
            const double v = o.vel[ic] + (0.5 * a) * o.kick_dt;
            o.vel[ic] = v;
            k2 = v * v;
        }
    }
    return k2;
}

template <bool LDS, class T>
__device__ __forceinline__ void gather_body(const GatherArgs &P, const Blk &bk, unsigned char *eraw, double *red,
                                            int *run_start, int *run_off) {
    T *ebrick = (T *)eraw;
    const T *emesh = (const T *)P.emesh;
    const Geom &g = P.g;
    const int B0 = P.br.Bg[0], B1 = P.br.Bg[1], B2 = P.br.Bg[2];
    const PMBlock blk = pm_block(g, P.br.sb_gather, bk.x);
    const int o0 = blk.o[0], o1 = blk.o[1], o2 = blk.o[2];
    if (LDS) {
        const int nb = B0 * B1 * B2;
        const Split3 sp3 = split3_init(B0, B1);
        for (int t = threadIdx.x; t < nb; t += blockDim.x) {
            int tx, ty, tz;
            split3(sp3, t, tx, ty, tz);
            const size_t node = ((size_t)wrap3(o2 + tz, g.M[2]) * g.M[1] + wrap3(o1 + ty, g.M[1])) * g.M[0] +
                                wrap3(o0 + tx, g.M[0]);
            ebrick[3 * t + 0] = emesh[3 * node + 0];
            ebrick[3 * t + 1] = emesh[3 * node + 1];
            ebrick[3 * t + 2] = emesh[3 * node + 2];
        }
    }
    const int nrun = pm_runs(g, P.cell_start, blk, run_start, run_off);
    const int npart = run_off[nrun];
    double ksum = 0.0;
    const int sub = threadIdx.x % kGatherLanes;  // stencil quarter of this lane
    const int per_pass = blockDim.x / kGatherLanes;
    const int npass = (npart + per_pass - 1) / per_pass;
    for (int pass = 0; pass < npass; ++pass) {
        const int k = pass * per_pass + threadIdx.x / kGatherLanes;
        const bool live = k < npart;
        const int s = live ? pm_particle(run_start, run_off, nrun, k) : 0;
        T ex = 0, ey = 0, ez = 0;
        double4 p = make_double4(0.0, 0.0, 0.0, 0.0);
        if (live) {
            p = P.spos[s];
            Quarter<T> q;
            quarter<T>(g, p, sub, q);
            const int rx = q.ix - o0, ry = q.iy - o1, rz = q.iz - o2;
            const bool inside = rx >= 0 && ry >= 0 && rz >= 0 && rx + 5 < B0 && ry + 5 < B1 && rz + 5 < B2;
            if (LDS && inside) {
#pragma unroll
                for (int m = 0; m < 3; ++m) {
#pragma unroll
                    for (int b = 0; b < 3; ++b) {
                        // FP64: loads are not hoisted across rows, which keeps them
                        // within the 128-VGPR limit of the 1024-thread workgroup
                        if constexpr (sizeof(T) == 8) __builtin_amdgcn_sched_barrier(0);
                        const T wzy = q.wz[m] * q.wy[b];
                        const T *row = ebrick + 3 * (((rz + q.zo[m]) * B1 + (ry + q.yo[b])) * B0 + rx);
#pragma unroll
                        for (int c = 0; c < 6; ++c) {
                            const T ww = wzy * q.wx[c];
                            ex += row[3 * c + 0] * ww;
                            ey += row[3 * c + 1] * ww;
                            ez += row[3 * c + 2] * ww;
                        }
                    }
                }
            } else {
                for (int m = 0; m < 3; ++m) {
                    const int gz = wrap3(q.iz + q.zo[m], g.M[2]);
                    for (int b = 0; b < 3; ++b) {
                        const int gy = wrap3(q.iy + q.yo[b], g.M[1]);
                        const T wzy = q.wz[m] * q.wy[b];
                        const T *row = emesh + 3 * (((size_t)gz * g.M[1] + gy) * g.M[0]);
#pragma unroll
                        for (int c = 0; c < 6; ++c) {
                            const int gx = wrap3(q.ix + c, g.M[0]);
                            const T ww = wzy * q.wx[c];
                            ex += row[3 * gx + 0] * ww;
                            ey += row[3 * gx + 1] * ww;
                            ez += row[3 * gx + 2] * ww;
                        }
                    }
                }
            }
        }
        // sum the four quarters of each particle
        ex = quad_sum(ex);
        ey = quad_sum(ey);
        ez = quad_sum(ez);
        if (!live || sub > 2) continue;
        const double pp = (P.o.include_pp || P.o.acc_pp) ? P.acc_pp_s[3 * (size_t)s + sub] : 0.0;
        ksum += gather_out_c<T>(P, sub, P.perm[s], p.w, pp, ex, ey, ez);
    }
    if (P.o.vel) {
        ksum = wave_sum(ksum);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ksum;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int w = 0; w < kGatherThreads / 64; ++w) t += red[w];
            P.part[bk.x] = t;
        }
    }
}

// FP32 gather from the LDS field brick (layout at gather_plane_f32).  Four
// lanes a particle: lane bit 0 takes the x nodes 0-2 or 3-5 of every row, bit 1
// the z planes 0, 2, 4 or 5, 3, 1; each lane covers all six y rows of its
// three planes.  By the mirror identity w[5-k](x) = w[k](-x) every lane
// evaluates the same polynomials at +-x, y, -y and +-z.  The lanes of a
// particle read at its base address plus 0, 9, 16 and 25 banks (mod 32), so a
// half-wave whose 8 particles have distinct bases mod 16 of one parity reads
// 32 distinct banks at each instruction.  The visit order that arranges this
// buckets the block's particles by (3 rx + Sy ry) mod 16 (the base mod 16; Sz
// is 16 mod 32) and deals them out one bucket of each parity class at a time:
// position of rank g in bucket j of class c = (sum over the class of
// min(size, g)) + (buckets j' < j of the class with size > g), after the even
// class.  Measured at N = 1e6 (128^3, 3^3-cell blocks): 0.151 -> 0.136 ms with
// a plane stride of 0 (mod 32) and a 64-bank ordering (mod 32 keys).
// LDS false: no brick; every particle reads its nodes from the field mesh in
// global memory (through the caches), for fine meshes where a block's brick
// holds many nodes per particle.
template <bool LDS>
__device__ __forceinline__ void gather_body_f32(const GatherArgs &P, const Blk &bk, unsigned char *eraw, double *red,
                                                int *run_start, int *run_off, int *bucket) {
    float *eb = (float *)eraw;
    const float *emesh = (const float *)P.emesh;
    const Geom &g = P.g;
    const int B0 = P.br.Bg[0], B1 = P.br.Bg[1], B2 = P.br.Bg[2];
    const int Sy = 3 * B0, Sz = gather_plane_f32(B0, B1);
    const PMBlock blk = pm_block(g, P.br.sb_gather, bk.x);
    const int o0 = blk.o[0], o1 = blk.o[1], o2 = blk.o[2];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int nb = LDS ? B0 * B1 * B2 : 0;
    const Split3 sp3 = split3_init(max(B0, 1), max(B1, 1));
    // stage the brick up to twelve nodes a thread at a time, the loads
    // unconditional (indices past the brick clamped to its last node) so that
    // all twelve are in flight together: one node a trip waited ~12 memory
    // latencies a block at 23^3 nodes, with nothing to overlap them (one
    // workgroup a CU)
    constexpr int kSU = 12;
    for (int t0 = tid; t0 < nb; t0 += kSU * nt) {
        float v[kSU][3];
#pragma unroll
        for (int u = 0; u < kSU; ++u) {
            const int t = min(t0 + u * nt, nb - 1);
            int tx, ty, tz;
            split3(sp3, t, tx, ty, tz);
            const size_t node =
                ((size_t)wrap3(o2 + tz, g.M[2]) * g.M[1] + wrap3(o1 + ty, g.M[1])) * g.M[0] + wrap3(o0 + tx, g.M[0]);
            v[u][0] = emesh[3 * node + 0];
            v[u][1] = emesh[3 * node + 1];
            v[u][2] = emesh[3 * node + 2];
        }
#pragma unroll
        for (int u = 0; u < kSU; ++u) {
            const int t = t0 + u * nt;
            if (t < nb) {
                int tx, ty, tz;
                split3(sp3, t, tx, ty, tz);
                float *d = eb + 3 * tx + Sy * ty + Sz * tz;
                d[0] = v[u][0];
                d[1] = v[u][1];
                d[2] = v[u][2];
            }
        }
    }
    const int nrun = pm_runs(g, P.cell_start, blk, run_start, run_off);  // ends with a barrier
    const int npart = run_off[nrun];
    // visit order: ord[slot] = run << 10 | offset in the run of the particle
    // to visit (nrun <= 64; a block with a run of 1024 or more keeps the
    // sorted order)
    unsigned short *ord = (unsigned short *)(eraw + 4 * (size_t)Sz * B2);
    bool reorder = LDS && npart <= P.br.ord_cap;  // ord_cap <= 4 blockDim: four particles a thread
    int kr[4], ko[4];
    if (reorder) {
        if (tid < 17) bucket[tid] = 0;  // [16]: set when an offset does not fit 10 bits
        __syncthreads();
        int run = 0;
        double px4[4], py4[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {  // the four position loads first
            const int k = tid + q * nt;
            ko[q] = -1;
            if (k < npart) {
                const double4 p = P.spos[pm_particle_seq(run_start, run_off, nrun, k, run)];
                px4[q] = p.x;
                py4[q] = p.y;
                const int off = k - run_off[run];
                ko[q] = (run << 10) | off;
                if (off >= 1024) bucket[16] = 1;
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            kr[q] = -1;
            if (ko[q] >= 0) {
                const int rx = (int)floor(px4[q] * g.ih[0]) - o0, ry = (int)floor(py4[q] * g.ih[1]) - o1;
                const int key = (3 * rx + Sy * ry) & 15;
                kr[q] = (key << 12) | atomicAdd(&bucket[key], 1);
            }
        }
        __syncthreads();
        reorder = bucket[16] == 0;
    }
    if (reorder) {
        int neven = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) neven += bucket[2 * j];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (kr[q] < 0) continue;
            const int key = kr[q] >> 12, rank = kr[q] & 4095, par = key & 1, jb = key >> 1;
            int pos = par ? neven : 0;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int sz = bucket[2 * j + par];
                pos += min(sz, rank) + ((j < jb && sz > rank) ? 1 : 0);
            }
            ord[pos] = (unsigned short)ko[q];
        }
        __syncthreads();
    }
    double ksum = 0.0;
    int run_nat = 0;
    const int sub = tid % kGatherLanes;
    const int xh = sub & 1, zh = sub >> 1;
    const int per_pass = nt / kGatherLanes;
    const int npass = (npart + per_pass - 1) / per_pass;
    const GatherOut &o = P.o;
    const bool want_pp = o.include_pp || o.acc_pp;
    // sorted index of the particle at a slot of the visit order
    auto locate = [&](int slot) -> int {
        if (reorder) {
            const int v = ord[slot];
            return run_start[v >> 10] + (v & 1023);
        }
        return pm_particle_seq(run_start, run_off, nrun, slot, run_nat);
    };
    // each pass's position is loaded one pass ahead, and its epilogue operands
    // (caller index, PP acceleration) before the field sums, so the memory
    // latencies overlap the LDS work.  The loads are unconditional (slots past
    // the block's end read its last particle and discard it): the compiler
    // then waits for each by its exact count of later loads, where loads
    // under branches made it wait for all of them
    const int last = npart - 1;
    int slot = tid / kGatherLanes;
    int s = 0;
    double4 p = make_double4(0.0, 0.0, 0.0, 0.0);
    if (npart > 0) {
        s = locate(min(slot, last));
        p = P.spos[s];
    }
    for (int pass = 0; pass < npass; ++pass) {
        const bool live = slot < npart;
        const int ic = P.perm[s];
        const double pp = P.acc_pp_s[3 * (size_t)s + (sub < 3 ? sub : 2)];
        const int slot_n = slot + per_pass;
        const int s_n = locate(min(slot_n, last));
        const double4 p_n = P.spos[s_n];
        float ex = 0.0f, ey = 0.0f, ez = 0.0f;
        if (live) {
            float fx, fy, fz;
            const int ix = base_and_offset<float>(p.x, g.ih[0], fx);
            const int iy = base_and_offset<float>(p.y, g.ih[1], fy);
            const int iz = base_and_offset<float>(p.z, g.ih[2], fz);
            // x: nodes 0-2 take w0, w1, w2 at x; nodes 3-5 take w2, w1, w0 at -x
            const float sx = xh ? -fx : fx, sz = zh ? -fz : fz;
            const float a0 = w6_0<float>(sx), a1 = w6_1<float>(sx), a2 = w6_2<float>(sx);
            const float wx0 = xh ? a2 : a0, wx2 = xh ? a0 : a2;
            float wy[6];
            weights6<float>(fy, wy);
            // z: planes 0, 2, 4 take w0, w2 at z and w1 at -z; planes 5, 3, 1 the same at -z
            const float wz[3] = {w6_0<float>(sz), w6_2<float>(sz), w6_1<float>(-sz)};
            const int rx = ix - o0, ry = iy - o1, rz = iz - o2;
            const bool inside =
                LDS && rx >= 0 && ry >= 0 && rz >= 0 && rx + 5 < B0 && ry + 5 < B1 && rz + 5 < B2;
            if (inside) {
                const float *base = eb + 3 * (rx + 3 * xh) + Sy * ry + Sz * rz;
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    const float *plane = base + Sz * (zh ? 5 - 2 * m : 2 * m);
                    // the plane's six rows summed per x node (weights wy), then
                    // the three nodes (weights wx): 9 fewer FMAs a plane than
                    // row by row
                    float cs[9];
#pragma unroll
                    for (int u = 0; u < 9; ++u) cs[u] = wy[0] * plane[u];
#pragma unroll
                    for (int b = 1; b < 6; ++b) {
                        const float *row = plane + Sy * b;
#pragma unroll
                        for (int u = 0; u < 9; ++u) cs[u] = fmaf(wy[b], row[u], cs[u]);
                    }
                    const float px = fmaf(wx2, cs[6], fmaf(a1, cs[3], wx0 * cs[0]));
                    const float py = fmaf(wx2, cs[7], fmaf(a1, cs[4], wx0 * cs[1]));
                    const float pz = fmaf(wx2, cs[8], fmaf(a1, cs[5], wx0 * cs[2]));
                    ex = fmaf(wz[m], px, ex);
                    ey = fmaf(wz[m], py, ey);
                    ez = fmaf(wz[m], pz, ez);
                }
            } else {
                const float wxl[3] = {wx0, a1, wx2};
                for (int m = 0; m < 3; ++m) {
                    const int gz = wrap3(iz + (zh ? 5 - 2 * m : 2 * m), g.M[2]);
                    for (int b = 0; b < 6; ++b) {
                        const int gy = wrap3(iy + b, g.M[1]);
                        const float *row = emesh + 3 * (((size_t)gz * g.M[1] + gy) * g.M[0]);
                        const float wzy = wz[m] * wy[b];
                        for (int c = 0; c < 3; ++c) {
                            const int gx = wrap3(ix + 3 * xh + c, g.M[0]);
                            const float ww = wzy * wxl[c];
                            ex += row[3 * gx + 0] * ww;
                            ey += row[3 * gx + 1] * ww;
                            ez += row[3 * gx + 2] * ww;
                        }
                    }
                }
            }
        }
        // sum the four lanes of each particle
        ex = quad_sum(ex);
        ey = quad_sum(ey);
        ez = quad_sum(ez);
        if (live && sub < 3) ksum += gather_out_c<float>(P, sub, ic, p.w, want_pp ? pp : 0.0, ex, ey, ez);
        slot = slot_n;
        s = s_n;
        p = p_n;
    }
    if (P.o.vel) {
        ksum = wave_sum(ksum);
        if ((tid & 63) == 0) red[tid >> 6] = ksum;
        __syncthreads();
        if (tid == 0) {
            double t = 0.0;
            for (int w = 0; w < kGatherThreads / 64; ++w) t += red[w];
            P.part[bk.x] = t;
        }
    }
}

template <bool LDS, class T>
__global__ __launch_bounds__(kGatherThreads) void k_gather(GatherArgs P) {
    extern __shared__ __align__(16) unsigned char eraw[];
    __shared__ double red[kGatherThreads / 64];
    __shared__ int run_start[64], run_off[65];
    if constexpr (sizeof(T) == 4) {
        __shared__ int bucket[17];
        gather_body_f32<LDS>(P, blk_plain(), eraw, red, run_start, run_off, bucket);
    } else {
        gather_body<LDS, T>(P, blk_plain(), eraw, red, run_start, run_off);
    }
}
template <class T>
__global__ __launch_bounds__(kGatherThreads) void k_gather_seg(const GatherArgs *__restrict__ tab, SegGrid S) {
    extern __shared__ __align__(16) unsigned char eraw[];
    __shared__ double red[kGatherThreads / 64];
    __shared__ int run_start[64], run_off[65];
    Blk b;
    const int r = seg_locate(S, b);
    if (tab[r].br.lds_gather) {
        if constexpr (sizeof(T) == 4) {
            __shared__ int bucket[17];
            gather_body_f32<true>(tab[r], b, eraw, red, run_start, run_off, bucket);
        } else {
            gather_body<true, T>(tab[r], b, eraw, red, run_start, run_off);
        }
    } else {
        if constexpr (sizeof(T) == 4) {
            __shared__ int bucket[17];
            gather_body_f32<false>(tab[r], b, eraw, red, run_start, run_off, bucket);
        } else {
            gather_body<false, T>(tab[r], b, eraw, red, run_start, run_off);
        }
    }
}

static GatherArgs gather_args(mdf_p3m_ctx *c, const GatherOut &o) {
    GatherArgs P;
    P.g = c->g;
    P.br = c->brick;
    P.spos = c->spos;
    P.cell_start = c->cell_start;
    P.perm = c->perm;
    P.acc_pp_s = c->acc_pp_s;
    P.part = c->part_red;
    P.o = o;
    const bool mixed = c->prm.precision == MDF_PREC_MIXED;
    P.emesh = mixed ? (const void *)c->emeshf : (const void *)c->emesh;
    return P;
}

int pm_gather(mdf_p3m_ctx *c, int n, const GatherOut &o, hipStream_t s) {
    (void)n;
    const GatherArgs P = gather_args(c, o);
    const bool mixed = c->prm.precision == MDF_PREC_MIXED;
    const dim3 grid(c->brick.nblock_gather);
    if (c->brick.lds_gather) {
        const size_t lds = gather_lds_bytes(c->brick, mixed);
        if (mixed) {
            MDF_TRY(allow_lds(k_gather<true, float>, lds));
            hipLaunchKernelGGL((k_gather<true, float>), grid, dim3(kGatherThreads), lds, s, P);
        } else {
            MDF_TRY(allow_lds(k_gather<true, double>, lds));
            hipLaunchKernelGGL((k_gather<true, double>), grid, dim3(kGatherThreads), lds, s, P);
        }
    } else if (mixed) {
        hipLaunchKernelGGL((k_gather<false, float>), grid, dim3(kGatherThreads), 0, s, P);
    } else {
        hipLaunchKernelGGL((k_gather<false, double>), grid, dim3(kGatherThreads), 0, s, P);
    }
    MDF_LAUNCHED();
    return MDF_OK;
}

// ---- segmented stages (replica batches) ----
struct FillArgs {
    double *p;
    size_t n;
};
__global__ void k_fill_seg(const FillArgs *__restrict__ tab, SegGrid S) {
    Blk b;
    const int r = seg_locate(S, b);
    const FillArgs A = tab[r];
    for (size_t i = (size_t)b.x * blockDim.x + threadIdx.x; i < A.n; i += (size_t)b.nx * blockDim.x) A.p[i] = 0.0;
}
static int build_fill_rho(mdf_p3m_ctx *c, const StepIO &, void *args, int4 &dims, size_t &lds) {
    *(FillArgs *)args = FillArgs{c->rho, c->nmesh};
    size_t nb = (c->nmesh + 255) / 256;
    if (nb > 1024) nb = 1024;
    dims = make_int4((int)nb, 1, 1, 0);
    lds = 0;
    return MDF_OK;
}
static int launch_fill(const void *tab, const SegGrid &S, int total, size_t, int, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_seg, dim3(total), dim3(256), 0, s, (const FillArgs *)tab, S);
    MDF_LAUNCHED();
    return MDF_OK;
}
template <class T>
static int build_assign(mdf_p3m_ctx *c, const StepIO &, void *args, int4 &dims, size_t &lds) {
    *(AssignArgs *)args = assign_args(c);
    dims = make_int4(c->brick.nblock_assign, 1, 1, 0);
    lds = c->brick.lds_assign ? (c->prm.precision == MDF_PREC_MIXED ? 4 : 8) * (size_t)c->brick.B[0] * c->brick.B[1] *
                                    c->brick.B[2]
                              : 0;
    return MDF_OK;
}
template <class T>
static int launch_assign(const void *tab, const SegGrid &S, int total, size_t lds, int, hipStream_t s) {
    MDF_TRY(allow_lds(k_assign_seg<T>, lds));
    hipLaunchKernelGGL(k_assign_seg<T>, dim3(total), dim3(kPMThreads), lds, s, (const AssignArgs *)tab, S);
    MDF_LAUNCHED();
    return MDF_OK;
}
template <class T>
static int build_gather(mdf_p3m_ctx *c, const StepIO &io, void *args, int4 &dims, size_t &lds) {
    GatherOut o{};
    o.acc = io.acc;
    o.af = io.af;
    o.include_pp = 1;
    o.vel = io.vel;
    o.kick_dt = io.dt;
    *(GatherArgs *)args = gather_args(c, o);
    dims = make_int4(c->brick.nblock_gather, 1, 1, 0);
    lds = c->brick.lds_gather ? gather_lds_bytes(c->brick, sizeof(T) == 4) : 0;
    return MDF_OK;
}
template <class T>
static int launch_gather(const void *tab, const SegGrid &S, int total, size_t lds, int, hipStream_t s) {
    MDF_TRY(allow_lds(k_gather_seg<T>, lds));
    hipLaunchKernelGGL(k_gather_seg<T>, dim3(total), dim3(kGatherThreads), lds, s, (const GatherArgs *)tab, S);
    MDF_LAUNCHED();
    return MDF_OK;
}

void seg_stages_pm(std::vector<SegStage> &out, int precision, int part) {
    const bool mixed = precision == MDF_PREC_MIXED;
    if (part == 0) {
        out.push_back({"fill_rho", sizeof(FillArgs), build_fill_rho, launch_fill});
        if (mixed) out.push_back({"assign", sizeof(AssignArgs), build_assign<float>, launch_assign<float>});
        else out.push_back({"assign", sizeof(AssignArgs), build_assign<double>, launch_assign<double>});
    } else {
        if (mixed) out.push_back({"gather", sizeof(GatherArgs), build_gather<float>, launch_gather<float>});
        else out.push_back({"gather", sizeof(GatherArgs), build_gather<double>, launch_gather<double>});
    }
}

}  // namespace mdf
