// Dynamic structure factor S(q, omega) from the density modes n(q, t).
//
// Reference: n_of_q_w, s_of_q_w.py:3-20 (per axis: n(q, w) = dt * FFT_t n(q, t),
// S = |n(q, w)|^2 / (N T), T the total time); s_of_q_w_avg_mat,
// s_of_q_w_avg.py:5-20 (per run and q index i: the series n_q_t[:, 3 i, :],
// then the mean of S over the three axes); the run average of
// s_of_q_w_script.py:36-37 is a mean over the first axis of the result.
//
// The temporal DFT runs on the device as a mixed-radix Stockham transform in
// FP64, all series of all runs in one launch per stage:
// * the length factors as 4^a 2^b 3^c 5^d and then any other primes; radices
//   2..5 run one thread per butterfly with the R inputs in registers, larger
//   primes one thread per output (R terms);
// * stage s with radix R and Ns = product of the earlier radices reads
//   v_r = x[j + r N/R] (j < N/R), multiplies by exp(-2 pi i k r / (Ns R)),
//   k = j mod Ns, takes the R-point DFT and writes X_m to
//   (j / Ns) Ns R + k + m Ns: the output is in natural order after the last
//   stage, with no bit reversal;
// * every twiddle is sincospi of an exact rational (k r mod Ns R) / (Ns R), so
//   no angle error accumulates with the length.
// The first stage reads the reference's n_q_t layout (run, t, column, axis)
// directly; the others ping-pong between two workspace halves.  A last kernel
// forms |dt X|^2 / (N T) and, on request, the mean over the axes.
// Memory-bound: each stage streams the series once in and once out.
#include <mutex>

#include "mdf_internal.h"

namespace mdf {

namespace {

struct SqwIn {  // where element t of series s = (run, i, axis) is read
    const double2 *p;
    long long run_stride;  // elements between runs
    long long col_stride;  // between consecutive series i
    long long t_stride;    // between consecutive t
    int nq, raw;           // raw: the caller's (run, t, column, axis) array; else packed (s, t)
    int npd;
    __device__ __forceinline__ long long off(int s) const {
        if (!raw) return (long long)s * npd;
        const int a = s % 3, i = (s / 3) % nq, run = s / (3 * nq);
        return run * run_stride + i * col_stride + a;
    }
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// exp(-2 pi i num / den) for 0 <= num < den
__device__ __forceinline__ double2 cis_neg(long long num, long long den) {
    if (2 * num > den) num -= den;  // argument in [-1, 1]
    double s, c;
    sincospi(-2.0 * (double)num / (double)den, &s, &c);
    return make_double2(c, s);
}

template <int R>
__device__ __forceinline__ void small_dft(double2 *v) {
    if constexpr (R == 2) {
        const double2 a = v[0], b = v[1];
        v[0] = make_double2(a.x + b.x, a.y + b.y);
        v[1] = make_double2(a.x - b.x, a.y - b.y);
    } else if constexpr (R == 4) {
        const double2 a = v[0], b = v[1], c = v[2], d = v[3];
        const double2 s0 = make_double2(a.x + c.x, a.y + c.y), d0 = make_double2(a.x - c.x, a.y - c.y);
        const double2 s1 = make_double2(b.x + d.x, b.y + d.y), d1 = make_double2(b.x - d.x, b.y - d.y);
        v[0] = make_double2(s0.x + s1.x, s0.y + s1.y);
        v[2] = make_double2(s0.x - s1.x, s0.y - s1.y);
        // X1 = d0 - i d1, X3 = d0 + i d1
        v[1] = make_double2(d0.x + d1.y, d0.y - d1.x);
        v[3] = make_double2(d0.x - d1.y, d0.y + d1.x);
    } else {  // 3, 5: X_m = sum_r v_r w^(r m mod R)
        double2 w[R], x[R];
#pragma unroll
        for (int q = 0; q < R; ++q) w[q] = cis_neg(q, R);
#pragma unroll
        for (int m = 0; m < R; ++m) {
            double2 acc = v[0];
#pragma unroll
            for (int r = 1; r < R; ++r) {
                const double2 t = cmul(v[r], w[(r * m) % R]);
                acc.x += t.x;
                acc.y += t.y;
            }
            x[m] = acc;
        }
#pragma unroll
        for (int m = 0; m < R; ++m) v[m] = x[m];
    }
}

template <int R>
__global__ __launch_bounds__(256) void k_sqw_stage(SqwIn in, double2 *__restrict__ out, int npd, int ns,
                                                   long long total) {
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= total) return;
    const int m = npd / R;
    const int s = (int)(gid / m), j = (int)(gid % m);
    const int k = j % ns;
    const long long base = in.off(s);
    double2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = in.p[base + (long long)(j + r * m) * in.t_stride];
    if (ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[r] = cmul(v[r], cis_neg((long long)k * r, (long long)ns * R));
    }
    small_dft<R>(v);
    double2 *o = out + (long long)s * npd + (long long)(j / ns) * ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) o[(long long)r * ns] = v[r];
}

// radix R (a prime above 5): one thread per output element
__global__ __launch_bounds__(256) void k_sqw_stage_gen(SqwIn in, double2 *__restrict__ out, int npd, int ns, int R,
                                                       long long total) {
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= total) return;
    const int s = (int)(gid / npd), e = (int)(gid % npd);
    const long long nsr = (long long)ns * R;
    const int blk = (int)(e / nsr), rem = (int)(e % nsr);
    const int mo = rem / ns, k = rem % ns;
    const int m = npd / R, j = blk * ns + k;
    const long long base = in.off(s);
    const long long ph = k + (long long)mo * ns;  // phase step per r
    double2 acc = make_double2(0.0, 0.0);
    long long num = 0;
    for (int r = 0; r < R; ++r) {
        const double2 t = cmul(in.p[base + (long long)(j + r * m) * in.t_stride], cis_neg(num, nsr));
        acc.x += t.x;
        acc.y += t.y;
        num += ph;
        if (num >= nsr) num %= nsr;
    }
    out[(long long)s * npd + e] = acc;
}

// out: (run, i, w) means over the axes, or (run, i, w, axis)
__global__ __launch_bounds__(256) void k_sqw_finish(const double2 *__restrict__ X, int npd, long long nrow,
                                                    double scale, int mean_axes, double *__restrict__ out) {
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= nrow * npd) return;
    const long long row = gid / npd;  // (run, i)
    const int w = (int)(gid % npd);
    double sa[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double2 x = X[(row * 3 + a) * npd + w];
        sa[a] = (x.x * x.x + x.y * x.y) * scale;
    }
    if (mean_axes)
        out[gid] = (sa[0] + sa[1] + sa[2]) / 3.0;
    else
        for (int a = 0; a < 3; ++a) out[gid * 3 + a] = sa[a];
}

int factor_radices(int n, std::vector<int> &rs) {
    rs.clear();
    while (n % 4 == 0) rs.push_back(4), n /= 4;
    while (n % 2 == 0) rs.push_back(2), n /= 2;
    for (int p = 3; p <= 5; p += 2)
        while (n % p == 0) rs.push_back(p), n /= p;
    for (int p = 7; n > 1; p += 2) {
        if ((long long)p * p > n) p = n;  // n itself is prime
        while (n % p == 0) rs.push_back(p), n /= p;
    }
    return (int)rs.size();
}

}  // namespace

}  // namespace mdf

using namespace mdf;

extern "C" int mdf_sqw_workspace(int nrun, int npd, int nq, size_t *bytes) {
    if (nrun < 1 || npd < 1 || nq < 1 || !bytes) {
        set_error("mdf_sqw_workspace: nrun, npd, nq must be >= 1 and bytes non-NULL");
        return MDF_EINVAL;
    }
    *bytes = 2 * (size_t)nrun * nq * 3 * npd * sizeof(double2);
    return MDF_OK;
}

namespace mdf {
namespace {

// Every stage launches one thread per work item, at most nseries * npd of
// them, from a 32-bit block count: capped well inside the 2^32 threads of a
// HIP grid.
constexpr long long kSqwMaxItems = (1ll << 31) - 256;

int sqw_run(const double *nqt, int nrun, int npd, int nq_total, int q_stride, int nq, double N, double dt, double T,
            int mean_axes, double *out, void *work, size_t work_bytes, hipStream_t s);

}  // namespace
}  // namespace mdf

extern "C" int mdf_sqw(const double *nqt, int nrun, int npd, int nq_total, int q_stride, int nq, double N,
                       double dt, double T, int mean_axes, double *out, void *work, size_t work_bytes,
                       void *stream) {
    if (!work) {
        set_error("mdf_sqw: null workspace");
        return MDF_EINVAL;
    }
    return sqw_run(nqt, nrun, npd, nq_total, q_stride, nq, N, dt, T, mean_axes, out, work, work_bytes,
                   (hipStream_t)stream);
}

namespace mdf {

// ---- allocation-free form ----
// The stages' ping-pong buffers in a library-owned device array: each GPU the
// library runs on holds its own copy of MDF_SQW_STATIC_BYTES until the process
// ends.  Host code never passes the array's name to a kernel or copy (on the
// host it is a stand-in, not a device address): the address comes from
// hipGetSymbolAddress, taken once per device with that device current, and
// kept per device.
static __device__ __align__(256) unsigned char scratch[MDF_SQW_STATIC_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

int scratch_address(void **out) {
    int dev = -1;
    MDF_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("mdf_sqw_noalloc: device %d beyond the %d this library tracks", dev, kMaxDevices);
        return MDF_EINVAL;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("mdf_sqw_noalloc: hipGetSymbolAddress of the static scratch failed on device %d: %s", dev,
                      hipGetErrorString(e));
            return MDF_EHIP;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return MDF_OK;
}
}  // namespace

}  // namespace mdf

extern "C" int mdf_sqw_noalloc(const double *nqt, int nrun, int npd, int nq_total, int q_stride, int nq, double N,
                               double dt, double T, int mean_axes, double *out, void *stream) {
    size_t need = 0;
    MDF_TRY(mdf_sqw_workspace(nrun, npd, nq, &need));
    if (need > (size_t)MDF_SQW_STATIC_BYTES) {
        set_error("mdf_sqw_noalloc: %zu bytes of scratch needed, the static scratch holds %zu "
                  "(at most %zu series samples a call; use mdf_sqw with a workspace)",
                  need, (size_t)MDF_SQW_STATIC_BYTES, (size_t)MDF_SQW_STATIC_BYTES / 32);
        return MDF_EINVAL;
    }
    void *work = nullptr;
    MDF_TRY(scratch_address(&work));
    return sqw_run(nqt, nrun, npd, nq_total, q_stride, nq, N, dt, T, mean_axes, out, work,
                   (size_t)MDF_SQW_STATIC_BYTES, (hipStream_t)stream);
}

namespace mdf {
namespace {

int sqw_run(const double *nqt, int nrun, int npd, int nq_total, int q_stride, int nq, double N, double dt, double T,
            int mean_axes, double *out, void *work, size_t work_bytes, hipStream_t s) {
    if (!nqt || !out || !work || nrun < 1 || npd < 1 || nq < 1 || q_stride < 1 ||
        (long long)(nq - 1) * q_stride >= nq_total) {
        set_error("mdf_sqw: bad arguments (nrun %d, npd %d, nq_total %d, q_stride %d, nq %d)", nrun, npd, nq_total,
                  q_stride, nq);
        return MDF_EINVAL;
    }
    if (!(N * T != 0.0)) {
        set_error("mdf_sqw: N * T must be non-zero");
        return MDF_EINVAL;
    }
    size_t need = 0;
    MDF_TRY(mdf_sqw_workspace(nrun, npd, nq, &need));
    if (work_bytes < need) {
        set_error("mdf_sqw: workspace of %zu bytes, %zu needed", work_bytes, need);
        return MDF_EINVAL;
    }
    const long long nseries = (long long)nrun * nq * 3;
    if (nseries * npd > kSqwMaxItems) {
        set_error("mdf_sqw: %lld series samples exceed the %lld a call can take", nseries * npd, kSqwMaxItems);
        return MDF_EINVAL;
    }
    double2 *buf[2] = {(double2 *)work, (double2 *)work + nseries * npd};
    SqwIn in;
    in.p = (const double2 *)nqt;
    in.run_stride = (long long)npd * nq_total * 3;
    in.col_stride = (long long)q_stride * 3;
    in.t_stride = (long long)nq_total * 3;
    in.nq = nq;
    in.raw = 1;
    in.npd = npd;
    std::vector<int> rs;
    factor_radices(npd, rs);
    int cur = 0, ns = 1;
    if (rs.empty()) {  // npd == 1: X = x, copied through a radix-free "stage"
        rs.push_back(1);
    }
    for (int st = 0; st < (int)rs.size(); ++st) {
        const int R = rs[st];
        double2 *dst = buf[cur];
        const long long total = R == 1 ? nseries : (R <= 5 ? nseries * (npd / R) : nseries * npd);
        const unsigned nb = (unsigned)((total + 255) / 256);
        switch (R) {
            case 1: hipLaunchKernelGGL(k_sqw_stage_gen, dim3(nb), dim3(256), 0, s, in, dst, npd, 1, 1, total); break;
            case 2: hipLaunchKernelGGL(k_sqw_stage<2>, dim3(nb), dim3(256), 0, s, in, dst, npd, ns, total); break;
            case 3: hipLaunchKernelGGL(k_sqw_stage<3>, dim3(nb), dim3(256), 0, s, in, dst, npd, ns, total); break;
            case 4: hipLaunchKernelGGL(k_sqw_stage<4>, dim3(nb), dim3(256), 0, s, in, dst, npd, ns, total); break;
            case 5: hipLaunchKernelGGL(k_sqw_stage<5>, dim3(nb), dim3(256), 0, s, in, dst, npd, ns, total); break;
            default: hipLaunchKernelGGL(k_sqw_stage_gen, dim3(nb), dim3(256), 0, s, in, dst, npd, ns, R, total);
        }
        MDF_LAUNCHED();
        ns *= R;
        in.p = dst;
        in.raw = 0;
        in.t_stride = 1;
        cur ^= 1;
    }
    const long long nrow = (long long)nrun * nq;
    const unsigned nb = (unsigned)((nrow * npd + 255) / 256);
    const double scale = dt * dt / (N * T);
    hipLaunchKernelGGL(k_sqw_finish, dim3(nb), dim3(256), 0, s, in.p, npd, nrow, scale, mean_axes, out);
    MDF_LAUNCHED();
    return MDF_OK;
}

}  // namespace
}  // namespace mdf
