// Optimal influence function G(k) for ik-differentiation, on the device.
//
// Reference: gf_opt, yukawa_gf_opt.py:5-98.  For every mode k != 0 of the
// shifted grid (n <-> k = 2 pi (n - n_mid)/L, :12-33):
//   G(k) = sum_m U^2(k_m) R(k_m) (k . k_m) / [ (sum_m U^2(k_m))^2 k^2 ],
//   k_m = k + 2 pi m M / L, m in [-m_max, m_max]^3 (summed mz, my, mx as in :56-58),
//   U(k) = prod_d sinc(k_d h_d / 2)^p with sinc(0) = 1 (:68-84),
//   R(k) = 4 pi exp(-(kappa^2 + k^2)/(4 alpha^2)) / (kappa^2 + k^2) (:86),
// G(0) = 0, and the error sum A_pm = sum U_G^2 / ((sum U^2)^2 k^2) (:96).
// One thread per mode; the 2 m_max + 1 sinc factors of each axis are computed
// once per mode (they depend on one axis only).  The result is written to the
// context's natural-order half spectrum and, optionally, to a full array in
// the reference's fftshift order.
#include "mdf_internal.h"

namespace mdf {

constexpr int kMaxAlias = 7;  // m_max <= 3

struct GreenArgs {
    int M[3];
    int Mxh;
    double h[3], L[3];
    double kappa_sq, Gew_sq;
    int p;
    int mm[3];
    double *__restrict__ half;
    double *__restrict__ full;
    double *__restrict__ part;
    int q_mode;  // 1: reduce the error functional terms instead of writing G
};

__device__ __forceinline__ double ipow(double x, int p) {
    double r = 1.0;
    for (int i = 0; i < p; ++i) r *= x;
    return r;
}

__global__ __launch_bounds__(256) void k_green_build(GreenArgs P) {
    __shared__ double red[4];
    const size_t total = (size_t)P.M[0] * P.M[1] * P.M[2];
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double contrib = 0.0;
    if (idx < total) {
        const int n[3] = {(int)(idx % P.M[0]), (int)((idx / P.M[0]) % P.M[1]), (int)(idx / ((size_t)P.M[0] * P.M[1]))};
        int nsh[3];
        double k[3];
        double U[3][kMaxAlias];
        double km[3][kMaxAlias];
        for (int d = 0; d < 3; ++d) {
            const int mid = P.M[d] / 2;  // (M-1)/2 for odd M
            nsh[d] = n[d] - mid;
            k[d] = 2.0 * M_PI * nsh[d] / P.L[d];
            for (int m = -P.mm[d]; m <= P.mm[d]; ++m) {
                const double kM = 2.0 * M_PI * (nsh[d] + m * P.M[d]) / P.L[d];
                km[d][m + P.mm[d]] = kM;
                U[d][m + P.mm[d]] = kM != 0.0 ? sin(0.5 * kM * P.h[d]) / (0.5 * kM * P.h[d]) : 1.0;
            }
        }
        const double k_sq = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
        double G = 0.0;
        if (k_sq != 0.0) {
            double U_k_sq = 0.0, U_G_k = 0.0, R_sq = 0.0;
            for (int mz = 0; mz <= 2 * P.mm[2]; ++mz)
                for (int my = 0; my <= 2 * P.mm[1]; ++my)
                    for (int mx = 0; mx <= 2 * P.mm[0]; ++mx) {
                        const double kx = km[0][mx], ky = km[1][my], kz = km[2][mz];
                        const double k_M_sq = kx * kx + ky * ky + kz * kz;
                        const double Uk = ipow(U[0][mx] * U[1][my] * U[2][mz], P.p);
                        const double U2 = Uk * Uk;
                        const double GM = 4.0 * M_PI * exp(-0.25 * (P.kappa_sq + k_M_sq) / P.Gew_sq) /
                                          (P.kappa_sq + k_M_sq);
                        const double kdot = k[0] * kx + k[1] * ky + k[2] * kz;
                        U_G_k += U2 * GM * kdot;
                        U_k_sq += U2;
                        R_sq += k_M_sq * GM * GM;
                    }
            G = U_G_k / ((U_k_sq * U_k_sq) * k_sq);
            contrib = U_G_k * U_G_k / ((U_k_sq * U_k_sq) * k_sq);
            if (P.q_mode) contrib = R_sq - contrib;  // this k's term of the error functional
        }
        if (!P.q_mode) {
            if (P.full) P.full[idx] = G;
            // shifted index n holds natural index (n + M - mid) % M
            int nat[3];
            for (int d = 0; d < 3; ++d) nat[d] = (n[d] + P.M[d] - P.M[d] / 2) % P.M[d];
            if (nat[0] < P.Mxh) P.half[((size_t)nat[2] * P.M[1] + nat[1]) * P.Mxh + nat[0]] = G;
        }
    }
    contrib = wave_sum(contrib);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = contrib;
    __syncthreads();
    if (threadIdx.x == 0) P.part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ __launch_bounds__(256) void k_sum_partials(const double *__restrict__ part, int n, double *__restrict__ out) {
    __shared__ double red[4];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += part[i];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) *out = red[0] + red[1] + red[2] + red[3];
}

__global__ void k_green_from_full(const double *__restrict__ full, double *__restrict__ half, int M0, int M1, int M2,
                                  int Mxh) {
    const size_t total = (size_t)M2 * M1 * Mxh;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int nx = (int)(i % Mxh), ny = (int)((i / Mxh) % M1), nz = (int)(i / ((size_t)Mxh * M1));
    // natural n sits at shifted position (n + M/2) % M (numpy fftshift)
    const int sx = (nx + M0 / 2) % M0, sy = (ny + M1 / 2) % M1, sz = (nz + M2 / 2) % M2;
    half[i] = full[((size_t)sz * M1 + sy) * M0 + sx];
}

int build_green(mdf_p3m_ctx *c, int p, int mx, int my, int mz, double *G_full, double *A_pm, hipStream_t s) {
    if (mx < 0 || my < 0 || mz < 0 || mx > 3 || my > 3 || mz > 3 || p < 1 || p > 16) {
        set_error("build_green: need 0 <= m_max <= 3 and 1 <= p <= 16 (got p=%d, m=%d,%d,%d)", p, mx, my, mz);
        return MDF_EINVAL;
    }
    GreenArgs P;
    for (int d = 0; d < 3; ++d) {
        P.M[d] = c->g.M[d];
        P.h[d] = c->g.h[d];
        P.L[d] = c->g.L[d];
    }
    P.Mxh = c->g.Mxh;
    P.kappa_sq = c->g.kappa * c->g.kappa;
    P.Gew_sq = c->g.alpha * c->g.alpha;
    P.p = p;
    P.mm[0] = mx;
    P.mm[1] = my;
    P.mm[2] = mz;
    P.half = c->green;
    P.full = G_full;
    P.part = c->part_red;
    P.q_mode = 0;
    const size_t total = c->nmesh;
    const unsigned nb = (unsigned)((total + 255) / 256);
    if (nb > (unsigned)kMaxPartials) {
        set_error("build_green: mesh too large");
        return MDF_EINVAL;
    }
    hipLaunchKernelGGL(k_green_build, dim3(nb), dim3(256), 0, s, P);
    MDF_LAUNCHED();
    if (A_pm) {
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, c->part_red, (int)nb, A_pm);
        MDF_LAUNCHED();
    }
    c->have_green = true;
    return green_to_float(c, s);
}

// The Hockney-Eastwood error functional of the optimal influence function,
// Q_sum = sum_{k != 0} [ sum_m |k_m|^2 G(k_m)^2 - A_k ], A_k the k term of
// A_pm (the same alias sums as build_green).  For N random unit charges in a
// box of volume V the mesh force error is sqrt(N Q_sum) / V.
int pm_error_sum(mdf_p3m_ctx *c, int p, int mx, int my, int mz, double *Q_sum, hipStream_t s) {
    if (mx < 0 || my < 0 || mz < 0 || mx > 3 || my > 3 || mz > 3 || p < 1 || p > 16 || !Q_sum) {
        set_error("pm_error_sum: need 0 <= m_max <= 3, 1 <= p <= 16 and an output");
        return MDF_EINVAL;
    }
    GreenArgs P;
    for (int d = 0; d < 3; ++d) {
        P.M[d] = c->g.M[d];
        P.h[d] = c->g.h[d];
        P.L[d] = c->g.L[d];
    }
    P.Mxh = c->g.Mxh;
    P.kappa_sq = c->g.kappa * c->g.kappa;
    P.Gew_sq = c->g.alpha * c->g.alpha;
    P.p = p;
    P.mm[0] = mx;
    P.mm[1] = my;
    P.mm[2] = mz;
    P.half = nullptr;
    P.full = nullptr;
    P.part = c->part_red;
    P.q_mode = 1;
    const unsigned nb = (unsigned)((c->nmesh + 255) / 256);
    if (nb > (unsigned)kMaxPartials) {
        set_error("pm_error_sum: mesh too large");
        return MDF_EINVAL;
    }
    hipLaunchKernelGGL(k_green_build, dim3(nb), dim3(256), 0, s, P);
    MDF_LAUNCHED();
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, c->part_red, (int)nb, Q_sum);
    MDF_LAUNCHED();
    return MDF_OK;
}

int set_green(mdf_p3m_ctx *c, const double *G_full, hipStream_t s) {
    const size_t total = c->nspec;
    hipLaunchKernelGGL(k_green_from_full, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, G_full, c->green,
                       c->g.M[0], c->g.M[1], c->g.M[2], c->g.Mxh);
    MDF_LAUNCHED();
    c->have_green = true;
    return green_to_float(c, s);
}

__global__ void k_to_float(const double *__restrict__ src, float *__restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = (float)src[i];
}

int green_to_float(mdf_p3m_ctx *c, hipStream_t s) {
    if (!c->greenf) return MDF_OK;
    const size_t n = c->nspec;
    size_t nb = (n + 255) / 256;
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(k_to_float, dim3((unsigned)nb), dim3(256), 0, s, c->green, c->greenf, n);
    MDF_LAUNCHED();
    return MDF_OK;
}

}  // namespace mdf
