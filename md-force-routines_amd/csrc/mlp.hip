// Learned short-range correction: a pair MLP on BF16 MFMA (gfx950).
//
// The reduced P3M setting (short rc, coarse mesh) misses part of the force.
// A small network predicts, for every ordered pair (i, j) with r_ij < rc, a
// radial correction g_ij; particle i's force gains sum_j g_ij u_ij with
// u_ij = (x_i - x_j) / r_ij.  The network sees 16 features per pair:
//   f_0..f_11  Gaussians exp(-((r - mu_k) / sigma)^2), mu_k = mu0 + k dmu
//   f_12       local density of the pair: (rho_i + rho_j) / 2, rho the
//              particle count of the particle's PP cell over its mean, minus 1
//   f_13..f_15 kappa, alpha and rc of the setting (scaled run constants)
// and has three hidden layers of 64 SiLU units and one linear output:
//   z1 = x W1 + b1, h1 = silu(z1); z2 = h1 W2 + b2, h2 = silu(z2);
//   z3 = h2 W3 + b3, h3 = silu(z3); g = out_scale * (h3 . w4 + b4).
// Arithmetic: features, weights (W1..W3, w4) and activations x, h1..h3 are
// rounded to BF16 (round to nearest even); products accumulate in FP32 on
// v_mfma_f32_16x16x32_bf16; biases, pre-activations and the output are FP32.
// The backward pass (training) follows the same rounding: each dz_l is
// rounded to BF16 before it enters the weight-gradient products
// (v_mfma_f32_16x16x16bf16_1k over 16 pairs) and the products with W^T.
//
// Work layout.  A wave owns tiles of 16 pairs.  For the 16x16x32 products
// lane l holds row (or column) l % 16 and the 8 K-elements of group l / 16;
// A and B operands use the same (group, element) -> k assignment, so the
// product is the same whatever order the unit walks K in.  Results come back
// as rows 4 (l / 16) + i, column l % 16; hidden activations go through a
// per-wave LDS buffer to reach the A-operand layout of the next layer.  The
// weights are staged per workgroup in LDS as BF16, both transposed (B
// operands of the forward products) and row-major (B operands of the
// products with W^T in the backward pass).
//
// Pair lists are built from the context's sub-cell list (one thread per
// particle, two passes: count, then write at scanned offsets); the pair
// count stays on the device, and the pair kernels loop over tiles up to it,
// so nothing needs a host synchronisation.
#include <vector>

#include "mdf_internal.h"

namespace mdf {

typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 b8v __attribute__((ext_vector_type(8)));

constexpr int kIn = 16, kH = 64;
// packed FP32 parameter vector
constexpr int kOffW1 = 0;                       // [16][64]
constexpr int kOffB1 = kOffW1 + kIn * kH;       // [64]
constexpr int kOffW2 = kOffB1 + kH;             // [64][64]
constexpr int kOffB2 = kOffW2 + kH * kH;
constexpr int kOffW3 = kOffB2 + kH;
constexpr int kOffB3 = kOffW3 + kH * kH;
constexpr int kOffW4 = kOffB3 + kH;             // [64]
constexpr int kOffB4 = kOffW4 + kH;             // [1]
constexpr int kNParam = kOffB4 + 1;             // 9473
constexpr int kMLPThreads = 256;
constexpr int kMLPWaves = kMLPThreads / 64;

struct MLPPair {
    int i, j;           // sorted particle indices
    float r, ux, uy, uz;
};

}  // namespace mdf

struct mdf_mlp {
    int max_pairs = 0, n_max = 0;
    mdf_mlp_meta meta{};
    float *param = nullptr;     // kNParam (FP32 master copy)
    float *adam_m = nullptr, *adam_v = nullptr;
    float *grad = nullptr;      // kNParam
    float *gpart = nullptr;     // per-wave gradient partials
    int ngrid = 0;              // workgroups of the pair kernels
    mdf::MLPPair *pairs = nullptr;
    float *g = nullptr;         // per-pair output
    int *cnt = nullptr;         // n_max + 1 (pair counts -> offsets)
    int *scan_tmp = nullptr;
    float *rho = nullptr;       // per sorted particle density feature
    double *fpred = nullptr;    // (n, 3) sorted predicted correction (training)
    float *err = nullptr;       // (n, 3) sorted residual of the prediction
    double *loss_part = nullptr;
    double *loss = nullptr;     // device scalar
    int *flag = nullptr;        // [0] pair-list overflow
    int step = 0;
};

namespace mdf {

// ------------------------------------------------------------- helpers --
__device__ __forceinline__ unsigned short bf16_bits(float f) {
    unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_val(unsigned short b) { return __uint_as_float((unsigned)b << 16); }
__device__ __forceinline__ float bf16_round(float f) { return bf16_val(bf16_bits(f)); }

__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + __expf(-z)); }

// 8 BF16 (as 16 bytes) -> operand of the 16x16x32 product
__device__ __forceinline__ b8v as_b8(uint4 v) { return __builtin_bit_cast(b8v, v); }
__device__ __forceinline__ s4v as_s4(uint2 v) { return __builtin_bit_cast(s4v, v); }

// Orders LDS writes of a wave before its later LDS reads of other lanes'
// data (LDS instructions of one wave complete in order).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Weights and activation tiles in LDS (BF16 bits).  The forward kernel uses
// the first part only (39 KB, 4 workgroups per CU); the backward kernel adds
// the row-major weights and the dz tile (63 KB).
struct alignas(16) MLPFwdShared {
    unsigned short w1t[kH][32];   // W1^T, K padded to 32 with zeros
    unsigned short w2t[kH][kH];   // W2^T
    unsigned short w3t[kH][kH];   // W3^T
    float b1[kH], b2[kH], b3[kH], w4[kH];
    unsigned short xb[kMLPWaves][16][kIn];   // per-wave input tile
    unsigned short h1[kMLPWaves][16][kH];
    unsigned short h2[kMLPWaves][16][kH];
};
struct alignas(16) MLPShared : MLPFwdShared {
    unsigned short w2[kH][kH];    // W2 (backward)
    unsigned short w3[kH][kH];    // W3 (backward)
    unsigned short d[kMLPWaves][16][kH];     // dz tile (backward)
};

template <class SH>
__device__ void stage_weights(SH &S, const float *__restrict__ P, bool backward) {
    for (int t = threadIdx.x; t < kH * 32; t += blockDim.x) {
        const int n = t / 32, k = t % 32;
        S.w1t[n][k] = k < kIn ? bf16_bits(P[kOffW1 + k * kH + n]) : (unsigned short)0;
    }
    for (int t = threadIdx.x; t < kH * kH; t += blockDim.x) {
        const int a = t / kH, b = t % kH;
        S.w2t[b][a] = bf16_bits(P[kOffW2 + t]);  // t = a*64 + b: W2[a][b]
        S.w3t[b][a] = bf16_bits(P[kOffW3 + t]);
        if constexpr (sizeof(SH) > sizeof(MLPFwdShared)) {
            if (backward) {
                S.w2[a][b] = bf16_bits(P[kOffW2 + t]);
                S.w3[a][b] = bf16_bits(P[kOffW3 + t]);
            }
        }
    }
    for (int t = threadIdx.x; t < kH; t += blockDim.x) {
        S.b1[t] = P[kOffB1 + t];
        S.b2[t] = P[kOffB2 + t];
        S.b3[t] = P[kOffB3 + t];
        S.w4[t] = bf16_round(P[kOffW4 + t]);
    }
}

struct MLPArgs {
    const MLPPair *__restrict__ pairs;
    const int *__restrict__ off;      // off[n] = total pairs
    int n;
    int max_pairs;                    // list capacity: pairs beyond it were not written
    const float *__restrict__ rho;
    const float *__restrict__ param;
    mdf_mlp_meta meta;
    float *__restrict__ g;            // forward output
    // backward
    const float *__restrict__ err;    // (n,3) sorted residuals
    float coef;                       // dL/dg = coef * err_i . u
    float *__restrict__ gpart;        // per-wave partials [wave][kNParam]
};

// Features of pair m of the tile, in the A layout of the first product:
// group 0 -> f_0..f_7, group 1 -> f_8..f_15, groups 2, 3 -> the zero padding.
__device__ __forceinline__ uint4 pair_features(const MLPArgs &P, const MLPPair &pr, bool valid, int grp,
                                               float f[8]) {
#pragma unroll
    for (int e = 0; e < 8; ++e) f[e] = 0.0f;
    if (valid && grp < 2) {
        const mdf_mlp_meta &M = P.meta;
        if (grp == 0) {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const float t = (pr.r - (M.mu0 + e * M.dmu)) * M.inv_sigma;
                f[e] = __expf(-t * t);
            }
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float t = (pr.r - (M.mu0 + (8 + e) * M.dmu)) * M.inv_sigma;
                f[e] = __expf(-t * t);
            }
            f[4] = 0.5f * (P.rho[pr.i] + P.rho[pr.j]);
            f[5] = M.f_kappa;
            f[6] = M.f_alpha;
            f[7] = M.f_rc;
        }
    }
    unsigned short b[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) b[e] = bf16_bits(f[e]);
    uint4 v;
    v.x = b[0] | ((unsigned)b[1] << 16);
    v.y = b[2] | ((unsigned)b[3] << 16);
    v.z = b[4] | ((unsigned)b[5] << 16);
    v.w = b[6] | ((unsigned)b[7] << 16);
    return v;
}

// One hidden layer: acc[nb] = A . W (K = 64, two 16x16x32 steps) from the
// per-wave activation buffer H (16 x 64) and W^T in LDS.
__device__ __forceinline__ void layer64(const unsigned short (*H)[kH], const unsigned short (*WT)[kH], int m,
                                        int grp, f4v acc[4]) {
    const uint4 a0 = *(const uint4 *)&H[m][8 * grp];
    const uint4 a1 = *(const uint4 *)&H[m][32 + 8 * grp];
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
        const uint4 w0 = *(const uint4 *)&WT[16 * nb + m][8 * grp];
        const uint4 w1 = *(const uint4 *)&WT[16 * nb + m][32 + 8 * grp];
        f4v c = {0.f, 0.f, 0.f, 0.f};
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_b8(a0), as_b8(w0), c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_b8(a1), as_b8(w1), c, 0, 0, 0);
        acc[nb] = c;
    }
}

// Forward of one tile; keeps what the backward pass needs.
struct Fwd {
    f4v z1[4], z2[4], z3[4];
    float h3[4][4];  // BF16-rounded, C layout [nb][i]
    float out[4];    // network output of rows 4*grp + i (valid in every lane of the group)
};

template <class SH>
__device__ __forceinline__ void forward_tile(const MLPArgs &P, SH &S, int wave, int p0, int total, Fwd &F) {
    const int lane = threadIdx.x & 63, m = lane & 15, grp = lane >> 4;
    const int p = p0 + m;
    const bool valid = p < total;
    MLPPair pr{0, 0, 1.0f, 0.f, 0.f, 0.f};
    if (valid) pr = P.pairs[p];
    float f[8];
    const uint4 xa = pair_features(P, pr, valid, grp, f);
    if (grp < 2) {
#pragma unroll
        for (int e = 0; e < 8; ++e) S.xb[wave][m][8 * grp + e] = bf16_bits(f[e]);
    }
    // layer 1 (K = 16 padded to 32)
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
        const uint4 w = *(const uint4 *)&S.w1t[16 * nb + m][8 * grp];
        f4v c = {0.f, 0.f, 0.f, 0.f};
        F.z1[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_b8(xa), as_b8(w), c, 0, 0, 0);
    }
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
        const int col = 16 * nb + m;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            F.z1[nb][i] += S.b1[col];
            S.h1[wave][4 * grp + i][col] = bf16_bits(F.z1[nb][i] * sigm(F.z1[nb][i]));
        }
    }
    wave_lds_sync();
    layer64(S.h1[wave], S.w2t, m, grp, F.z2);
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
        const int col = 16 * nb + m;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            F.z2[nb][i] += S.b2[col];
            S.h2[wave][4 * grp + i][col] = bf16_bits(F.z2[nb][i] * sigm(F.z2[nb][i]));
        }
    }
    wave_lds_sync();
    layer64(S.h2[wave], S.w3t, m, grp, F.z3);
    float part[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
        const int col = 16 * nb + m;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            F.z3[nb][i] += S.b3[col];
            F.h3[nb][i] = bf16_round(F.z3[nb][i] * sigm(F.z3[nb][i]));
            part[i] += F.h3[nb][i] * S.w4[col];
        }
    }
    const float b4 = P.param[kOffB4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float v = part[i];
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) v += __shfl_xor(v, o, 64);
        F.out[i] = v + b4;
    }
}

__global__ __launch_bounds__(kMLPThreads) void k_mlp_forward(MLPArgs P) {
    __shared__ MLPFwdShared S;
    stage_weights(S, P.param, false);
    __syncthreads();
    const int total = min(P.off[P.n], P.max_pairs);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, m = lane & 15, grp = lane >> 4;
    const int ntile = (total + 15) / 16;
    for (int t = blockIdx.x * kMLPWaves + wave; t < ntile; t += gridDim.x * kMLPWaves) {
        Fwd F;
        forward_tile(P, S, wave, 16 * t, total, F);
        if (m == 0) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int p = 16 * t + 4 * grp + i;
                if (p < total) P.g[p] = P.meta.out_scale * F.out[i];
            }
        }
        wave_lds_sync();
    }
}

__device__ __forceinline__ float dsilu(float z) {
    const float s = sigm(z);
    return s * (1.0f + z * (1.0f - s));
}

__device__ __forceinline__ uint2 pack4(const unsigned short b[4]) {
    uint2 v;
    v.x = b[0] | ((unsigned)b[1] << 16);
    v.y = b[2] | ((unsigned)b[3] << 16);
    return v;
}

// Backward: per-wave gradient partials of the loss sum_i |F_i - target_i|^2
// scaled by coef/2 (dL/dg_p = coef * err_i . u_p).
__global__ __launch_bounds__(kMLPThreads) void k_mlp_backward(MLPArgs P) {
    __shared__ MLPShared S;
    stage_weights(S, P.param, true);
    __syncthreads();
    const int total = min(P.off[P.n], P.max_pairs);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, m = lane & 15, grp = lane >> 4;
    const int ntile = (total + 15) / 16;
    f4v gW1[4], gW2[4][4], gW3[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        gW1[a] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int b = 0; b < 4; ++b) gW2[a][b] = gW3[a][b] = f4v{0.f, 0.f, 0.f, 0.f};
    }
    float gb1[4] = {0, 0, 0, 0}, gb2[4] = {0, 0, 0, 0}, gb3[4] = {0, 0, 0, 0}, gw4[4] = {0, 0, 0, 0}, gb4 = 0.f;
    const float oscale = P.meta.out_scale;
    for (int t = blockIdx.x * kMLPWaves + wave; t < ntile; t += gridDim.x * kMLPWaves) {
        Fwd F;
        const int p0 = 16 * t;
        forward_tile(P, S, wave, p0, total, F);
        // dL/d(out) of rows 4*grp + i
        float dout[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int p = p0 + 4 * grp + i;
            float d = 0.f;
            if (p < total) {
                const MLPPair pr = P.pairs[p];
                const float *e = P.err + 3 * (size_t)pr.i;
                d = P.coef * oscale * (e[0] * pr.ux + e[1] * pr.uy + e[2] * pr.uz);
            }
            dout[i] = d;
        }
        if (m == 0) gb4 += dout[0] + dout[1] + dout[2] + dout[3];
        // layer 3: dz3 = dout w4 silu'(z3)
        unsigned short dz[4][4];
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            const int col = 16 * nb + m;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                gw4[nb] += dout[i] * F.h3[nb][i];
                dz[nb][i] = bf16_bits(dout[i] * S.w4[col] * dsilu(F.z3[nb][i]));
                gb3[nb] += bf16_val(dz[nb][i]);
                S.d[wave][4 * grp + i][col] = dz[nb][i];
            }
        }
        wave_lds_sync();
        // gW3 += h2^T dz3 (M = W3 rows, K = 16 pairs, N = W3 cols)
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
            unsigned short a[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) a[e] = S.h2[wave][4 * grp + e][16 * mb + m];
            const s4v A = as_s4(pack4(a));
#pragma unroll
            for (int nb = 0; nb < 4; ++nb)
                gW3[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A, as_s4(pack4(dz[nb])), gW3[mb][nb], 0, 0, 0);
        }
        // dh2 = dz3 W3^T  -> dz2
        f4v dh[4];
        layer64(S.d[wave], S.w3, m, grp, dh);
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            const int col = 16 * nb + m;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                dz[nb][i] = bf16_bits(dh[nb][i] * dsilu(F.z2[nb][i]));
                gb2[nb] += bf16_val(dz[nb][i]);
            }
            (void)col;
        }
        wave_lds_sync();
#pragma unroll
        for (int nb = 0; nb < 4; ++nb)
#pragma unroll
            for (int i = 0; i < 4; ++i) S.d[wave][4 * grp + i][16 * nb + m] = dz[nb][i];
        wave_lds_sync();
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
            unsigned short a[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) a[e] = S.h1[wave][4 * grp + e][16 * mb + m];
            const s4v A = as_s4(pack4(a));
#pragma unroll
            for (int nb = 0; nb < 4; ++nb)
                gW2[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A, as_s4(pack4(dz[nb])), gW2[mb][nb], 0, 0, 0);
        }
        // dh1 = dz2 W2^T -> dz1
        layer64(S.d[wave], S.w2, m, grp, dh);
#pragma unroll
        for (int nb = 0; nb < 4; ++nb)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                dz[nb][i] = bf16_bits(dh[nb][i] * dsilu(F.z1[nb][i]));
                gb1[nb] += bf16_val(dz[nb][i]);
            }
        // gW1 += x^T dz1 (M = 16 input features)
        {
            unsigned short a[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) a[e] = S.xb[wave][4 * grp + e][m];
            const s4v A = as_s4(pack4(a));
#pragma unroll
            for (int nb = 0; nb < 4; ++nb)
                gW1[nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A, as_s4(pack4(dz[nb])), gW1[nb], 0, 0, 0);
        }
        wave_lds_sync();
    }
    // write this wave's partials: C layout row 4*grp + i, column 16*nb + m
    float *G = P.gpart + (size_t)(blockIdx.x * kMLPWaves + wave) * kNParam;
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
        const int col = 16 * nb + m;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = 4 * grp + i;
            G[kOffW1 + row * kH + col] = gW1[nb][i];
#pragma unroll
            for (int mb = 0; mb < 4; ++mb) {
                G[kOffW2 + (16 * mb + row) * kH + col] = gW2[mb][nb][i];
                G[kOffW3 + (16 * mb + row) * kH + col] = gW3[mb][nb][i];
            }
        }
    }
    // bias and w4 partials: sum over the four lane groups
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
        float v1 = gb1[nb], v2 = gb2[nb], v3 = gb3[nb], v4 = gw4[nb];
#pragma unroll
        for (int o = 16; o < 64; o <<= 1) {
            v1 += __shfl_xor(v1, o, 64);
            v2 += __shfl_xor(v2, o, 64);
            v3 += __shfl_xor(v3, o, 64);
            v4 += __shfl_xor(v4, o, 64);
        }
        if (grp == 0) {
            G[kOffB1 + 16 * nb + m] = v1;
            G[kOffB2 + 16 * nb + m] = v2;
            G[kOffB3 + 16 * nb + m] = v3;
            G[kOffW4 + 16 * nb + m] = v4;
        }
    }
    float v = gb4;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) G[kOffB4] = v;
}

__global__ void k_fill_float_zero_mlp(float *__restrict__ a, float *__restrict__ b) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < kNParam) {
        a[k] = 0.0f;
        b[k] = 0.0f;
    }
}

__global__ void k_mlp_grad_reduce(const float *__restrict__ gpart, int nw, float *__restrict__ grad) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= kNParam) return;
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += gpart[(size_t)w * kNParam + k];
    grad[k] = (float)s;
}

__global__ void k_mlp_adam(float *__restrict__ p, float *__restrict__ m, float *__restrict__ v,
                           const float *__restrict__ g, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= kNParam) return;
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.0f - b1) * gk;
    const float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    p[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

// ---------------------------------------------------------- pair lists --
__device__ __forceinline__ void wrapc(int c, int n, double L, int &cw, double &sh) {
    if (c < 0) {
        cw = c + n;
        sh = -L;
    } else if (c >= n) {
        cw = c - n;
        sh = L;
    } else {
        cw = c;
        sh = 0.0;
    }
}

__device__ __forceinline__ int subcell_of(double x, double w, int n) {
    int c = (int)floor(x / w);
    return c < 0 ? 0 : (c >= n ? n - 1 : c);
}

// count (WRITE = false) or write (WRITE = true) the ordered pairs of sorted particle s
template <bool WRITE>
__global__ void k_mlp_pairs(Geom g, const double4 *__restrict__ spos, const int *__restrict__ cell_start, int n,
                            double rc2, int *__restrict__ cnt, MLPPair *__restrict__ out, int max_pairs,
                            int *__restrict__ flag) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const double4 pi = spos[s];
    const int cx = subcell_of(pi.x, g.w[0], g.nsub[0]);
    const int cy = subcell_of(pi.y, g.w[1], g.nsub[1]);
    const int cz = subcell_of(pi.z, g.w[2], g.nsub[2]);
    int k = 0;
    int base = 0;
    if (WRITE) base = cnt[s];
    for (int dz = -2; dz <= 2; ++dz) {
        int wz;
        double sz;
        wrapc(cz + dz, g.nsub[2], g.L[2], wz, sz);
        for (int dy = -2; dy <= 2; ++dy) {
            int wy;
            double sy;
            wrapc(cy + dy, g.nsub[1], g.L[1], wy, sy);
            for (int dx = -2; dx <= 2; ++dx) {
                int wx;
                double sx;
                wrapc(cx + dx, g.nsub[0], g.L[0], wx, sx);
                const int c = wx + g.nsub[0] * (wy + g.nsub[1] * wz);
                for (int j = cell_start[c]; j < cell_start[c + 1]; ++j) {
                    if (j == s) continue;
                    const double4 pj = spos[j];
                    const double ddx = pi.x - (pj.x + sx), ddy = pi.y - (pj.y + sy), ddz = pi.z - (pj.z + sz);
                    const double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
                    if (!(r2 < rc2) || !(r2 > 0.0)) continue;
                    if (WRITE) {
                        const int p = base + k;
                        if (p < max_pairs) {
                            const double r = sqrt(r2);
                            MLPPair q;
                            q.i = s;
                            q.j = j;
                            q.r = (float)r;
                            q.ux = (float)(ddx / r);
                            q.uy = (float)(ddy / r);
                            q.uz = (float)(ddz / r);
                            out[p] = q;
                        } else {
                            atomicOr(flag, 1);
                        }
                    }
                    ++k;
                }
            }
        }
    }
    if (!WRITE) cnt[s] = k;
}

// density feature per sorted particle: PP-cell (2x2x2 sub-cells) count over its mean, minus 1
__global__ void k_mlp_density(Geom g, const double4 *__restrict__ spos, const int *__restrict__ cell_start, int n,
                              float *__restrict__ rho) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const double4 p = spos[s];
    const int bx = subcell_of(p.x, g.w[0], g.nsub[0]) >> 1;
    const int by = subcell_of(p.y, g.w[1], g.nsub[1]) >> 1;
    const int bz = subcell_of(p.z, g.w[2], g.nsub[2]) >> 1;
    int c = 0;
    for (int h = 0; h < 8; ++h) {
        const int x = 2 * bx + (h & 1), y = 2 * by + ((h >> 1) & 1), z = 2 * bz + (h >> 2);
        const int cell = x + g.nsub[0] * (y + g.nsub[1] * z);
        c += cell_start[cell + 1] - cell_start[cell];
    }
    const double nblk = (double)g.nblk[0] * g.nblk[1] * g.nblk[2];
    rho[s] = (float)((double)c * nblk / (double)n - 1.0);
}

// exclusive scan of cnt[0..n) into cnt (cnt[n] = total), tiles of 1024
constexpr int kTile = 1024;

__device__ __forceinline__ int block_scan_256(int v, int *ws) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    if (lane == 63) ws[wv] = v;
    __syncthreads();
    int add = 0;
    for (int w = 0; w < wv; ++w) add += ws[w];
    __syncthreads();
    return v + add;
}

__global__ __launch_bounds__(256) void k_scan_a(const int *__restrict__ v, int n, int *__restrict__ tsum) {
    __shared__ int ws[4];
    const int base = blockIdx.x * kTile + threadIdx.x * 4;
    int s = 0;
    for (int q = 0; q < 4; ++q) s += base + q < n ? v[base + q] : 0;
    const int incl = block_scan_256(s, ws);
    if (threadIdx.x == 255) tsum[blockIdx.x] = incl;
}

__global__ __launch_bounds__(1024) void k_scan_b(int *__restrict__ tsum, int nt) {
    // one block; nt <= 8192
    __shared__ int ws[16];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    int v[8], s = 0;
    for (int q = 0; q < 8; ++q) {
        const int k = t * 8 + q;
        v[q] = k < nt ? tsum[k] : 0;
        s += v[q];
    }
    int incl = s;
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(incl, o, 64);
        if (lane >= o) incl += u;
    }
    if (lane == 63) ws[wv] = incl;
    __syncthreads();
    int add = 0;
    for (int w = 0; w < wv; ++w) add += ws[w];
    int run = incl - s + add;
    for (int q = 0; q < 8; ++q) {
        const int k = t * 8 + q;
        if (k < nt) tsum[k] = run;
        run += v[q];
    }
    if (t == 1023) tsum[nt] = run;
}

__global__ __launch_bounds__(256) void k_scan_c(int *__restrict__ v, int n, const int *__restrict__ tsum, int nt) {
    __shared__ int ws[4];
    const int base = blockIdx.x * kTile + threadIdx.x * 4;
    int x[4], s = 0;
    for (int q = 0; q < 4; ++q) {
        x[q] = base + q < n ? v[base + q] : 0;
        s += x[q];
    }
    const int incl = block_scan_256(s, ws);
    int run = tsum[blockIdx.x] + incl - s;
    for (int q = 0; q < 4; ++q) {
        if (base + q < n) v[base + q] = run;
        run += x[q];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) v[n] = tsum[nt];
}

// per sorted particle: F = sum_p g_p u_p over its pairs; MODE 0: acc += F
// (sorted PP acceleration), MODE 1: fpred = F
template <int MODE>
__global__ void k_mlp_force(const MLPPair *__restrict__ pairs, const int *__restrict__ off, const float *__restrict__ g,
                            int n, int max_pairs, double *__restrict__ out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    const int pend = min(off[s + 1], max_pairs);
    for (int p = off[s]; p < pend; ++p) {
        const double gp = g[p];
        fx += gp * pairs[p].ux;
        fy += gp * pairs[p].uy;
        fz += gp * pairs[p].uz;
    }
    if (MODE == 0) {
        out[3 * (size_t)s + 0] += fx;
        out[3 * (size_t)s + 1] += fy;
        out[3 * (size_t)s + 2] += fz;
    } else {
        out[3 * (size_t)s + 0] = fx;
        out[3 * (size_t)s + 1] = fy;
        out[3 * (size_t)s + 2] = fz;
    }
}

// residual err_s = fpred_s - target[perm[s]] and per-block sums of |err|^2
__global__ __launch_bounds__(256) void k_mlp_residual(const double *__restrict__ fpred, const double *__restrict__ target,
                                                      const int *__restrict__ perm, int n, float *__restrict__ err,
                                                      double *__restrict__ part) {
    __shared__ double red[4];
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    double l = 0.0;
    if (s < n) {
        const int i = perm[s];
        for (int d = 0; d < 3; ++d) {
            const double e = fpred[3 * (size_t)s + d] - target[3 * (size_t)i + d];
            err[3 * (size_t)s + d] = (float)e;
            l += e * e;
        }
    }
    l = wave_sum(l);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void k_mlp_loss(const double *__restrict__ part, int nb, double scale, double *__restrict__ loss) {
    __shared__ double red[4];
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) s += part[i];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) *loss = scale * (red[0] + red[1] + red[2] + red[3]);
}

// ---------------------------------------------------------- launchers --
static MLPArgs mlp_args(mdf_mlp *m, int n) {
    MLPArgs A{};
    A.pairs = m->pairs;
    A.off = m->cnt;
    A.n = n;
    A.max_pairs = m->max_pairs;
    A.rho = m->rho;
    A.param = m->param;
    A.meta = m->meta;
    A.g = m->g;
    A.err = m->err;
    A.gpart = m->gpart;
    return A;
}

static int check_mlp(mdf_mlp *m, mdf_p3m_ctx *c, int n) {
    if (!m || !c) {
        set_error("mlp: null argument");
        return MDF_EINVAL;
    }
    if (n < 0 || n > m->n_max || n > c->prm.n_max) {
        set_error("mlp: particle count %d outside the network's (%d) or context's (%d) capacity", n, m->n_max,
                  c->prm.n_max);
        return MDF_EINVAL;
    }
    return MDF_OK;
}

// Pair list and density features of the context's current binning.
int mlp_build_pairs(mdf_mlp *m, mdf_p3m_ctx *c, int n, hipStream_t s) {
    if (n <= 0) return fill_zero_i(m->cnt, 1, s);
    const unsigned nb = (unsigned)((n + 255) / 256);
    const double rc2 = c->g.rc * c->g.rc;
    hipLaunchKernelGGL(k_mlp_density, dim3(nb), dim3(256), 0, s, c->g, c->spos, c->cell_start, n, m->rho);
    hipLaunchKernelGGL((k_mlp_pairs<false>), dim3(nb), dim3(256), 0, s, c->g, c->spos, c->cell_start, n, rc2, m->cnt,
                       m->pairs, m->max_pairs, m->flag);
    const int nt = (n + kTile - 1) / kTile;
    hipLaunchKernelGGL(k_scan_a, dim3(nt), dim3(256), 0, s, m->cnt, n, m->scan_tmp);
    hipLaunchKernelGGL(k_scan_b, dim3(1), dim3(1024), 0, s, m->scan_tmp, nt);
    hipLaunchKernelGGL(k_scan_c, dim3(nt), dim3(256), 0, s, m->cnt, n, m->scan_tmp, nt);
    hipLaunchKernelGGL((k_mlp_pairs<true>), dim3(nb), dim3(256), 0, s, c->g, c->spos, c->cell_start, n, rc2, m->cnt,
                       m->pairs, m->max_pairs, m->flag);
    MDF_LAUNCHED();
    return MDF_OK;
}

int mlp_forward(mdf_mlp *m, int n, hipStream_t s) {
    MLPArgs A = mlp_args(m, n);
    hipLaunchKernelGGL(k_mlp_forward, dim3(m->ngrid), dim3(kMLPThreads), 0, s, A);
    MDF_LAUNCHED();
    return MDF_OK;
}

// Inference inside a step: pairs, network, and the correction added to the
// sorted PP acceleration (unit charges, before the af scaling of the gather).
int mlp_correct(mdf_mlp *m, mdf_p3m_ctx *c, int n, hipStream_t s) {
    if (n <= 0) return MDF_OK;
    MDF_TRY(mlp_build_pairs(m, c, n, s));
    MDF_TRY(mlp_forward(m, n, s));
    hipLaunchKernelGGL(k_mlp_force<0>, dim3((n + 255) / 256), dim3(256), 0, s, m->pairs, m->cnt, m->g, n, m->max_pairs, c->acc_pp_s);
    MDF_LAUNCHED();
    return MDF_OK;
}

int *mlp_flag(mdf_mlp *m) { return m->flag; }

// Loss and gradient for targets (n,3) in caller order; the pair list must be built.
static int mlp_grad(mdf_mlp *m, mdf_p3m_ctx *c, const double *target, int n, hipStream_t s) {
    MDF_TRY(mlp_forward(m, n, s));
    const unsigned nb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_mlp_force<1>, dim3(nb), dim3(256), 0, s, m->pairs, m->cnt, m->g, n, m->max_pairs, m->fpred);
    hipLaunchKernelGGL(k_mlp_residual, dim3(nb), dim3(256), 0, s, m->fpred, target, c->perm, n, m->err, m->loss_part);
    hipLaunchKernelGGL(k_mlp_loss, dim3(1), dim3(256), 0, s, m->loss_part, (int)nb, 1.0 / (double)n, m->loss);
    MLPArgs A = mlp_args(m, n);
    A.coef = 2.0f / (float)n;  // d/dg of (1/n) sum_i |F_i - t_i|^2
    hipLaunchKernelGGL(k_mlp_backward, dim3(m->ngrid), dim3(kMLPThreads), 0, s, A);
    hipLaunchKernelGGL(k_mlp_grad_reduce, dim3((kNParam + 255) / 256), dim3(256), 0, s, m->gpart,
                       m->ngrid * kMLPWaves, m->grad);
    MDF_LAUNCHED();
    return MDF_OK;
}

}  // namespace mdf

using namespace mdf;

extern "C" {

int mdf_mlp_nparam(void) { return kNParam; }

int mdf_mlp_create(int n_max, int max_pairs, const mdf_mlp_meta *meta, mdf_mlp **out) {
    if (!out || !meta || n_max <= 0 || max_pairs <= 0) {
        set_error("mdf_mlp_create: bad argument");
        return MDF_EINVAL;
    }
    *out = nullptr;
    mdf_mlp *m = new (std::nothrow) mdf_mlp();
    if (!m) {
        set_error("out of host memory");
        return MDF_ENOMEM;
    }
    m->n_max = n_max;
    m->max_pairs = max_pairs;
    m->meta = *meta;
    m->ngrid = 1024;
    const int nt = (n_max + kTile - 1) / kTile;
    if (nt > 8192) {
        delete m;
        set_error("mdf_mlp_create: n_max too large");
        return MDF_EINVAL;
    }
    bool ok = hipMalloc((void **)&m->param, kNParam * sizeof(float)) == hipSuccess &&
              hipMalloc((void **)&m->adam_m, kNParam * sizeof(float)) == hipSuccess &&
              hipMalloc((void **)&m->adam_v, kNParam * sizeof(float)) == hipSuccess &&
              hipMalloc((void **)&m->grad, kNParam * sizeof(float)) == hipSuccess &&
              hipMalloc((void **)&m->gpart, (size_t)m->ngrid * kMLPWaves * kNParam * sizeof(float)) == hipSuccess &&
              hipMalloc((void **)&m->pairs, (size_t)max_pairs * sizeof(MLPPair)) == hipSuccess &&
              hipMalloc((void **)&m->g, (size_t)max_pairs * sizeof(float)) == hipSuccess &&
              hipMalloc((void **)&m->cnt, ((size_t)n_max + 1) * sizeof(int)) == hipSuccess &&
              hipMalloc((void **)&m->scan_tmp, ((size_t)nt + 1) * sizeof(int)) == hipSuccess &&
              hipMalloc((void **)&m->rho, (size_t)n_max * sizeof(float)) == hipSuccess &&
              hipMalloc((void **)&m->fpred, 3 * (size_t)n_max * sizeof(double)) == hipSuccess &&
              hipMalloc((void **)&m->err, 3 * (size_t)n_max * sizeof(float)) == hipSuccess &&
              hipMalloc((void **)&m->loss_part, ((size_t)n_max / 256 + 1) * sizeof(double)) == hipSuccess &&
              hipMalloc((void **)&m->loss, sizeof(double)) == hipSuccess &&
              hipMalloc((void **)&m->flag, sizeof(int)) == hipSuccess;
    if (ok) {
        ok = hipMemset(m->param, 0, kNParam * sizeof(float)) == hipSuccess &&
             hipMemset(m->adam_m, 0, kNParam * sizeof(float)) == hipSuccess &&
             hipMemset(m->adam_v, 0, kNParam * sizeof(float)) == hipSuccess &&
             hipMemset(m->flag, 0, sizeof(int)) == hipSuccess && hipMemset(m->cnt, 0, sizeof(int)) == hipSuccess &&
             hipDeviceSynchronize() == hipSuccess;
    }
    if (!ok) {
        mdf_mlp_destroy(m);
        set_error("mdf_mlp_create: device allocation failed");
        return MDF_ENOMEM;
    }
    *out = m;
    return MDF_OK;
}

int mdf_mlp_destroy(mdf_mlp *m) {
    if (!m) return MDF_OK;
    (void)hipDeviceSynchronize();
    void *ptrs[] = {m->param, m->adam_m, m->adam_v, m->grad, m->gpart, m->pairs, m->g, m->cnt,
                    m->scan_tmp, m->rho, m->fpred, m->err, m->loss_part, m->loss, m->flag};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    delete m;
    return MDF_OK;
}

int mdf_mlp_set_meta(mdf_mlp *m, const mdf_mlp_meta *meta) {
    if (!m || !meta) {
        set_error("mdf_mlp_set_meta: null argument");
        return MDF_EINVAL;
    }
    m->meta = *meta;
    return MDF_OK;
}

int mdf_mlp_params(mdf_mlp *m, const float *in, float *out, int reset_adam, void *stream) {
    if (!m) {
        set_error("mdf_mlp_params: null network");
        return MDF_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    if (in) MDF_HIP(hipMemcpyAsync(m->param, in, kNParam * sizeof(float), hipMemcpyDeviceToDevice, s));
    if (out) MDF_HIP(hipMemcpyAsync(out, m->param, kNParam * sizeof(float), hipMemcpyDeviceToDevice, s));
    if (reset_adam) {
        hipLaunchKernelGGL(k_fill_float_zero_mlp, dim3((kNParam + 255) / 256), dim3(256), 0, s, m->adam_m, m->adam_v);
        MDF_LAUNCHED();
        m->step = 0;
    }
    return MDF_OK;
}

int mdf_mlp_pairs(mdf_mlp *m, mdf_p3m_ctx *ctx, const double *pos, int n, int *pairs_out, float *g_out,
                  float *rho_out, int *perm_out, void *stream) {
    MDF_TRY(check_mlp(m, ctx, n));
    hipStream_t s = (hipStream_t)stream;
    MDF_TRY(bin_particles(ctx, pos, nullptr, n, s));
    MDF_TRY(mlp_build_pairs(m, ctx, n, s));
    MDF_TRY(mlp_forward(m, n, s));
    int total = 0, flag = 0;
    MDF_HIP(hipMemcpyAsync(&total, m->cnt + n, sizeof(int), hipMemcpyDeviceToHost, s));
    MDF_HIP(hipMemcpyAsync(&flag, m->flag, sizeof(int), hipMemcpyDeviceToHost, s));
    MDF_HIP(hipStreamSynchronize(s));
    if (flag) {
        MDF_TRY(fill_zero_i(m->flag, 1, s));
        set_error("mdf_mlp_pairs: %d pairs exceed max_pairs = %d", total, m->max_pairs);
        return MDF_ESTATE;
    }
    if (pairs_out) {
        MDF_HIP(hipMemcpyAsync(pairs_out, &total, sizeof(int), hipMemcpyHostToDevice, s));
        MDF_HIP(hipMemcpyAsync(pairs_out + 1, m->pairs, (size_t)total * sizeof(MLPPair), hipMemcpyDeviceToDevice, s));
    }
    if (g_out) MDF_HIP(hipMemcpyAsync(g_out, m->g, (size_t)total * sizeof(float), hipMemcpyDeviceToDevice, s));
    if (rho_out) MDF_HIP(hipMemcpyAsync(rho_out, m->rho, (size_t)n * sizeof(float), hipMemcpyDeviceToDevice, s));
    if (perm_out) MDF_HIP(hipMemcpyAsync(perm_out, ctx->perm, (size_t)n * sizeof(int), hipMemcpyDeviceToDevice, s));
    MDF_HIP(hipStreamSynchronize(s));
    return total;
}

int mdf_mlp_correction(mdf_mlp *m, mdf_p3m_ctx *ctx, const double *pos, int n, double *f_out, void *stream) {
    MDF_TRY(check_mlp(m, ctx, n));
    if (!f_out && n > 0) {
        set_error("mdf_mlp_correction: null output");
        return MDF_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    if (n <= 0) return MDF_OK;
    MDF_TRY(bin_particles(ctx, pos, nullptr, n, s));
    MDF_TRY(mlp_build_pairs(m, ctx, n, s));
    MDF_TRY(mlp_forward(m, n, s));
    hipLaunchKernelGGL(k_mlp_force<1>, dim3((n + 255) / 256), dim3(256), 0, s, m->pairs, m->cnt, m->g, n, m->max_pairs, m->fpred);
    MDF_LAUNCHED();
    MDF_TRY(unpermute3(ctx, m->fpred, f_out, n, s));
    return check_overflow(ctx, m, s);
}

int mdf_mlp_grad(mdf_mlp *m, mdf_p3m_ctx *ctx, const double *pos, const double *target, int n, float *grad_out,
                 double *loss_out, int build, void *stream) {
    MDF_TRY(check_mlp(m, ctx, n));
    if (!target || n <= 0) {
        set_error("mdf_mlp_grad: need targets and n > 0");
        return MDF_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    if (build) {
        MDF_TRY(bin_particles(ctx, pos, nullptr, n, s));
        MDF_TRY(mlp_build_pairs(m, ctx, n, s));
        MDF_TRY(check_overflow(ctx, m, s));
    }
    MDF_TRY(mlp_grad(m, ctx, target, n, s));
    if (grad_out) MDF_HIP(hipMemcpyAsync(grad_out, m->grad, kNParam * sizeof(float), hipMemcpyDeviceToDevice, s));
    if (loss_out) MDF_HIP(hipMemcpyAsync(loss_out, m->loss, sizeof(double), hipMemcpyDeviceToDevice, s));
    return MDF_OK;
}

int mdf_mlp_adam(mdf_mlp *m, float lr, float beta1, float beta2, float eps, void *stream) {
    if (!m) {
        set_error("mdf_mlp_adam: null network");
        return MDF_EINVAL;
    }
    ++m->step;
    const float c1 = 1.0f - powf(beta1, (float)m->step), c2 = 1.0f - powf(beta2, (float)m->step);
    hipLaunchKernelGGL(k_mlp_adam, dim3((kNParam + 255) / 256), dim3(256), 0, (hipStream_t)stream, m->param,
                       m->adam_m, m->adam_v, m->grad, lr, beta1, beta2, eps, c1, c2);
    MDF_LAUNCHED();
    return MDF_OK;
}

int mdf_mlp_stats(mdf_mlp *m, int n, int *out2) {
    if (!m || !out2 || n < 0 || n > m->n_max) {
        set_error("mdf_mlp_stats: bad argument");
        return MDF_EINVAL;
    }
    MDF_HIP(hipMemcpy(&out2[0], m->cnt + n, sizeof(int), hipMemcpyDeviceToHost));
    MDF_HIP(hipMemcpy(&out2[1], m->flag, sizeof(int), hipMemcpyDeviceToHost));
    return MDF_OK;
}

int mdf_p3m_attach_mlp(mdf_p3m_ctx *ctx, mdf_mlp *m) {
    if (!ctx) {
        set_error("mdf_p3m_attach_mlp: null context");
        return MDF_EINVAL;
    }
    if (m && m->n_max < ctx->prm.n_max) {
        set_error("mdf_p3m_attach_mlp: network n_max %d below the context's %d", m->n_max, ctx->prm.n_max);
        return MDF_EINVAL;
    }
    ctx->mlp = m;
    return MDF_OK;
}

}  // extern "C"
