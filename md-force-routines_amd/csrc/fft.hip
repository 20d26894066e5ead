// Mesh solve: 3D real FFT, optimal-Green multiply, PM energy,
// ik-differentiation and three inverse transforms, all hand-written.
//
// Reference: particle_mesh_fft_r, yukawa_pm_ord6.py:442-480:
//   rho_k = fftshift(fftn(rho)); phi_k = G_k rho_k;
//   U_f = 1/2 sum_k |rho_k|^2 G_k / V;
//   E_a = real(M^3 ifftn(ifftshift(-i k_a phi_k)) / V),  a = x, y, z.
// Here the spectrum is kept as the half spectrum kx in [0, Mx/2] in natural
// (unshifted) order, so the shift copies disappear and the Green function is
// stored once in the same order.  Equivalences used:
// * real(ifftn(.)) of the reference equals a Hermitian (C2R) inverse of the
//   same spectrum with the Nyquist plane of each field's own derivative axis
//   set to zero (-i k_a phi at k_a = -pi M/L has no Hermitian partner, so the
//   real part cancels it), and with the imaginary part of the kx = 0 and
//   kx = Mx/2 columns dropped;
// * the full-spectrum energy sum equals the half-spectrum sum with weight 2
//   on 0 < kx < Mx/2 and 1 on kx = 0, Mx/2.
//
// Passes (FP64 complex, radix-2 Stockham in LDS, twiddles from a table):
//   1. x, real-to-complex: two real rows packed as one complex row.
//   2. y, forward, in place on the half spectrum.
//   3. z, forward + G multiply + energy + three ik fields + three inverse z,
//      one kernel, written to three half spectra.
//   4. y, inverse, three fields.
//   5. x, complex-to-real: (Ex + iEy) of a row and (Ez of two rows) per
//      complex transform, written interleaved (Ex, Ey, Ez) per mesh node.
// The 1/V of the reference's M^3 ifftn / V is folded into step 3.
#include "mdf_internal.h"

namespace mdf {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Radix-2 Stockham transform of nl lines of length m = 2^lm stored contiguously
// in LDS (line stride m).  a holds the input, b is scratch; returns the buffer
// holding the result.  tw[t] = exp(-2 pi i t / m), t < m/2.  DIR = -1 forward
// (exp(-i...)), +1 inverse (unnormalised).  Ends with a barrier.
template <int DIR>
__device__ double2 *stockham(double2 *a, double2 *b, int nl, int m, int lm, const double2 *__restrict__ tw) {
    const int half = m >> 1;
    const int total = nl * half;
    for (int s = 0; s < lm; ++s) {
        const int Ns = 1 << s;
        for (int t = threadIdx.x; t < total; t += blockDim.x) {
            const int line = t >> (lm - 1);
            const int j = t & (half - 1);
            const double2 *src = a + line * m;
            double2 *dst = b + line * m;
            const double2 u = src[j];
            double2 v = src[j + half];
            const int k = j & (Ns - 1);
            double2 w = tw[k << (lm - 1 - s)];
            if (DIR > 0) w.y = -w.y;
            v = cmul(v, w);
            const int o = ((j - k) << 1) + k;
            dst[o] = make_double2(u.x + v.x, u.y + v.y);
            dst[o + Ns] = make_double2(u.x - v.x, u.y - v.y);
        }
        __syncthreads();
        double2 *t2 = a;
        a = b;
        b = t2;
    }
    return a;
}

// ---- pass 1: x, real to complex -------------------------------------------
// Block: nl complex lines = 2*nl consecutive real rows.
__global__ __launch_bounds__(256) void k_fft_x_r2c(const double *__restrict__ rho, double2 *__restrict__ ck, int m,
                                                   int lm, int Mxh, int nl, const double2 *__restrict__ tw) {
    extern __shared__ double2 sm[];
    double2 *a = sm, *b = sm + nl * m;
    const size_t row0 = (size_t)blockIdx.x * 2 * nl;
    for (int t = threadIdx.x; t < nl * m; t += blockDim.x) {
        const int l = t / m, x = t - l * m;
        const size_t ra = row0 + 2 * l;
        a[t] = make_double2(rho[ra * m + x], rho[(ra + 1) * m + x]);
    }
    __syncthreads();
    const double2 *r = stockham<-1>(a, b, nl, m, lm, tw);
    for (int t = threadIdx.x; t < nl * Mxh; t += blockDim.x) {
        const int l = t / Mxh, k = t - l * Mxh;
        const double2 zk = r[l * m + k];
        const double2 zc = r[l * m + ((m - k) & (m - 1))];
        const size_t ra = row0 + 2 * l;
        // A = (Z_k + conj Z_{m-k})/2, B = (Z_k - conj Z_{m-k})/(2i)
        ck[ra * Mxh + k] = make_double2(0.5 * (zk.x + zc.x), 0.5 * (zk.y - zc.y));
        ck[(ra + 1) * Mxh + k] = make_double2(0.5 * (zk.y + zc.y), -0.5 * (zk.x - zc.x));
    }
}

// ---- passes 2 and 4: complex lines along a strided axis --------------------
// Element e of line (o, kx) sits at o*os + e*es + kx; a block transforms kt
// consecutive kx of one outer index o (blockIdx.y) in field blockIdx.z.
template <int DIR>
__global__ __launch_bounds__(256) void k_fft_strided(double2 *__restrict__ data, size_t fstride, int m, int lm,
                                                     int es, int os, int Mxh, int kt,
                                                     const double2 *__restrict__ tw) {
    extern __shared__ double2 sm[];
    const int kx0 = blockIdx.x * kt;
    const int nk = min(kt, Mxh - kx0);
    double2 *a = sm, *b = sm + kt * m;
    double2 *base = data + blockIdx.z * fstride + (size_t)blockIdx.y * os + kx0;
    for (int t = threadIdx.x; t < nk * m; t += blockDim.x) {
        const int e = t / nk, kk = t - e * nk;
        a[kk * m + e] = base[(size_t)e * es + kk];
    }
    __syncthreads();
    const double2 *r = stockham<DIR>(a, b, nk, m, lm, tw);
    for (int t = threadIdx.x; t < nk * m; t += blockDim.x) {
        const int e = t / nk, kk = t - e * nk;
        base[(size_t)e * es + kk] = r[kk * m + e];
    }
}

// ---- pass 3: z forward, Green, energy, ik, z inverse ----------------------
struct GreenZArgs {
    const double2 *__restrict__ ck;
    double2 *__restrict__ ek;          // 3 half spectra
    const double *__restrict__ green;  // natural-order half spectrum
    const double2 *__restrict__ tw;
    double *__restrict__ part;         // energy partial per block
    size_t nspec;
    int M[3];
    int Mxh, lm, kt;
    double kfac[3];                    // 2 pi / L per axis
    double inv_V;
};

__device__ __forceinline__ double kval(int n, int M, double f) { return f * (double)(n < M / 2 ? n : n - M); }

__global__ __launch_bounds__(256) void k_fft_z_green(GreenZArgs P) {
    extern __shared__ double2 sm[];
    __shared__ double red[4];
    const int m = P.M[2], My = P.M[1], Mxh = P.Mxh, kt = P.kt;
    const int y = blockIdx.y;
    const int kx0 = blockIdx.x * kt;
    const int nk = min(kt, Mxh - kx0);
    const int es = My * Mxh;
    const size_t off = (size_t)y * Mxh + kx0;
    double2 *a = sm, *b = sm + kt * m, *phi = sm + 2 * kt * m;
    for (int t = threadIdx.x; t < nk * m; t += blockDim.x) {
        const int e = t / nk, kk = t - e * nk;
        a[kk * m + e] = P.ck[off + (size_t)e * es + kk];
    }
    __syncthreads();
    const double2 *r = stockham<-1>(a, b, nk, m, P.lm, P.tw);
    double esum = 0.0;
    for (int t = threadIdx.x; t < nk * m; t += blockDim.x) {
        const int kk = t / m, kz = t - kk * m;
        const int kx = kx0 + kk;
        const double gk = P.green[off + (size_t)kz * es + kk];
        const double2 v = r[t];
        const double wgt = (kx == 0 || 2 * kx == P.M[0]) ? 1.0 : 2.0;
        esum += wgt * gk * (v.x * v.x + v.y * v.y);
        phi[t] = make_double2(gk * v.x * P.inv_V, gk * v.y * P.inv_V);
    }
    esum = wave_sum(esum);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = esum;
    __syncthreads();
    if (threadIdx.x == 0) P.part[blockIdx.y * gridDim.x + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
    for (int ax = 0; ax < 3; ++ax) {
        for (int t = threadIdx.x; t < nk * m; t += blockDim.x) {
            const int kk = t / m, kz = t - kk * m;
            const int n = ax == 0 ? kx0 + kk : (ax == 1 ? y : kz);
            const int Mn = P.M[ax];
            const double k = (2 * n == Mn) ? 0.0 : kval(n, Mn, P.kfac[ax]);
            const double2 f = phi[t];
            a[t] = make_double2(k * f.y, -k * f.x);  // -i k phi
        }
        __syncthreads();
        const double2 *q = stockham<1>(a, b, nk, m, P.lm, P.tw);
        double2 *dst = P.ek + ax * P.nspec + off;
        for (int t = threadIdx.x; t < nk * m; t += blockDim.x) {
            const int e = t / nk, kk = t - e * nk;
            dst[(size_t)e * es + kk] = q[kk * m + e];
        }
        __syncthreads();
    }
}

// ---- pass 5: x, complex to real ------------------------------------------
// Block: np row pairs (rows R, R+1 of the (z, y) row list); 3 lines per pair.
__device__ __forceinline__ double2 herm_get(const double2 *__restrict__ row, int k, int m) {
    const int h = m >> 1;
    if (k == 0 || k == h) return make_double2(row[k].x, 0.0);
    if (k < h) return row[k];
    const double2 v = row[m - k];
    return make_double2(v.x, -v.y);
}

__global__ __launch_bounds__(256) void k_fft_x_c2r(const double2 *__restrict__ ek, double *__restrict__ emesh, int m,
                                                   int lm, int Mxh, int np, size_t nspec,
                                                   const double2 *__restrict__ tw) {
    extern __shared__ double2 sm[];
    const int nl = 3 * np;
    double2 *a = sm, *b = sm + nl * m;
    const size_t R0 = (size_t)blockIdx.x * 2 * np;
    for (int t = threadIdx.x; t < np * m; t += blockDim.x) {
        const int p = t / m, k = t - p * m;
        const size_t R = R0 + 2 * p;
        const double2 *ex0 = ek + R * Mxh, *ex1 = ex0 + Mxh;
        const double2 *ey0 = ek + nspec + R * Mxh, *ey1 = ey0 + Mxh;
        const double2 *ez0 = ek + 2 * nspec + R * Mxh, *ez1 = ez0 + Mxh;
        double2 X = herm_get(ex0, k, m), Y = herm_get(ey0, k, m);
        a[(3 * p + 0) * m + k] = make_double2(X.x - Y.y, X.y + Y.x);
        X = herm_get(ex1, k, m);
        Y = herm_get(ey1, k, m);
        a[(3 * p + 1) * m + k] = make_double2(X.x - Y.y, X.y + Y.x);
        X = herm_get(ez0, k, m);
        Y = herm_get(ez1, k, m);
        a[(3 * p + 2) * m + k] = make_double2(X.x - Y.y, X.y + Y.x);
    }
    __syncthreads();
    const double2 *r = stockham<1>(a, b, nl, m, lm, tw);
    for (int t = threadIdx.x; t < np * m; t += blockDim.x) {
        const int p = t / m, x = t - p * m;
        const size_t R = R0 + 2 * p;
        const double2 l0 = r[(3 * p + 0) * m + x];
        const double2 l1 = r[(3 * p + 1) * m + x];
        const double2 l2 = r[(3 * p + 2) * m + x];
        double *n0 = emesh + 3 * (R * m + x);
        double *n1 = emesh + 3 * ((R + 1) * m + x);
        n0[0] = l0.x;
        n0[1] = l0.y;
        n0[2] = l2.x;
        n1[0] = l1.x;
        n1[1] = l1.y;
        n1[2] = l2.y;
    }
}

static int ilog2(int m) {
    int l = 0;
    while ((1 << l) < m) ++l;
    return l;
}

int pm_solve(mdf_p3m_ctx *c, hipStream_t s) {
    const int Mx = c->g.M[0], My = c->g.M[1], Mz = c->g.M[2], Mxh = c->g.Mxh;
    const int lx = ilog2(Mx), ly = ilog2(My), lz = ilog2(Mz);
    // pass 1
    {
        int nl = 512 / Mx;
        if (nl < 1) nl = 1;
        while (nl > 1 && (My * Mz) % (2 * nl)) nl >>= 1;
        const size_t lds = 2 * sizeof(double2) * nl * Mx;
        hipLaunchKernelGGL(k_fft_x_r2c, dim3((My * Mz) / (2 * nl)), dim3(256), lds, s, c->rho, c->ck, Mx, lx, Mxh,
                           nl, c->tw[0]);
        MDF_LAUNCHED();
    }
    // pass 2: y forward
    const int kty = My >= 512 ? 4 : (My >= 256 ? 8 : 16);
    {
        const size_t lds = 2 * sizeof(double2) * kty * My;
        dim3 grid((Mxh + kty - 1) / kty, Mz, 1);
        hipLaunchKernelGGL(k_fft_strided<-1>, grid, dim3(256), lds, s, c->ck, (size_t)0, My, ly, Mxh, My * Mxh, Mxh,
                           kty, c->tw[1]);
        MDF_LAUNCHED();
    }
    // pass 3: z + Green + ik
    {
        GreenZArgs P;
        P.ck = c->ck;
        P.ek = c->ek;
        P.green = c->green;
        P.tw = c->tw[2];
        P.part = c->part_pm;
        P.nspec = c->nspec;
        P.M[0] = Mx;
        P.M[1] = My;
        P.M[2] = Mz;
        P.Mxh = Mxh;
        P.lm = lz;
        P.kt = Mz >= 512 ? 2 : (Mz >= 256 ? 4 : 8);
        for (int d = 0; d < 3; ++d) P.kfac[d] = 2.0 * M_PI / c->g.L[d];
        P.inv_V = 1.0 / (c->g.L[0] * c->g.L[1] * c->g.L[2]);
        const size_t lds = 3 * sizeof(double2) * P.kt * Mz;
        dim3 grid((Mxh + P.kt - 1) / P.kt, My, 1);
        c->n_part_pm = grid.x * grid.y;
        hipLaunchKernelGGL(k_fft_z_green, grid, dim3(256), lds, s, P);
        MDF_LAUNCHED();
    }
    // pass 4: y inverse, 3 fields
    {
        const size_t lds = 2 * sizeof(double2) * kty * My;
        dim3 grid((Mxh + kty - 1) / kty, Mz, 3);
        hipLaunchKernelGGL(k_fft_strided<1>, grid, dim3(256), lds, s, c->ek, c->nspec, My, ly, Mxh, My * Mxh, Mxh,
                           kty, c->tw[1]);
        MDF_LAUNCHED();
    }
    // pass 5: x complex to real
    {
        int np = 256 / Mx;
        if (np < 1) np = 1;
        while (np > 1 && (My * Mz) % (2 * np)) np >>= 1;
        const size_t lds = 2 * sizeof(double2) * 3 * np * Mx;
        hipLaunchKernelGGL(k_fft_x_c2r, dim3((My * Mz) / (2 * np)), dim3(256), lds, s, c->ek, c->emesh, Mx, lx, Mxh,
                           np, c->nspec, c->tw[0]);
        MDF_LAUNCHED();
    }
    return MDF_OK;
}

}  // namespace mdf
