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
// Passes (complex FP64, or FP32 in the mixed mode; radix-4/2 Stockham in LDS,
// twiddles from a table; the PM energy is always summed in FP64):
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

// Complex type of a real type T (double2 / float2).
template <class T>
struct Cx;
template <>
struct Cx<double> {
    using t = double2;
    static __device__ __forceinline__ double2 make(double a, double b) { return make_double2(a, b); }
};
template <>
struct Cx<float> {
    using t = float2;
    static __device__ __forceinline__ float2 make(float a, float b) { return make_float2(a, b); }
};

template <class C>
__device__ __forceinline__ C cmul(C a, C b) {
    C r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}

// Stockham transform of nl lines of length m = 2^lm stored contiguously in
// LDS (line stride m): radix-4 stages, then one radix-2 stage when lm is odd
// (each stage reads and writes every element once, behind a barrier, so
// radix 4 halves the LDS passes and barriers of a radix-2 transform).  a
// holds the input, b is scratch; returns the buffer holding the result.
// tw[t] = exp(-2 pi i t / m), t < m/2.  DIR = -1 forward (exp(-i...)), +1
// inverse (unnormalised).  A stage of span Ns (the product of the radices
// before it) takes element j of each of the R quarter (half) lines, turns
// them by exp(DIR 2 pi i r k / (R Ns)), k = j mod Ns, applies the R-point
// DFT and writes them to (j - k) R + k + r Ns.  Ends with a barrier.
template <int DIR, class C>
__device__ C *stockham(C *a, C *b, int nl, int m, int lm, const C *__restrict__ tw) {
    int s = 0;  // log2 of the span Ns
    if (lm >= 2) {
        const int q = m >> 2;
        const int total = nl * q;
        for (; s + 2 <= lm; s += 2) {
            const int Ns = 1 << s;
            const int st = lm - 2 - s;  // m / (4 Ns) = 2^st
            for (int t = threadIdx.x; t < total; t += blockDim.x) {
                const int line = t >> (lm - 2);
                const int j = t & (q - 1);
                const C *src = a + line * m;
                C *dst = b + line * m;
                const int k = j & (Ns - 1);
                C a0 = src[j], a1 = src[j + q], a2 = src[j + 2 * q], a3 = src[j + 3 * q];
                if (s > 0) {
                    C w1 = tw[k << st], w2 = tw[(2 * k) << st];  // indices < m/2
                    if (DIR > 0) {
                        w1.y = -w1.y;
                        w2.y = -w2.y;
                    }
                    const C w3 = cmul(w1, w2);
                    a1 = cmul(a1, w1);
                    a2 = cmul(a2, w2);
                    a3 = cmul(a3, w3);
                }
                C b0, b1, b2, b3;
                b0.x = a0.x + a2.x;
                b0.y = a0.y + a2.y;
                b1.x = a0.x - a2.x;
                b1.y = a0.y - a2.y;
                b2.x = a1.x + a3.x;
                b2.y = a1.y + a3.y;
                b3.x = a1.x - a3.x;
                b3.y = a1.y - a3.y;
                const int o = ((j - k) << 2) + k;
                C y;
                y.x = b0.x + b2.x;
                y.y = b0.y + b2.y;
                dst[o] = y;
                y.x = b0.x - b2.x;
                y.y = b0.y - b2.y;
                dst[o + 2 * Ns] = y;
                // forward: y1 = b1 - i b3, y3 = b1 + i b3; inverse the other way round
                C yp, ym;
                yp.x = b1.x - b3.y;  // b1 + i b3
                yp.y = b1.y + b3.x;
                ym.x = b1.x + b3.y;  // b1 - i b3
                ym.y = b1.y - b3.x;
                dst[o + Ns] = DIR < 0 ? ym : yp;
                dst[o + 3 * Ns] = DIR < 0 ? yp : ym;
            }
            __syncthreads();
            C *t2 = a;
            a = b;
            b = t2;
        }
    }
    if (s < lm) {  // last stage, radix 2 (span m/2)
        const int half = m >> 1;
        const int total = nl * half;
        const int Ns = 1 << s;
        for (int t = threadIdx.x; t < total; t += blockDim.x) {
            const int line = t >> (lm - 1);
            const int j = t & (half - 1);
            const C *src = a + line * m;
            C *dst = b + line * m;
            const C u = src[j];
            C v = src[j + half];
            const int k = j & (Ns - 1);
            C w = tw[k << (lm - 1 - s)];
            if (DIR > 0) w.y = -w.y;
            v = cmul(v, w);
            const int o = ((j - k) << 1) + k;
            C p, r;
            p.x = u.x + v.x;
            p.y = u.y + v.y;
            r.x = u.x - v.x;
            r.y = u.y - v.y;
            dst[o] = p;
            dst[o + Ns] = r;
        }
        __syncthreads();
        C *t2 = a;
        a = b;
        b = t2;
    }
    return a;
}

// ---- pass 1: x, real to complex -------------------------------------------
// Block: nl complex lines = 2*nl consecutive real rows.  The real input is
// always the FP64 charge mesh; T is the arithmetic of the transform.
template <class T>
__device__ __forceinline__ void r2c_body(const double *__restrict__ rho, typename Cx<T>::t *__restrict__ ck, int m,
                                         int lm, int Mxh, int nl, const typename Cx<T>::t *__restrict__ tw, int bx,
                                         unsigned char *smraw) {
    using C = typename Cx<T>::t;
    C *sm = (C *)smraw;
    C *a = sm, *b = sm + nl * m;
    const size_t row0 = (size_t)bx * 2 * nl;
    for (int t = threadIdx.x; t < nl * m; t += blockDim.x) {
        const int l = t / m, x = t - l * m;
        const size_t ra = row0 + 2 * l;
        a[t] = Cx<T>::make((T)rho[ra * m + x], (T)rho[(ra + 1) * m + x]);
    }
    __syncthreads();
    const C *r = stockham<-1>(a, b, nl, m, lm, tw);
    for (int t = threadIdx.x; t < nl * Mxh; t += blockDim.x) {
        const int l = t / Mxh, k = t - l * Mxh;
        const C zk = r[l * m + k];
        const C zc = r[l * m + ((m - k) & (m - 1))];
        const size_t ra = row0 + 2 * l;
        // A = (Z_k + conj Z_{m-k})/2, B = (Z_k - conj Z_{m-k})/(2i)
        ck[ra * Mxh + k] = Cx<T>::make((T)0.5 * (zk.x + zc.x), (T)0.5 * (zk.y - zc.y));
        ck[(ra + 1) * Mxh + k] = Cx<T>::make((T)0.5 * (zk.y + zc.y), (T)-0.5 * (zk.x - zc.x));
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_fft_x_r2c(const double *__restrict__ rho, typename Cx<T>::t *__restrict__ ck,
                                                   int m, int lm, int Mxh, int nl,
                                                   const typename Cx<T>::t *__restrict__ tw) {
    extern __shared__ __align__(16) unsigned char smraw[];
    r2c_body<T>(rho, ck, m, lm, Mxh, nl, tw, blockIdx.x, smraw);
}

// ---- passes 2 and 4: complex lines along a strided axis --------------------
// Element e of line (o, kx) sits at o*os + e*es + kx; a block transforms kt
// consecutive kx of one outer index o (blockIdx.y) in field blockIdx.z.
template <int DIR, class C>
__device__ __forceinline__ void strided_body(C *__restrict__ data, size_t fstride, int m, int lm, int es, int os,
                                             int Mxh, int kt, const C *__restrict__ tw, int bx, int by, int bz,
                                             unsigned char *smraw) {
    C *sm = (C *)smraw;
    const int kx0 = bx * kt;
    const int nk = min(kt, Mxh - kx0);
    C *a = sm, *b = sm + kt * m;
    C *base = data + bz * fstride + (size_t)by * os + kx0;
    for (int t = threadIdx.x; t < nk * m; t += blockDim.x) {
        const int e = t / nk, kk = t - e * nk;
        a[kk * m + e] = base[(size_t)e * es + kk];
    }
    __syncthreads();
    const C *r = stockham<DIR>(a, b, nk, m, lm, tw);
    for (int t = threadIdx.x; t < nk * m; t += blockDim.x) {
        const int e = t / nk, kk = t - e * nk;
        base[(size_t)e * es + kk] = r[kk * m + e];
    }
}

template <int DIR, class C>
__global__ __launch_bounds__(256) void k_fft_strided(C *__restrict__ data, size_t fstride, int m, int lm, int es, int os,
                                                     int Mxh, int kt, const C *__restrict__ tw) {
    extern __shared__ __align__(16) unsigned char smraw[];
    strided_body<DIR, C>(data, fstride, m, lm, es, os, Mxh, kt, tw, blockIdx.x, blockIdx.y, blockIdx.z, smraw);
}

// ---- pass 3: z forward, Green, energy, ik, z inverse ----------------------
// The energy partial is accumulated in FP64 in both arithmetics.
template <class T>
struct GreenZArgs {
    using C = typename Cx<T>::t;
    const C *__restrict__ ck;
    C *__restrict__ ek;                // 3 half spectra
    const T *__restrict__ green;       // natural-order half spectrum
    const C *__restrict__ tw;
    double *__restrict__ part;         // energy partial per block
    size_t nspec;
    int M[3];
    int Mxh, lm, kt;
    double kfac[3];                    // 2 pi / L per axis
    double inv_V;
};

__device__ __forceinline__ double kval(int n, int M, double f) { return f * (double)(n < M / 2 ? n : n - M); }

template <class T>
__device__ __forceinline__ void zgreen_body(const GreenZArgs<T> &P, int bx, int by, int nbx, unsigned char *smraw,
                                            double *red) {
    using C = typename Cx<T>::t;
    C *sm = (C *)smraw;
    const int m = P.M[2], My = P.M[1], Mxh = P.Mxh, kt = P.kt;
    const int y = by;
    const int kx0 = bx * kt;
    const int nk = min(kt, Mxh - kx0);
    const int es = My * Mxh;
    const size_t off = (size_t)y * Mxh + kx0;
    C *a = sm, *b = sm + kt * m, *phi = sm + 2 * kt * m;
    for (int t = threadIdx.x; t < nk * m; t += blockDim.x) {
        const int e = t / nk, kk = t - e * nk;
        a[kk * m + e] = P.ck[off + (size_t)e * es + kk];
    }
    __syncthreads();
    const C *r = stockham<-1>(a, b, nk, m, P.lm, P.tw);
    double esum = 0.0;
    const T inv_V = (T)P.inv_V;
    for (int t = threadIdx.x; t < nk * m; t += blockDim.x) {
        const int kk = t / m, kz = t - kk * m;
        const int kx = kx0 + kk;
        const T gk = P.green[off + (size_t)kz * es + kk];
        const C v = r[t];
        const double wgt = (kx == 0 || 2 * kx == P.M[0]) ? 1.0 : 2.0;
        esum += wgt * (double)gk * ((double)v.x * (double)v.x + (double)v.y * (double)v.y);
        phi[t] = Cx<T>::make(gk * v.x * inv_V, gk * v.y * inv_V);
    }
    esum = wave_sum(esum);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = esum;
    __syncthreads();
    if (threadIdx.x == 0) P.part[by * nbx + bx] = red[0] + red[1] + red[2] + red[3];
    for (int ax = 0; ax < 3; ++ax) {
        for (int t = threadIdx.x; t < nk * m; t += blockDim.x) {
            const int kk = t / m, kz = t - kk * m;
            const int n = ax == 0 ? kx0 + kk : (ax == 1 ? y : kz);
            const int Mn = P.M[ax];
            const T k = (T)((2 * n == Mn) ? 0.0 : kval(n, Mn, P.kfac[ax]));
            const C f = phi[t];
            a[t] = Cx<T>::make(k * f.y, -k * f.x);  // -i k phi
        }
        __syncthreads();
        const C *q = stockham<1>(a, b, nk, m, P.lm, P.tw);
        C *dst = P.ek + ax * P.nspec + off;
        for (int t = threadIdx.x; t < nk * m; t += blockDim.x) {
            const int e = t / nk, kk = t - e * nk;
            dst[(size_t)e * es + kk] = q[kk * m + e];
        }
        __syncthreads();
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_fft_z_green(GreenZArgs<T> P) {
    extern __shared__ __align__(16) unsigned char smraw[];
    __shared__ double red[4];
    zgreen_body<T>(P, blockIdx.x, blockIdx.y, gridDim.x, smraw, red);
}

// ---- pass 5: x, complex to real ------------------------------------------
// Block: np row pairs (rows R, R+1 of the (z, y) row list); 3 lines per pair.
// Output: interleaved (Ex, Ey, Ez) per node in arithmetic T.
template <class C>
__device__ __forceinline__ C herm_get(const C *__restrict__ row, int k, int m) {
    const int h = m >> 1;
    C v;
    if (k == 0 || k == h) {
        v = row[k];
        v.y = 0;
        return v;
    }
    if (k < h) return row[k];
    v = row[m - k];
    v.y = -v.y;
    return v;
}

template <class T>
__device__ __forceinline__ void c2r_body(const typename Cx<T>::t *__restrict__ ek, T *__restrict__ emesh, int m, int lm,
                                         int Mxh, int np, size_t nspec, const typename Cx<T>::t *__restrict__ tw,
                                         int bx, unsigned char *smraw) {
    using C = typename Cx<T>::t;
    C *sm = (C *)smraw;
    const int nl = 3 * np;
    C *a = sm, *b = sm + nl * m;
    const size_t R0 = (size_t)bx * 2 * np;
    for (int t = threadIdx.x; t < np * m; t += blockDim.x) {
        const int p = t / m, k = t - p * m;
        const size_t R = R0 + 2 * p;
        const C *ex0 = ek + R * Mxh, *ex1 = ex0 + Mxh;
        const C *ey0 = ek + nspec + R * Mxh, *ey1 = ey0 + Mxh;
        const C *ez0 = ek + 2 * nspec + R * Mxh, *ez1 = ez0 + Mxh;
        C X = herm_get(ex0, k, m), Y = herm_get(ey0, k, m);
        a[(3 * p + 0) * m + k] = Cx<T>::make(X.x - Y.y, X.y + Y.x);
        X = herm_get(ex1, k, m);
        Y = herm_get(ey1, k, m);
        a[(3 * p + 1) * m + k] = Cx<T>::make(X.x - Y.y, X.y + Y.x);
        X = herm_get(ez0, k, m);
        Y = herm_get(ez1, k, m);
        a[(3 * p + 2) * m + k] = Cx<T>::make(X.x - Y.y, X.y + Y.x);
    }
    __syncthreads();
    const C *r = stockham<1>(a, b, nl, m, lm, tw);
    for (int t = threadIdx.x; t < np * m; t += blockDim.x) {
        const int p = t / m, x = t - p * m;
        const size_t R = R0 + 2 * p;
        const C l0 = r[(3 * p + 0) * m + x];
        const C l1 = r[(3 * p + 1) * m + x];
        const C l2 = r[(3 * p + 2) * m + x];
        T *n0 = emesh + 3 * (R * m + x);
        T *n1 = emesh + 3 * ((R + 1) * m + x);
        n0[0] = l0.x;
        n0[1] = l0.y;
        n0[2] = l2.x;
        n1[0] = l1.x;
        n1[1] = l1.y;
        n1[2] = l2.y;
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_fft_x_c2r(const typename Cx<T>::t *__restrict__ ek, T *__restrict__ emesh,
                                                   int m, int lm, int Mxh, int np, size_t nspec,
                                                   const typename Cx<T>::t *__restrict__ tw) {
    extern __shared__ __align__(16) unsigned char smraw[];
    c2r_body<T>(ek, emesh, m, lm, Mxh, np, nspec, tw, blockIdx.x, smraw);
}

// ---- register radix-8 transforms of strided lines (passes 2, 3, 4) --------
// A line of m = 2^lm >= 8 elements is held by P = m/8 threads, eight
// elements each: thread p holds elements p + kP, k = 0..7.  The transform runs
// as Stockham stages of radix 8 (then one of radix 4 or 2 when 3 does not
// divide lm).  A stage is one set of butterflies per thread on its registers
// (radix 8: one; radix 4: two, on registers k even and odd; radix 2: four, on
// registers c and c + 4); the stages exchange through LDS, where a block's KT
// lines sit interleaved (element e of line l at e * KT + l, so that the reads
// of p + kP are consecutive across the lines of a wave).  Every stage reads
// positions p + kP, and the last stage's outputs land on those same
// positions, so the first stage takes its input from registers and the last
// leaves its output in them: two LDS exchanges for m = 512 against the six
// reads and writes of the LDS Stockham transform, and one buffer of m KT
// elements instead of two (16 columns a block at 512 against 4 there).
// Stage of radix R and span Ns: butterfly j, k = j mod Ns, input x_q =
// x[j + q m/R] turned by exp(DIR 2 pi i q k / (R Ns)), R-point DFT, output r
// to (j - k) R + k + r Ns, as the LDS transform above.

// exp(DIR 2 pi i t / m), 0 <= t < m, from the forward table of the first m/2
template <int DIR, class C>
__device__ __forceinline__ C twd(const C *__restrict__ tw, int t, int half) {
    C w = tw[t < half ? t : t - half];
    if (t >= half) {
        w.x = -w.x;
        w.y = -w.y;
    }
    if (DIR > 0) w.y = -w.y;
    return w;
}
template <class C>
__device__ __forceinline__ C cadd(C a, C b) {
    a.x += b.x;
    a.y += b.y;
    return a;
}
template <class C>
__device__ __forceinline__ C csub(C a, C b) {
    a.x -= b.x;
    a.y -= b.y;
    return a;
}
// a times DIR i
template <int DIR, class C>
__device__ __forceinline__ C muli(C a) {
    C r;
    if (DIR < 0) {
        r.x = a.y;
        r.y = -a.x;
    } else {
        r.x = -a.y;
        r.y = a.x;
    }
    return r;
}
// 4-point DFT in place, natural order
template <int DIR, class C>
__device__ __forceinline__ void dft4(C &x0, C &x1, C &x2, C &x3) {
    const C c0 = cadd(x0, x2), c1 = csub(x0, x2), c2 = cadd(x1, x3), c3 = muli<DIR>(csub(x1, x3));
    x0 = cadd(c0, c2);
    x1 = cadd(c1, c3);
    x2 = csub(c0, c2);
    x3 = csub(c1, c3);
}
// 8-point DFT in place, natural order: even outputs the DFT4 of x_k + x_{k+4},
// odd ones that of (x_k - x_{k+4}) exp(DIR 2 pi i k / 8)
template <int DIR, class C>
__device__ __forceinline__ void dft8(C *x) {
    using T = decltype(x[0].x);
    const T r2 = (T)0.70710678118654752440;
    C a0 = cadd(x[0], x[4]), a1 = cadd(x[1], x[5]), a2 = cadd(x[2], x[6]), a3 = cadd(x[3], x[7]);
    C b0 = csub(x[0], x[4]), b1 = csub(x[1], x[5]), b2 = csub(x[2], x[6]), b3 = csub(x[3], x[7]);
    const C i1 = muli<DIR>(b1), i3 = muli<DIR>(b3);
    b1.x = r2 * (b1.x + i1.x);
    b1.y = r2 * (b1.y + i1.y);
    b2 = muli<DIR>(b2);
    b3.x = r2 * (i3.x - b3.x);
    b3.y = r2 * (i3.y - b3.y);
    dft4<DIR>(a0, a1, a2, a3);
    dft4<DIR>(b0, b1, b2, b3);
    x[0] = a0;
    x[1] = b0;
    x[2] = a1;
    x[3] = b1;
    x[4] = a2;
    x[5] = b2;
    x[6] = a3;
    x[7] = b3;
}

// LDS index of element e of a thread's line: interleaved lines (e * KT + l)
// for the column passes, whose threads run line-fastest, and consecutive
// lines of stride m + 16 for the row passes, whose threads run
// element-fastest (the 16-element pad puts the lines of one wave on
// different banks).
struct ColIdx {
    int l, KT;
    __device__ __forceinline__ int operator()(int e) const { return e * KT + l; }
};
struct RowIdx {
    int base;
    __device__ __forceinline__ int operator()(int e) const { return base + e; }
};

// The transform of one line (all threads of the block call it: it holds
// barriers).  v: this thread's elements p + kP, in and out.
template <int DIR, class C, class Idx>
__device__ void reg_fft(C *v, int lm, int p, Idx at, C *lds, const C *__restrict__ tw) {
    const int m = 1 << lm, P = m >> 3, half = m >> 1;
    const int n8 = lm / 3, rem = lm - 3 * n8;
    const int nst = n8 + (rem ? 1 : 0);
    int s = 0;  // log2 of the span Ns
    for (int st = 0; st < nst; ++st) {
        const int R = st < n8 ? 8 : (rem == 2 ? 4 : 2);
        const bool last = st == nst - 1;
        const int Ns = 1 << s;
        if (st > 0) {
#pragma unroll
            for (int k = 0; k < 8; ++k) v[k] = lds[at(p + k * P)];
        }
        if (R == 8) {
            const int j = p, k = j & (Ns - 1);
            if (s > 0) {
                const int t1 = k << (lm - 3 - s);
                const C w1 = twd<DIR>(tw, t1, half), w2 = twd<DIR>(tw, 2 * t1, half), w4 = twd<DIR>(tw, 4 * t1, half);
                const C w3 = cmul(w1, w2);
                v[1] = cmul(v[1], w1);
                v[2] = cmul(v[2], w2);
                v[3] = cmul(v[3], w3);
                v[4] = cmul(v[4], w4);
                v[5] = cmul(v[5], cmul(w1, w4));
                v[6] = cmul(v[6], cmul(w2, w4));
                v[7] = cmul(v[7], cmul(w3, w4));
            }
            dft8<DIR>(v);
            if (!last) {
                __syncthreads();  // every thread has read this stage's inputs
#pragma unroll
                for (int r = 0; r < 8; ++r) lds[at((j - k) * 8 + k + r * Ns)] = v[r];
                __syncthreads();
            }
        } else if (R == 4) {
            int o[2];
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const int j = p + c * P, k = j & (Ns - 1);
                o[c] = (j - k) * 4 + k;
                if (s > 0) {
                    const int t1 = k << (lm - 2 - s);
                    const C w1 = twd<DIR>(tw, t1, half), w2 = twd<DIR>(tw, 2 * t1, half);
                    v[2 + c] = cmul(v[2 + c], w1);
                    v[4 + c] = cmul(v[4 + c], w2);
                    v[6 + c] = cmul(v[6 + c], cmul(w1, w2));
                }
                dft4<DIR>(v[c], v[2 + c], v[4 + c], v[6 + c]);
            }
            if (!last) {
                __syncthreads();
#pragma unroll
                for (int c = 0; c < 2; ++c)
#pragma unroll
                    for (int r = 0; r < 4; ++r) lds[at(o[c] + r * Ns)] = v[2 * r + c];
                __syncthreads();
            }
        } else {
            int o[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int j = p + c * P, k = j & (Ns - 1);
                o[c] = (j - k) * 2 + k;
                if (s > 0) v[c + 4] = cmul(v[c + 4], twd<DIR>(tw, k << (lm - 1 - s), half));
                const C a = v[c], b = v[c + 4];
                v[c] = cadd(a, b);
                v[c + 4] = csub(a, b);
            }
            if (!last) {
                __syncthreads();
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    lds[at(o[c])] = v[c];
                    lds[at(o[c] + Ns)] = v[c + 4];
                }
                __syncthreads();
            }
        }
        s += R == 8 ? 3 : (R == 4 ? 2 : 1);
    }
}

// Passes 2 and 4: lines along a strided axis.  Element e of line (o, kx) at
// o*os + e*es + kx; a block takes KT consecutive kx of outer index o
// (blockIdx.y) in field blockIdx.z; blockDim = KT * m/8 (at most 1024 in
// FP32, 512 in FP64, whose registers would not fit 1024 threads).
template <int DIR, class C>
__global__ __launch_bounds__(sizeof(C) == 8 ? 1024 : 512) void k_fft_col(C *__restrict__ data, size_t fstride, int lm, int es, int os, int Mxh,
                                                  int KT, const C *__restrict__ tw) {
    extern __shared__ __align__(16) unsigned char smraw[];
    C *lds = (C *)smraw;
    const int l = threadIdx.x % KT, p = threadIdx.x / KT;
    const int P = (1 << lm) >> 3;
    const int kx = blockIdx.x * KT + l;
    const bool ok = kx < Mxh;
    C *base = data + blockIdx.z * fstride + (size_t)blockIdx.y * os + (ok ? kx : 0);
    C v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        v[k] = ok ? base[(size_t)(p + k * P) * es] : C{};
    }
    reg_fft<DIR>(v, lm, p, ColIdx{l, KT}, lds, tw);
    if (ok) {
#pragma unroll
        for (int k = 0; k < 8; ++k) base[(size_t)(p + k * P) * es] = v[k];
    }
}

// Pass 3 with register transforms: z forward, G, energy, three ik fields and
// three inverse z transforms; a block takes KT consecutive kx of one y.
template <class T>
__global__ __launch_bounds__(sizeof(T) == 4 ? 1024 : 512) void k_fft_zgreen_r(GreenZArgs<T> P) {
    using C = typename Cx<T>::t;
    extern __shared__ __align__(16) unsigned char smraw[];
    __shared__ double red[16];
    C *lds = (C *)smraw;
    const int KT = P.kt, lm = P.lm;
    const int l = threadIdx.x % KT, p = threadIdx.x / KT;
    const int Pn = (1 << lm) >> 3;
    const int y = blockIdx.y;
    const int kx = blockIdx.x * KT + l;
    const bool ok = kx < P.Mxh;
    const size_t es = (size_t)P.M[1] * P.Mxh;
    const size_t off = (size_t)y * P.Mxh + (ok ? kx : 0);
    C v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = ok ? P.ck[off + (size_t)(p + k * Pn) * es] : C{};
    reg_fft<-1>(v, lm, p, ColIdx{l, KT}, lds, P.tw);
    double esum = 0.0;
    const T inv_V = (T)P.inv_V;
    const double wgt = (kx == 0 || 2 * kx == P.M[0]) ? 1.0 : 2.0;
    C phi[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int kz = p + k * Pn;
        const T gk = ok ? P.green[off + (size_t)kz * es] : (T)0;
        esum += wgt * (double)gk * ((double)v[k].x * (double)v[k].x + (double)v[k].y * (double)v[k].y);
        phi[k] = Cx<T>::make(gk * v[k].x * inv_V, gk * v[k].y * inv_V);
    }
    esum = wave_sum(esum);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = esum;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)((blockDim.x + 63) >> 6); ++w) t += red[w];
        P.part[blockIdx.y * gridDim.x + blockIdx.x] = t;
    }
    for (int ax = 0; ax < 3; ++ax) {
        const int Mn = P.M[ax];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int kz = p + k * Pn;
            const int n = ax == 0 ? kx : (ax == 1 ? y : kz);
            const T kk = (T)((2 * n == Mn) ? 0.0 : kval(n, Mn, P.kfac[ax]));
            v[k] = Cx<T>::make(kk * phi[k].y, -kk * phi[k].x);  // -i k phi
        }
        reg_fft<1>(v, lm, p, ColIdx{l, KT}, lds, P.tw);
        if (ok) {
            C *dst = P.ek + ax * P.nspec + off;
#pragma unroll
            for (int k = 0; k < 8; ++k) dst[(size_t)(p + k * Pn) * es] = v[k];
        }
    }
}

// Pass 1 with register transforms: rows 2q, 2q+1 of the FP64 charge mesh
// as one complex line (blockDim = KT * m/8, line-major threads), forward,
// then through LDS to the Hermitian split of the two rows' half spectra.
template <class T>
__global__ __launch_bounds__(sizeof(T) == 4 ? 1024 : 512) void k_fft_row_r2c(const double *__restrict__ rho,
                                                                                typename Cx<T>::t *__restrict__ ck,
                                                                                int lm, int Mxh, int KT, int npair,
                                                                                const typename Cx<T>::t *__restrict__ tw) {
    using C = typename Cx<T>::t;
    extern __shared__ __align__(16) unsigned char smraw[];
    C *lds = (C *)smraw;
    const int m = 1 << lm, P = m >> 3;
    const int l = threadIdx.x / P, p = threadIdx.x - l * P;
    const int q = blockIdx.x * KT + l;
    const bool ok = q < npair;
    const size_t ra = 2 * (size_t)(ok ? q : 0);
    C v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int x = p + k * P;
        v[k] = ok ? Cx<T>::make((T)rho[ra * m + x], (T)rho[(ra + 1) * m + x]) : C{};
    }
    const RowIdx at{l * (m + 16)};
    reg_fft<-1>(v, lm, p, at, lds, tw);
    __syncthreads();  // every thread has read the last stage's inputs
#pragma unroll
    for (int k = 0; k < 8; ++k) lds[at(p + k * P)] = v[k];
    __syncthreads();
    if (!ok) return;
    for (int k = p; k < Mxh; k += P) {
        const C zk = lds[at(k)];
        const C zc = lds[at((m - k) & (m - 1))];
        // A = (Z_k + conj Z_{m-k})/2, B = (Z_k - conj Z_{m-k})/(2i)
        ck[ra * Mxh + k] = Cx<T>::make((T)0.5 * (zk.x + zc.x), (T)0.5 * (zk.y - zc.y));
        ck[(ra + 1) * Mxh + k] = Cx<T>::make((T)0.5 * (zk.y + zc.y), (T)-0.5 * (zk.x - zc.x));
    }
}

// Pass 5 with register transforms: per row pair (R, R+1) three complex lines
// (Ex_R + i Ey_R, Ex_R+1 + i Ey_R+1, Ez_R + i Ez_R+1), each element read
// straight from the half spectra, inverse, written to the interleaved
// (Ex, Ey, Ez) mesh: the line's real and imaginary parts go to their
// components (blockDim = 3 NP * m/8, line-major threads).
template <class T>
__global__ __launch_bounds__(sizeof(T) == 4 ? 1024 : 512) void k_fft_row_c2r(const typename Cx<T>::t *__restrict__ ek,
                                                                                T *__restrict__ emesh, int lm, int Mxh,
                                                                                int NP, int npair, size_t nspec,
                                                                                const typename Cx<T>::t *__restrict__ tw) {
    using C = typename Cx<T>::t;
    extern __shared__ __align__(16) unsigned char smraw[];
    C *lds = (C *)smraw;
    const int m = 1 << lm, P = m >> 3;
    const int L = threadIdx.x / P, p = threadIdx.x - L * P;
    const int pr = L / 3, f = L - 3 * pr;  // row pair in the block, line kind
    const int q = blockIdx.x * NP + pr;
    const bool ok = q < npair;
    const size_t R = 2 * (size_t)(ok ? q : 0);
    // line f = 0: Ex_R + i Ey_R; 1: Ex_R+1 + i Ey_R+1; 2: Ez_R + i Ez_R+1
    const C *a = f < 2 ? ek + (R + f) * Mxh : ek + 2 * nspec + R * Mxh;
    const C *b = f < 2 ? ek + nspec + (R + f) * Mxh : ek + 2 * nspec + (R + 1) * Mxh;
    C v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int x = p + k * P;
        if (ok) {
            const C X = herm_get(a, x, m), Y = herm_get(b, x, m);
            v[k] = Cx<T>::make(X.x - Y.y, X.y + Y.x);
        } else {
            v[k] = C{};
        }
    }
    reg_fft<1>(v, lm, p, RowIdx{L * (m + 16)}, lds, tw);
    if (!ok) return;
    T *n0 = emesh + 3 * (R * m), *n1 = emesh + 3 * ((R + 1) * m);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int x = p + k * P;
        if (f == 0) {
            n0[3 * x + 0] = v[k].x;
            n0[3 * x + 1] = v[k].y;
        } else if (f == 1) {
            n1[3 * x + 0] = v[k].x;
            n1[3 * x + 1] = v[k].y;
        } else {
            n0[3 * x + 2] = v[k].x;
            n1[3 * x + 2] = v[k].y;
        }
    }
}

// Row pairs a block of the row passes: lines of m/8 threads and m + 16 LDS
// elements, at most 1024 threads (FP32) or 512 (FP64) and 64 KB; lpp lines a
// pair (1 in pass 1, 3 in pass 5).
static int row_pairs(int m, size_t cs, int lpp) {
    const int maxt = cs == 8 ? 1024 : 512, P = m / 8;
    int np = 16;
    while (np > 1 && (np * lpp * P > maxt || (size_t)np * lpp * (m + 16) * cs > 64 * 1024)) --np;
    return np;
}

// Lines a block of the register transforms: 16 (or as many as make one full
// wave of m/8 threads a line, for m < 32), fewer when the block would exceed
// 1024 threads (FP32) or 512 (FP64) or the LDS buffer (m KT elements) 64 KB;
// the block stays a whole number of waves.
static int reg_lines(int m, size_t cs) {
    const int maxt = cs == 8 ? 1024 : 512, P = m / 8;
    int kt = P >= 4 ? 16 : 64 / P;
    while (kt * P > 64 && (kt * P > maxt || (size_t)m * kt * cs > 64 * 1024)) kt >>= 1;
    return kt;
}

static int ilog2(int m) {
    int l = 0;
    while ((1 << l) < m) ++l;
    return l;
}

// Parameters of the five passes of one solve.
template <class T>
struct FftPlan {
    using C = typename Cx<T>::t;
    const double *rho;
    C *ck, *ek;
    T *emesh;
    const C *tw0, *tw1, *tw2;
    int Mx, My, Mz, Mxh, lx, ly, lz;
    size_t nspec;
    int nl, g1;          // pass 1: lines per block, blocks
    size_t lds1;
    int kty, g2x;        // passes 2 and 4
    size_t lds2;
    GreenZArgs<T> gz;    // pass 3
    int g3x;
    size_t lds3;
    int np, g5;          // pass 5
    size_t lds5;
};

template <class T>
static FftPlan<T> fft_plan(mdf_p3m_ctx *c, const typename Cx<T>::t *tw0, const typename Cx<T>::t *tw1,
                           const typename Cx<T>::t *tw2, typename Cx<T>::t *ck, typename Cx<T>::t *ek, const T *green,
                           T *emesh) {
    using C = typename Cx<T>::t;
    FftPlan<T> F;
    F.rho = c->rho;
    F.ck = ck;
    F.ek = ek;
    F.emesh = emesh;
    F.tw0 = tw0;
    F.tw1 = tw1;
    F.tw2 = tw2;
    const int Mx = c->g.M[0], My = c->g.M[1], Mz = c->g.M[2], Mxh = c->g.Mxh;
    F.Mx = Mx;
    F.My = My;
    F.Mz = Mz;
    F.Mxh = Mxh;
    F.lx = ilog2(Mx);
    F.ly = ilog2(My);
    F.lz = ilog2(Mz);
    F.nspec = c->nspec;
    const size_t cs = sizeof(C);
    int nl = 512 / Mx;
    if (nl < 1) nl = 1;
    while (nl > 1 && (My * Mz) % (2 * nl)) nl >>= 1;
    F.nl = nl;
    F.g1 = (My * Mz) / (2 * nl);
    F.lds1 = 2 * cs * nl * Mx;
    // columns a block: 8 below 256 (16 measured 2-3 us slower at 128^3:
    // half as many blocks, each a longer chain of barrier-separated stages)
    F.kty = My >= 512 ? 4 : 8;
    F.g2x = (Mxh + F.kty - 1) / F.kty;
    F.lds2 = 2 * cs * F.kty * My;
    GreenZArgs<T> &P = F.gz;
    P.ck = ck;
    P.ek = ek;
    P.green = green;
    P.tw = tw2;
    P.part = c->part_pm;
    P.nspec = c->nspec;
    P.M[0] = Mx;
    P.M[1] = My;
    P.M[2] = Mz;
    P.Mxh = Mxh;
    P.lm = F.lz;
    P.kt = Mz >= 512 ? 2 : (Mz >= 256 ? 4 : 8);
    if (sizeof(T) == 4 && Mz >= 256) P.kt *= 2;  // FP32 at 128^3: 8 columns took the pass 6 us below 16
    for (int d = 0; d < 3; ++d) P.kfac[d] = 2.0 * M_PI / c->g.L[d];
    P.inv_V = 1.0 / (c->g.L[0] * c->g.L[1] * c->g.L[2]);
    F.g3x = (Mxh + P.kt - 1) / P.kt;
    F.lds3 = 3 * cs * P.kt * Mz;
    int np = 256 / Mx;
    if (np < 1) np = 1;
    while (np > 1 && (My * Mz) % (2 * np)) np >>= 1;
    F.np = np;
    F.g5 = (My * Mz) / (2 * np);
    F.lds5 = 2 * cs * 3 * np * Mx;
    return F;
}

template <class T>
static int pm_solve_t(mdf_p3m_ctx *c, const typename Cx<T>::t *tw0, const typename Cx<T>::t *tw1,
                      const typename Cx<T>::t *tw2, typename Cx<T>::t *ck, typename Cx<T>::t *ek, const T *green,
                      T *emesh, hipStream_t s) {
    using C = typename Cx<T>::t;
    const FftPlan<T> F = fft_plan<T>(c, tw0, tw1, tw2, ck, ek, green, emesh);
    // x: register transforms of row pairs (rows of 8 or more elements)
    const bool reg_x = F.lx >= 3;
    const int npair = F.My * F.Mz / 2, P8 = F.Mx / 8;
    if (reg_x) {
        const int np1 = row_pairs(F.Mx, sizeof(C), 1);
        const size_t lds = (size_t)np1 * (F.Mx + 16) * sizeof(C);
        MDF_TRY(allow_lds(k_fft_row_r2c<T>, lds));
        hipLaunchKernelGGL(k_fft_row_r2c<T>, dim3((npair + np1 - 1) / np1), dim3(np1 * P8), lds, s, c->rho, ck,
                           F.lx, F.Mxh, np1, npair, tw0);
    } else {
        hipLaunchKernelGGL(k_fft_x_r2c<T>, dim3(F.g1), dim3(256), F.lds1, s, c->rho, ck, F.Mx, F.lx, F.Mxh, F.nl,
                           tw0);
    }
    MDF_LAUNCHED();
    // y and z: register transforms (lines of 8 or more elements)
    const bool reg_y = F.ly >= 3, reg_z = F.lz >= 3;
    const int kty = reg_lines(F.My, sizeof(C)), ktz = reg_lines(F.Mz, sizeof(C));
    const int gy = (F.Mxh + kty - 1) / kty, gz = (F.Mxh + ktz - 1) / ktz;
    const size_t ldsy = (size_t)F.My * kty * sizeof(C), ldsz = (size_t)F.Mz * ktz * sizeof(C);
    if (reg_y) {
        MDF_TRY(allow_lds(k_fft_col<-1, C>, ldsy));
        hipLaunchKernelGGL((k_fft_col<-1, C>), dim3(gy, F.Mz, 1), dim3(kty * (F.My / 8)), ldsy, s, ck, (size_t)0,
                           F.ly, F.Mxh, F.My * F.Mxh, F.Mxh, kty, tw1);
    } else {
        hipLaunchKernelGGL((k_fft_strided<-1, C>), dim3(F.g2x, F.Mz, 1), dim3(256), F.lds2, s, ck, (size_t)0, F.My,
                           F.ly, F.Mxh, F.My * F.Mxh, F.Mxh, F.kty, tw1);
    }
    MDF_LAUNCHED();
    if (reg_z) {
        GreenZArgs<T> G = F.gz;
        G.kt = ktz;
        c->n_part_pm = gz * F.My;
        MDF_TRY(allow_lds(k_fft_zgreen_r<T>, ldsz));
        hipLaunchKernelGGL(k_fft_zgreen_r<T>, dim3(gz, F.My), dim3(ktz * (F.Mz / 8)), ldsz, s, G);
    } else {
        c->n_part_pm = F.g3x * F.My;
        hipLaunchKernelGGL(k_fft_z_green<T>, dim3(F.g3x, F.My), dim3(256), F.lds3, s, F.gz);
    }
    MDF_LAUNCHED();
    if (reg_y) {
        MDF_TRY(allow_lds(k_fft_col<1, C>, ldsy));
        hipLaunchKernelGGL((k_fft_col<1, C>), dim3(gy, F.Mz, 3), dim3(kty * (F.My / 8)), ldsy, s, ek, c->nspec, F.ly,
                           F.Mxh, F.My * F.Mxh, F.Mxh, kty, tw1);
    } else {
        hipLaunchKernelGGL((k_fft_strided<1, C>), dim3(F.g2x, F.Mz, 3), dim3(256), F.lds2, s, ek, c->nspec, F.My,
                           F.ly, F.Mxh, F.My * F.Mxh, F.Mxh, F.kty, tw1);
    }
    MDF_LAUNCHED();
    if (reg_x) {
        const int np5 = row_pairs(F.Mx, sizeof(C), 3);
        const size_t lds = (size_t)np5 * 3 * (F.Mx + 16) * sizeof(C);
        MDF_TRY(allow_lds(k_fft_row_c2r<T>, lds));
        hipLaunchKernelGGL(k_fft_row_c2r<T>, dim3((npair + np5 - 1) / np5), dim3(np5 * 3 * P8), lds, s, ek, emesh,
                           F.lx, F.Mxh, np5, npair, c->nspec, tw0);
    } else {
        hipLaunchKernelGGL(k_fft_x_c2r<T>, dim3(F.g5), dim3(256), F.lds5, s, ek, emesh, F.Mx, F.lx, F.Mxh, F.np,
                           c->nspec, tw0);
    }
    MDF_LAUNCHED();
    return MDF_OK;
}

// ---- segmented stages (replica batches) ----
template <class T>
__global__ __launch_bounds__(256) void k_fft_r2c_seg(const FftPlan<T> *__restrict__ tab, SegGrid S) {
    extern __shared__ __align__(16) unsigned char smraw[];
    Blk b;
    const FftPlan<T> &A = tab[seg_locate(S, b)];
    r2c_body<T>(A.rho, A.ck, A.Mx, A.lx, A.Mxh, A.nl, A.tw0, b.x, smraw);
}
template <int DIR, class T>
__global__ __launch_bounds__(256) void k_fft_y_seg(const FftPlan<T> *__restrict__ tab, SegGrid S) {
    extern __shared__ __align__(16) unsigned char smraw[];
    Blk b;
    const FftPlan<T> &A = tab[seg_locate(S, b)];
    strided_body<DIR, typename Cx<T>::t>(DIR < 0 ? A.ck : A.ek, DIR < 0 ? (size_t)0 : A.nspec, A.My, A.ly, A.Mxh,
                                         A.My * A.Mxh, A.Mxh, A.kty, A.tw1, b.x, b.y, b.z, smraw);
}
template <class T>
__global__ __launch_bounds__(256) void k_fft_zgreen_seg(const FftPlan<T> *__restrict__ tab, SegGrid S) {
    extern __shared__ __align__(16) unsigned char smraw[];
    __shared__ double red[4];
    Blk b;
    const FftPlan<T> &A = tab[seg_locate(S, b)];
    zgreen_body<T>(A.gz, b.x, b.y, b.nx, smraw, red);
}
template <class T>
__global__ __launch_bounds__(256) void k_fft_c2r_seg(const FftPlan<T> *__restrict__ tab, SegGrid S) {
    extern __shared__ __align__(16) unsigned char smraw[];
    Blk b;
    const FftPlan<T> &A = tab[seg_locate(S, b)];
    c2r_body<T>(A.ek, A.emesh, A.Mx, A.lx, A.Mxh, A.np, A.nspec, A.tw0, b.x, smraw);
}

template <class T>
static FftPlan<T> ctx_plan(mdf_p3m_ctx *c) {
    if (sizeof(T) == 4)
        return fft_plan<T>(c, (const typename Cx<T>::t *)c->twf[0], (const typename Cx<T>::t *)c->twf[1],
                           (const typename Cx<T>::t *)c->twf[2], (typename Cx<T>::t *)c->ckf,
                           (typename Cx<T>::t *)c->ekf, (const T *)c->greenf, (T *)c->emeshf);
    return fft_plan<T>(c, (const typename Cx<T>::t *)c->tw[0], (const typename Cx<T>::t *)c->tw[1],
                       (const typename Cx<T>::t *)c->tw[2], (typename Cx<T>::t *)c->ck, (typename Cx<T>::t *)c->ek,
                       (const T *)c->green, (T *)c->emesh);
}

template <class T, int PASS>
static int build_fft(mdf_p3m_ctx *c, const StepIO &, void *args, int4 &dims, size_t &lds) {
    if ((sizeof(T) == 4) != (c->prm.precision == MDF_PREC_MIXED)) {
        set_error("replica batch: mixed precisions");
        return MDF_EINVAL;
    }
    const FftPlan<T> F = ctx_plan<T>(c);
    *(FftPlan<T> *)args = F;
    switch (PASS) {
        case 1: dims = make_int4(F.g1, 1, 1, 0); lds = F.lds1; break;
        case 2: dims = make_int4(F.g2x, F.Mz, 1, 0); lds = F.lds2; break;
        case 3:
            dims = make_int4(F.g3x, F.My, 1, 0);
            lds = F.lds3;
            c->n_part_pm = F.g3x * F.My;
            break;
        case 4: dims = make_int4(F.g2x, F.Mz, 3, 0); lds = F.lds2; break;
        default: dims = make_int4(F.g5, 1, 1, 0); lds = F.lds5; break;
    }
    return MDF_OK;
}

template <class T, int PASS>
static int launch_fft(const void *tab, const SegGrid &S, int total, size_t lds, int, hipStream_t s) {
    const FftPlan<T> *t = (const FftPlan<T> *)tab;
    switch (PASS) {
        case 1: hipLaunchKernelGGL(k_fft_r2c_seg<T>, dim3(total), dim3(256), lds, s, t, S); break;
        case 2: hipLaunchKernelGGL((k_fft_y_seg<-1, T>), dim3(total), dim3(256), lds, s, t, S); break;
        case 3: hipLaunchKernelGGL(k_fft_zgreen_seg<T>, dim3(total), dim3(256), lds, s, t, S); break;
        case 4: hipLaunchKernelGGL((k_fft_y_seg<1, T>), dim3(total), dim3(256), lds, s, t, S); break;
        default: hipLaunchKernelGGL(k_fft_c2r_seg<T>, dim3(total), dim3(256), lds, s, t, S); break;
    }
    MDF_LAUNCHED();
    return MDF_OK;
}

template <class T>
static void push_fft(std::vector<SegStage> &out) {
    out.push_back({"fft_r2c", sizeof(FftPlan<T>), build_fft<T, 1>, launch_fft<T, 1>});
    out.push_back({"fft_y_fwd", sizeof(FftPlan<T>), build_fft<T, 2>, launch_fft<T, 2>});
    out.push_back({"fft_z_green", sizeof(FftPlan<T>), build_fft<T, 3>, launch_fft<T, 3>});
    out.push_back({"fft_y_inv", sizeof(FftPlan<T>), build_fft<T, 4>, launch_fft<T, 4>});
    out.push_back({"fft_c2r", sizeof(FftPlan<T>), build_fft<T, 5>, launch_fft<T, 5>});
}

void seg_stages_fft(std::vector<SegStage> &out, int precision) {
    if (precision == MDF_PREC_MIXED) push_fft<float>(out);
    else push_fft<double>(out);
}

int pm_solve(mdf_p3m_ctx *c, hipStream_t s) {
    if (c->prm.precision == MDF_PREC_MIXED)
        return pm_solve_t<float>(c, c->twf[0], c->twf[1], c->twf[2], c->ckf, c->ekf, c->greenf, c->emeshf, s);
    return pm_solve_t<double>(c, c->tw[0], c->tw[1], c->tw[2], c->ck, c->ek, c->green, c->emesh, s);
}

}  // namespace mdf
