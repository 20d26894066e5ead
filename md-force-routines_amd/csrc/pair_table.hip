// Host-side construction of the PP pair-term tables (see pp.hip).
//
// g1(r) = r phi(r) and g2(r) = r^2 f(r) of the reference pair terms
// (yukawa_pp.py:115-124) are fitted on [0, rc] by piecewise Chebyshev
// series: npiece equal pieces of a given degree, coefficients from the
// discrete cosine transform of long-double values at the Chebyshev nodes of
// each piece.  FP64 tables keep the Chebyshev coefficients (Clenshaw in the
// kernel); FP32 tables hold each piece in the power basis of the same local
// variable t in [-1, 1] (Horner in the kernel: 2 x degree FMAs for both
// terms).  At these low degrees the power-basis coefficients of the smooth
// pieces stay within a few times the function: FP32 Horner measured
// 5e-8 .. 9e-8 of max|g| at degrees 4 .. 7 (FP32 Clenshaw at degree 7:
// 7e-8 .. 1.3e-7; kappa 0.5 .. 3, alpha 0.25 .. 1.1, rc 3 .. 14).  The fit,
// with its coefficients rounded to the storage type, is checked against
// direct long-double evaluation at 64 points per piece; if the worst error
// exceeds tol * max|g| on [0, rc], the piece count is doubled (up to 512).
// FP64 tables use degree 13 and tol 4e-15.  FP32 tables use degree 4 and
// tol 1e-7: 96 pieces (3.8 KB) at kappa 2, alpha 0.4, rc 7, where the PP
// kernel measured 2.16 / 1.82 / 1.67 / 1.98 ms at degrees 7 / 5 / 4 / 3
// (N = 1e6): fewer FMAs and table reads per pair beat the larger table
// down to degree 4.
#include <cmath>
#include <cstring>
#include <vector>

#include "mdf_internal.h"

namespace mdf {

static void pair_g(long double r, long double kappa, long double alpha, long double &g1, long double &g2) {
    const long double b = kappa / (2.0L * alpha);
    const long double A = expl(kappa * r) * erfcl(alpha * r + b);   // e^{kr} erfc(ar + b)
    const long double B = expl(-kappa * r) * erfcl(alpha * r - b);  // e^{-kr} erfc(ar - b)
    const long double pi = 3.141592653589793238462643383279502884L;
    g1 = 0.5L * (A + B);
    g2 = 0.5L * (A * (1.0L - kappa * r) + B * (1.0L + kappa * r)) +
         (2.0L * alpha / sqrtl(pi)) * r * expl(-(alpha * r) * (alpha * r) - b * b);
}

static long double clenshaw_ld(const double2 *c, int deg, int comp, bool as_float, long double t) {
    long double b1 = 0.0L, b2 = 0.0L;
    auto coef = [&](int k) -> long double {
        const double v = comp ? c[k].y : c[k].x;
        return as_float ? (long double)(float)v : (long double)v;
    };
    for (int k = deg; k >= 1; --k) {
        const long double ck = coef(k);
        const long double nb = 2.0L * t * b1 - b2 + ck;
        b2 = b1;
        b1 = nb;
    }
    return t * b1 - b2 + coef(0);
}

static long double horner_ld(const double2 *c, int deg, int comp, long double t) {
    long double a = (long double)(float)(comp ? c[deg].y : c[deg].x);
    for (int k = deg - 1; k >= 0; --k) a = a * t + (long double)(float)(comp ? c[k].y : c[k].x);
    return a;
}

// Power-basis coefficients m[j] of sum_k c[k] T_k(t), k, j <= deg
static void cheb_to_power(const long double *c, int deg, long double *m) {
    long double T[32][32] = {};
    T[0][0] = 1.0L;
    if (deg >= 1) T[1][1] = 1.0L;
    for (int k = 2; k <= deg; ++k)
        for (int j = 0; j <= k; ++j) T[k][j] = (j ? 2.0L * T[k - 1][j - 1] : 0.0L) - T[k - 2][j];
    for (int j = 0; j <= deg; ++j) {
        long double s = 0.0L;
        for (int k = j; k <= deg; ++k) s += c[k] * T[k][j];
        m[j] = s;
    }
}

int build_pair_table(double kappa, double alpha, double rc, int deg, double tol, bool as_float,
                     std::vector<double2> &tab, int &npiece) {
    const int nc = deg + 1;
    if (nc > 32) {
        set_error("pair-table degree %d too large", deg);
        return MDF_EINVAL;
    }
    const long double pi = 3.141592653589793238462643383279502884L;
    double scale = alpha * rc;
    if (0.5 * kappa * rc > scale) scale = 0.5 * kappa * rc;
    int P = (int)ceil(scale / 0.6);
    if (P < 1) P = 1;
    for (; P <= 512; P *= 2) {
        const long double w = (long double)rc / P;
        tab.assign((size_t)P * nc, make_double2(0.0, 0.0));
        for (int p = 0; p < P; ++p) {
            long double v1[32], v2[32];
            for (int j = 0; j < nc; ++j) {
                const long double t = cosl(pi * (j + 0.5L) / nc);
                pair_g(p * w + (t + 1.0L) * 0.5L * w, kappa, alpha, v1[j], v2[j]);
            }
            long double c1[32], c2[32];
            for (int k = 0; k < nc; ++k) {
                long double s1 = 0.0L, s2 = 0.0L;
                for (int j = 0; j < nc; ++j) {
                    const long double cs = cosl(pi * k * (j + 0.5L) / nc);
                    s1 += v1[j] * cs;
                    s2 += v2[j] * cs;
                }
                const long double f = (k == 0 ? 1.0L : 2.0L) / nc;
                c1[k] = f * s1;
                c2[k] = f * s2;
            }
            if (as_float) {
                long double m1[32], m2[32];
                cheb_to_power(c1, deg, m1);
                cheb_to_power(c2, deg, m2);
                for (int k = 0; k < nc; ++k) {
                    c1[k] = m1[k];
                    c2[k] = m2[k];
                }
            }
            for (int k = 0; k < nc; ++k) tab[(size_t)p * nc + k] = make_double2((double)c1[k], (double)c2[k]);
        }
        // check
        long double m1 = 0.0L, m2 = 0.0L, e1 = 0.0L, e2 = 0.0L;
        for (int p = 0; p < P; ++p) {
            for (int s = 0; s <= 64; ++s) {
                const long double t = -1.0L + 2.0L * s / 64.0L;
                long double g1, g2;
                pair_g(p * w + (t + 1.0L) * 0.5L * w, kappa, alpha, g1, g2);
                const double2 *c = &tab[(size_t)p * nc];
                const long double a1 = as_float ? horner_ld(c, deg, 0, t) : clenshaw_ld(c, deg, 0, false, t);
                const long double a2 = as_float ? horner_ld(c, deg, 1, t) : clenshaw_ld(c, deg, 1, false, t);
                m1 = fmaxl(m1, fabsl(g1));
                m2 = fmaxl(m2, fabsl(g2));
                e1 = fmaxl(e1, fabsl(a1 - g1));
                e2 = fmaxl(e2, fabsl(a2 - g2));
            }
        }
        if (e1 <= (long double)tol * m1 && e2 <= (long double)tol * m2) {
            npiece = P;
            return MDF_OK;
        }
    }
    set_error("pair-term table did not reach %g with 512 pieces (kappa=%g alpha=%g rc=%g)", tol, kappa, alpha, rc);
    return MDF_EINVAL;
}

}  // namespace mdf

namespace mdf {

// FP32 table in the scaled squared distance of the mixed-precision PP
// kernel.  That kernel keeps tile coordinates in units of 1 / sigma, with
// sigma = 2^(E/2) / rc, so the table is a function of s' = sigma^2 r^2 and
// the cutoff r = rc is s' = 2^E exactly.  A piece is 2^-m of an octave of s',
// so the float bits of s' give the piece directly: its exponent and top m
// mantissa bits, counted from s_lo (the power of two at or above the s' of
// r = 1/2); the low 23 - m mantissa bits give the position t in the piece.
// Pieces of equal relative width follow the r^-1 and r^-3 growth of the pair
// terms at small r with one degree, and no square root is needed to find a
// pair's piece.  Per piece the table holds, as degree-deg power series of t in
// [-1, 1),
//   phi(s') = g1(r) / r               (pair energy)
//   F(s')   = g2(r) / r^3 / sigma     (force on i: F (x_i - x_j) sigma)
// stored for the variable f - tc = t / 2^(m+1) (coefficient k times
// 2^((m+1) k), exact in float): the kernel forms f - tc with one subtraction,
// and Horner in it gives t's results bit for bit (power-of-two scalings).
// One more piece, all zeros, starts at s' = 2^E: the kernel clamps s' into
// [s_lo, 2^E] (one med3), so every pair at or beyond the cutoff reads zeros
// and needs no cutoff test of its own.  The pieces are fitted at the
// Chebyshev nodes in long double; the fit with coefficients rounded to float
// is checked at 64 points per piece, with the same measure as the r-indexed
// table: the errors of g1 = r phi and g2 = r^3 F (unscaled) against tol times
// max|g1| and max|g2| on [0, rc]; m grows (to at most 6) until it passes.
// Pairs closer than s_lo use the r-indexed table instead.
int build_pair_table_s(double kappa, double alpha, double rc, int deg, double tol, int E, std::vector<double2> &tab,
                       SPieces &sp) {
    const int nc = deg + 1;
    if (nc > 32) {
        set_error("pair-table degree %d too large", deg);
        return MDF_EINVAL;
    }
    if (E < -60 || E > 120) {
        set_error("s-indexed pair table: scale exponent %d out of range", E);
        return MDF_EINVAL;
    }
    const long double pi = 3.141592653589793238462643383279502884L;
    const long double sig2 = ldexpl(1.0L, E) / ((long double)rc * (long double)rc);  // s' = sig2 r^2
    const long double sig = sqrtl(sig2);
    const int e_lo = (int)ceill(log2l(0.25L * sig2));
    const float s_lo = std::ldexp(1.0f, e_lo);
    const float s_top = std::ldexp(1.0f, E);
    for (int m = 3; m <= 6; ++m) {
        const int shift = 23 - m;
        unsigned lo_bits, top_bits;
        std::memcpy(&lo_bits, &s_lo, 4);
        std::memcpy(&top_bits, &s_top, 4);
        const unsigned base = lo_bits >> shift;
        const int P = (int)((top_bits >> shift) - base);  // pieces below 2^E; piece P is the zero piece
        tab.assign((size_t)(P + 1) * nc, make_double2(0.0, 0.0));
        auto piece_s = [&](int p, long double t) {  // s' at t in [-1, 1] of piece p
            unsigned b0 = (base + (unsigned)p) << shift, b1 = (base + (unsigned)p + 1) << shift;
            float f0, f1;
            std::memcpy(&f0, &b0, 4);
            std::memcpy(&f1, &b1, 4);
            return (long double)f0 + (t + 1.0L) * 0.5L * ((long double)f1 - (long double)f0);
        };
        auto funcs = [&](long double sp_, long double &phi, long double &F) {  // of s'
            const long double s = sp_ / sig2, r = sqrtl(s);
            long double g1, g2;
            pair_g(r, kappa, alpha, g1, g2);
            phi = g1 / r;
            F = g2 / (r * s) / sig;
        };
        long double g1max = 0.0L, g2max = 0.0L;  // on [0, rc], as the r-indexed table's check
        for (int q = 0; q <= 4096; ++q) {
            long double g1, g2;
            pair_g((long double)rc * q / 4096.0L, kappa, alpha, g1, g2);
            g1max = fmaxl(g1max, fabsl(g1));
            g2max = fmaxl(g2max, fabsl(g2));
        }
        bool ok = true;
        for (int p = 0; p < P && ok; ++p) {
            long double v1[32], v2[32];
            for (int j = 0; j < nc; ++j) funcs(piece_s(p, cosl(pi * (j + 0.5L) / nc)), v1[j], v2[j]);
            long double c1[32], c2[32];
            for (int k = 0; k < nc; ++k) {
                long double s1 = 0.0L, s2 = 0.0L;
                for (int j = 0; j < nc; ++j) {
                    const long double cs = cosl(pi * k * (j + 0.5L) / nc);
                    s1 += v1[j] * cs;
                    s2 += v2[j] * cs;
                }
                const long double f = (k == 0 ? 1.0L : 2.0L) / nc;
                c1[k] = f * s1;
                c2[k] = f * s2;
            }
            long double m1[32], m2[32];
            cheb_to_power(c1, deg, m1);
            cheb_to_power(c2, deg, m2);
            double2 *c = &tab[(size_t)p * nc];
            for (int k = 0; k < nc; ++k) c[k] = make_double2((double)m1[k], (double)m2[k]);
            long double e1 = 0.0L, e2 = 0.0L;
            for (int q = 0; q <= 64; ++q) {
                const long double t = -1.0L + 2.0L * q / 64.0L;
                const long double sv = piece_s(p, t), s = sv / sig2, r = sqrtl(s);
                long double a, b;
                funcs(sv, a, b);
                e1 = fmaxl(e1, fabsl(horner_ld(c, deg, 0, t) - a) * r);
                e2 = fmaxl(e2, fabsl(horner_ld(c, deg, 1, t) - b) * sig * r * s);
            }
            ok = e1 <= (long double)tol * g1max && e2 <= (long double)tol * g2max;
        }
        if (ok) {
            // f in [1, 1 + 2^-m): t = 2^(m+1) (f - tc), tc = 1 + 2^-(m+1)
            for (int p = 0; p < P; ++p)
                for (int k = 0; k < nc; ++k) {
                    const double sc = std::ldexp(1.0, (m + 1) * k);
                    tab[(size_t)p * nc + k].x *= sc;
                    tab[(size_t)p * nc + k].y *= sc;
                }
            sp.npiece = P + 1;
            sp.base = base;
            sp.shift = shift;
            sp.s_lo = s_lo;
            sp.s_top = s_top;
            sp.tc = 1.0f + std::ldexp(1.0f, -(m + 1));
            sp.E = E;
            sp.qscale = (double)sig;
            sp.inv_q = (float)(1.0L / sig);
            return MDF_OK;
        }
    }
    set_error("s-indexed pair table did not reach %g with 64 pieces an octave (kappa=%g alpha=%g rc=%g)", tol, kappa,
              alpha, rc);
    return MDF_EINVAL;
}

}  // namespace mdf
