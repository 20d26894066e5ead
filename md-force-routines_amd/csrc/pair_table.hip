// Host-side construction of the PP pair-term tables (see pp.hip).
//
// g1(r) = r phi(r) and g2(r) = r^2 f(r) of the reference pair terms
// (yukawa_pp.py:143-152) are fitted on [0, rc] by piecewise Chebyshev
// series: npiece equal pieces of a given degree, coefficients from the
// discrete cosine transform of long-double values at the Chebyshev nodes of
// each piece.  The fit, with its coefficients rounded to the storage type
// (FP64, or FP32 for the mixed-precision kernel), is then checked against
// direct long-double evaluation at 64 points per piece; if the worst error
// exceeds tol * max|g| on [0, rc], the piece count is doubled (up to 512).
// FP64 tables use degree 13 and tol 4e-15; FP32 tables degree 7 and 1e-7.
#include <cmath>
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
            for (int k = 0; k < nc; ++k) {
                long double s1 = 0.0L, s2 = 0.0L;
                for (int j = 0; j < nc; ++j) {
                    const long double cs = cosl(pi * k * (j + 0.5L) / nc);
                    s1 += v1[j] * cs;
                    s2 += v2[j] * cs;
                }
                const long double f = (k == 0 ? 1.0L : 2.0L) / nc;
                tab[(size_t)p * nc + k] = make_double2((double)(f * s1), (double)(f * s2));
            }
        }
        // check
        long double m1 = 0.0L, m2 = 0.0L, e1 = 0.0L, e2 = 0.0L;
        for (int p = 0; p < P; ++p) {
            for (int s = 0; s <= 64; ++s) {
                const long double t = -1.0L + 2.0L * s / 64.0L;
                long double g1, g2;
                pair_g(p * w + (t + 1.0L) * 0.5L * w, kappa, alpha, g1, g2);
                const double2 *c = &tab[(size_t)p * nc];
                const long double a1 = clenshaw_ld(c, deg, 0, as_float, t), a2 = clenshaw_ld(c, deg, 1, as_float, t);
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
