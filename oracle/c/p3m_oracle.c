/*
 * Single-threaded C restatement of the reference P3M Yukawa force path
 * (FP64).  Test infrastructure: the checker for large N, and the CPU
 * baseline ("port") that bench.py times.  The product path never links it.
 *
 * Follows, function by function:
 *   orc_pp              yukawa_pp.py:6-139 (twin: yukawa_p3m.c:182-416)
 *   orc_charge_assign   yukawa_pm_ord6.py:5-95
 *   orc_field_gather    yukawa_pm_ord6.py:212-312
 *   orc_pm              yukawa_pm_ord6.py:439-484 (FFTW -> own radix-2 FFT)
 *   orc_gf_opt          yukawa_gf_opt.py:5-98
 *   orc_force_pot       yukawa_p3m_ord6.py:6-29
 *   orc_md_steps        velocity_verlet_p3m_ord6.py:5-26, thermostat_p3m_ord6.py:5-25,
 *                       yukawa_MD_p3m_ord6.py:195-219
 * Every entry point returns 0 on success and a negative value on error.
 *
 * Energy sums (U_pp over pairs, U_pm over modes) use Neumaier compensated
 * summation: a plain running sum over ~1e8 pair energies drifts by ~1e-12
 * relative, which is the size of the parity bar this checker enforces.
 *
 * Built with OpenMP (c/liboracle_p3m_omp.so, the all-core CPU baseline) the
 * same loops run on all threads: cells, particles, mesh lines and modes are
 * split over threads; the Newton-3 pair updates and the charge deposits go to
 * per-thread arrays summed afterwards in thread order.  The serial build
 * (c/liboracle_p3m.so) is the checker.
 */

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Threads of the OpenMP build (n > 0 sets them); 1 in the serial build. */
int orc_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

static int nthreads_now(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static int thread_id(void)
{
#ifdef _OPENMP
    return omp_get_thread_num();
#else
    return 0;
#endif
}

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct { double re, im; } cplx;

/* Neumaier compensated running sum. */
typedef struct { double s, c; } csum;
static void csum_add(csum *a, double v)
{
    double t = a->s + v;
    if (fabs(a->s) >= fabs(v)) a->c += (a->s - t) + v;
    else a->c += (v - t) + a->s;
    a->s = t;
}
static double csum_get(const csum *a) { return a->s + a->c; }

/* ---------------------------------------------------------------- PP ---- */

/* Pair energy and radial force exactly as written in yukawa_pp.py:115-124. */
static void pair_terms(double r, double kappa, double G, double *phi, double *fr)
{
    double a = 0.5 * kappa / G;
    *phi = (0.5 / r) * (exp(kappa * r) * erfc(G * r + a) + exp(-kappa * r) * erfc(G * r - a));
    double f1 = (0.5 / (r * r)) * exp(kappa * r) * erfc(G * r + a) * (1 - kappa * r);
    double f2 = (0.5 / (r * r)) * exp(-kappa * r) * erfc(G * r - a) * (1 + kappa * r);
    double f3 = (G / sqrt(M_PI) / r) * (exp(-(G * r + a) * (G * r + a)) * exp(kappa * r)
                                        + exp(-(G * r - a) * (G * r - a)) * exp(-kappa * r));
    *fr = f1 + f2 + f3;
}

int orc_pp(const double *pos, int n, const double *Lv, double kappa, double G, double rc,
           double *acc, double *U_out)
{
    const int empty = -50;
    int nd[3];
    double w[3];
    for (int d = 0; d < 3; ++d) {
        nd[d] = (int)floor(Lv[d] / rc);
        if (nd[d] < 1) return -1;
        w[d] = Lv[d] / nd[d];
    }
    int ncell = nd[0] * nd[1] * nd[2];
    int *head = malloc((size_t)ncell * sizeof(int));
    int *ls = malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    if (!head || !ls) { free(head); free(ls); return -2; }
    for (int c = 0; c < ncell; ++c) head[c] = empty;
    memset(acc, 0, (size_t)n * 3 * sizeof(double));
    /* head / linked list, yukawa_pp.py:44-52; out-of-range cells clamped */
    for (int i = 0; i < n; ++i) {
        int cc[3];
        for (int d = 0; d < 3; ++d) {
            int v = (int)floor(pos[3 * i + d] / w[d]);
            cc[d] = v < 0 ? 0 : (v >= nd[d] ? nd[d] - 1 : v);
        }
        int c = cc[0] + cc[1] * nd[0] + cc[2] * nd[0] * nd[1];
        ls[i] = head[c];
        head[c] = i;
    }
    const int nt = nthreads_now();
    double *tacc = acc;
    csum *tU = calloc((size_t)nt, sizeof(csum));
    if (nt > 1) tacc = calloc((size_t)nt * 3 * (size_t)n, sizeof(double));
    if (!tU || !tacc) { free(head); free(ls); free(tU); if (tacc != acc) free(tacc); return -2; }
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4)
#endif
    for (int cell = 0; cell < ncell; ++cell) {
        const int cx = cell % nd[0], cy = (cell / nd[0]) % nd[1], cz = cell / (nd[0] * nd[1]);
        const int tid = thread_id();
        double *acc_t = nt > 1 ? tacc + (size_t)tid * 3 * n : acc;
        csum *U = &tU[tid];
        int c = cx + cy * nd[0] + cz * nd[0] * nd[1];
        for (int czN = cz - 1; czN <= cz + 1; ++czN)
        for (int cyN = cy - 1; cyN <= cy + 1; ++cyN)
        for (int cxN = cx - 1; cxN <= cx + 1; ++cxN) {
            int cn[3] = {cxN, cyN, czN};
            double sh[3];
            for (int d = 0; d < 3; ++d) {
                if (cn[d] < 0) { cn[d] += nd[d]; sh[d] = -Lv[d]; }
                else if (cn[d] >= nd[d]) { cn[d] -= nd[d]; sh[d] = Lv[d]; }
                else sh[d] = 0.0;
            }
            int cN = cn[0] + cn[1] * nd[0] + cn[2] * nd[0] * nd[1];
            for (int i = head[c]; i != empty; i = ls[i]) {
                for (int j = head[cN]; j != empty; j = ls[j]) {
                    if (i < j) {
                        double dx = pos[3 * i + 0] - (pos[3 * j + 0] + sh[0]);
                        double dy = pos[3 * i + 1] - (pos[3 * j + 1] + sh[1]);
                        double dz = pos[3 * i + 2] - (pos[3 * j + 2] + sh[2]);
                        double r = sqrt(dx * dx + dy * dy + dz * dz);
                        if (r < rc) {
                            double phi, fr;
                            pair_terms(r, kappa, G, &phi, &fr);
                            csum_add(U, phi);
                            acc_t[3 * i + 0] += fr * dx / r;
                            acc_t[3 * i + 1] += fr * dy / r;
                            acc_t[3 * i + 2] += fr * dz / r;
                            acc_t[3 * j + 0] -= fr * dx / r;
                            acc_t[3 * j + 1] -= fr * dy / r;
                            acc_t[3 * j + 2] -= fr * dz / r;
                        }
                    }
                }
            }
        }
    }
    if (nt > 1) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
        for (long k = 0; k < 3L * n; ++k) {
            double a = 0.0;
            for (int t = 0; t < nt; ++t) a += tacc[(size_t)t * 3 * n + k];
            acc[k] = a;
        }
        free(tacc);
    }
    csum Ut = {0.0, 0.0};
    for (int t = 0; t < nt; ++t) {
        csum_add(&Ut, tU[t].s);
        csum_add(&Ut, tU[t].c);
    }
    free(tU);
    free(head);
    free(ls);
    *U_out = csum_get(&Ut);
    return 0;
}

/* --------------------------------------------------- assignment / gather -- */

/* Weights of yukawa_pm_ord6.py:29-48 for offset x in [-1/2, 1/2). */
static void weights6(double x, double *w)
{
    double x2 = x * x, x3 = x2 * x, x4 = x3 * x, x5 = x4 * x;
    w[0] = (1 - 10 * x + 40 * x2 - 80 * x3 + 80 * x4 - 32 * x5) / 3840;
    w[1] = (237 - 750 * x + 840 * x2 - 240 * x3 - 240 * x4 + 160 * x5) / 3840;
    w[2] = (841 - 770 * x - 440 * x2 + 560 * x3 + 80 * x4 - 160 * x5) / 1920;
    w[3] = (841 + 770 * x - 440 * x2 - 560 * x3 + 80 * x4 + 160 * x5) / 1920;
    w[4] = (237 + 750 * x + 840 * x2 + 240 * x3 - 240 * x4 - 160 * x5) / 3840;
    w[5] = (1 + 10 * x + 40 * x2 + 80 * x3 + 80 * x4 + 32 * x5) / 3840;
}

static int wrap1(int i, int M) { return i < 0 ? i + M : (i > M - 1 ? i - M : i); }

static void stencil(const double *p, const int *M, const double *h, int base[3], double w[3][6])
{
    for (int d = 0; d < 3; ++d) {
        int i0 = (int)floor(p[d] / h[d]);
        double x = (p[d] - (i0 + 0.5) * h[d]) / h[d];
        weights6(x, w[d]);
        base[d] = i0;
    }
}

int orc_charge_assign(const double *pos, const double *Z, int n, int Mx, int My, int Mz,
                      double hx, double hy, double hz, double *rho)
{
    int M[3] = {Mx, My, Mz};
    double h[3] = {hx, hy, hz};
    const size_t mv = (size_t)Mx * My * Mz;
    memset(rho, 0, mv * sizeof(double));
    const int nt = nthreads_now();
    double *trho = rho;
    if (nt > 1) {
        trho = calloc((size_t)nt * mv, sizeof(double));
        if (!trho) return -2;
    }
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int ip = 0; ip < n; ++ip) {
        double *rho_t = nt > 1 ? trho + (size_t)thread_id() * mv : rho;
        int b[3];
        double w[3][6];
        stencil(pos + 3 * ip, M, h, b, w);
        for (int g = 0; g < 6; ++g) {
            int rg = wrap1(b[2] + g, Mz);
            for (int i = 0; i < 6; ++i) {
                int ri = wrap1(b[1] + i, My);
                for (int j = 0; j < 6; ++j) {
                    int rj = wrap1(b[0] + j, Mx);
                    rho_t[((size_t)rg * My + ri) * Mx + rj] += Z[ip] * w[2][g] * w[1][i] * w[0][j];
                }
            }
        }
    }
    if (nt > 1) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
        for (long k = 0; k < (long)mv; ++k) {
            double a = 0.0;
            for (int t = 0; t < nt; ++t) a += trho[(size_t)t * mv + k];
            rho[k] = a;
        }
        free(trho);
    }
    return 0;
}

int orc_field_gather(const double *Ex, const double *Ey, const double *Ez, const double *pos,
                     const double *Z, int n, int Mx, int My, int Mz, double hx, double hy,
                     double hz, double *acc, double *E_p)
{
    int M[3] = {Mx, My, Mz};
    double h[3] = {hx, hy, hz};
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int ip = 0; ip < n; ++ip) {
        int b[3];
        double w[3][6];
        stencil(pos + 3 * ip, M, h, b, w);
        double ex = 0, ey = 0, ez = 0;
        for (int g = 0; g < 6; ++g) {
            int rg = wrap1(b[2] + g, Mz);
            for (int i = 0; i < 6; ++i) {
                int ri = wrap1(b[1] + i, My);
                for (int j = 0; j < 6; ++j) {
                    int rj = wrap1(b[0] + j, Mx);
                    size_t k = ((size_t)rg * My + ri) * Mx + rj;
                    double ww = w[2][g] * w[1][i] * w[0][j];
                    ex += Ex[k] * ww;
                    ey += Ey[k] * ww;
                    ez += Ez[k] * ww;
                }
            }
        }
        if (E_p) { E_p[3 * ip] = ex; E_p[3 * ip + 1] = ey; E_p[3 * ip + 2] = ez; }
        acc[3 * ip] = Z[ip] * ex;
        acc[3 * ip + 1] = Z[ip] * ey;
        acc[3 * ip + 2] = Z[ip] * ez;
    }
    return 0;
}

/* ----------------------------------------------------------------- FFT ---- */

/* In-place 1D DFT of length m, sign -1 forward, +1 inverse (unnormalised).
 * Radix-2 iterative Cooley-Tukey for powers of two, direct O(m^2) sum
 * otherwise.  tmp must hold m entries. */
static void fft1d(cplx *x, int m, int sign, cplx *tmp)
{
    if (m <= 1) return;
    if ((m & (m - 1)) == 0) {
        for (int i = 1, j = 0; i < m; ++i) {
            int bit = m >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) { cplx t = x[i]; x[i] = x[j]; x[j] = t; }
        }
        for (int len = 2; len <= m; len <<= 1) {
            int half = len >> 1;
            for (int k = 0; k < half; ++k) {
                double ang = sign * 2.0 * M_PI * k / len;
                double wr = cos(ang), wi = sin(ang);
                for (int s = 0; s < m; s += len) {
                    cplx u = x[s + k], v = x[s + k + half];
                    double vr = v.re * wr - v.im * wi, vi = v.re * wi + v.im * wr;
                    x[s + k].re = u.re + vr; x[s + k].im = u.im + vi;
                    x[s + k + half].re = u.re - vr; x[s + k + half].im = u.im - vi;
                }
            }
        }
    } else {
        for (int k = 0; k < m; ++k) {
            double sr = 0, si = 0;
            for (int j = 0; j < m; ++j) {
                double ang = sign * 2.0 * M_PI * (double)(((long)j * k) % m) / m;
                sr += x[j].re * cos(ang) - x[j].im * sin(ang);
                si += x[j].re * sin(ang) + x[j].im * cos(ang);
            }
            tmp[k].re = sr; tmp[k].im = si;
        }
        memcpy(x, tmp, (size_t)m * sizeof(cplx));
    }
}

/* 3D transform of c[Mz][My][Mx] along all axes. */
static int fft3d(cplx *c, int Mx, int My, int Mz, int sign)
{
    int mm = Mx > My ? Mx : My;
    mm = mm > Mz ? mm : Mz;
    const int nt = nthreads_now();
    cplx *lines = malloc((size_t)nt * 2 * mm * sizeof(cplx));  /* per thread: line, tmp */
    if (!lines) return -2;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int zy = 0; zy < Mz * My; ++zy) {
        cplx *tmp = lines + (size_t)thread_id() * 2 * mm + mm;
        fft1d(c + (size_t)zy * Mx, Mx, sign, tmp);
    }
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int zx = 0; zx < Mz * Mx; ++zx) {
        const int z = zx / Mx, x = zx % Mx;
        cplx *line = lines + (size_t)thread_id() * 2 * mm, *tmp = line + mm;
        for (int y = 0; y < My; ++y) line[y] = c[((size_t)z * My + y) * Mx + x];
        fft1d(line, My, sign, tmp);
        for (int y = 0; y < My; ++y) c[((size_t)z * My + y) * Mx + x] = line[y];
    }
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int yx = 0; yx < My * Mx; ++yx) {
        const int y = yx / Mx, x = yx % Mx;
        cplx *line = lines + (size_t)thread_id() * 2 * mm, *tmp = line + mm;
        for (int z = 0; z < Mz; ++z) line[z] = c[((size_t)z * My + y) * Mx + x];
        fft1d(line, Mz, sign, tmp);
        for (int z = 0; z < Mz; ++z) c[((size_t)z * My + y) * Mx + x] = line[z];
    }
    free(lines);
    return 0;
}

int orc_fft3d(double *c_interleaved, int Mx, int My, int Mz, int sign)
{
    return fft3d((cplx *)c_interleaved, Mx, My, Mz, sign);
}

/* ---------------------------------------------------------- Green fn ---- */

int orc_gf_opt(double kappa, double Gew, int p, int mx_max, int my_max, int mz_max, int Mx,
               int My, int Mz, double hx, double hy, double hz, double Lx, double Ly, double Lz,
               double *G_k, double *A_pm_out)
{
    double kappa_sq = kappa * kappa, Gew_sq = Gew * Gew;
    int nx_mid = Mx / 2, ny_mid = My / 2, nz_mid = Mz / 2; /* (M-1)/2 for odd M */
    double A_pm = 0.0;
    /* modes are independent; only A_pm's summation order depends on the threads */
#pragma omp parallel for reduction(+ : A_pm) schedule(static)
    for (int nz = 0; nz < Mz; ++nz) {
        int nz_sh = nz - nz_mid;
        double kz = 2 * M_PI * nz_sh / Lz;
        for (int ny = 0; ny < My; ++ny) {
            int ny_sh = ny - ny_mid;
            double ky = 2 * M_PI * ny_sh / Ly;
            for (int nx = 0; nx < Mx; ++nx) {
                int nx_sh = nx - nx_mid;
                double kx = 2 * M_PI * nx_sh / Lx;
                double k_sq = kx * kx + ky * ky + kz * kz;
                size_t idx = ((size_t)nz * My + ny) * Mx + nx;
                G_k[idx] = 0.0;
                if (k_sq == 0.0) continue;
                double U_k_sq = 0.0, U_G_k = 0.0;
                for (int mz = -mz_max; mz <= mz_max; ++mz)
                for (int my = -my_max; my <= my_max; ++my)
                for (int mx = -mx_max; mx <= mx_max; ++mx) {
                    double kx_M = 2 * M_PI * (nx_sh + mx * Mx) / Lx;
                    double ky_M = 2 * M_PI * (ny_sh + my * My) / Ly;
                    double kz_M = 2 * M_PI * (nz_sh + mz * Mz) / Lz;
                    double k_M_sq = kx_M * kx_M + ky_M * ky_M + kz_M * kz_M;
                    double Ux = kx_M != 0.0 ? sin(0.5 * kx_M * hx) / (0.5 * kx_M * hx) : 1.0;
                    double Uy = ky_M != 0.0 ? sin(0.5 * ky_M * hy) / (0.5 * ky_M * hy) : 1.0;
                    double Uz = kz_M != 0.0 ? sin(0.5 * kz_M * hz) / (0.5 * kz_M * hz) : 1.0;
                    double U_k_M = pow(Ux * Uy * Uz, p);
                    double U_k_M_sq = U_k_M * U_k_M;
                    double G_k_M = 4 * M_PI * exp(-0.25 * (kappa_sq + k_M_sq) / Gew_sq) / (kappa_sq + k_M_sq);
                    double k_dot_k_M = kx * kx_M + ky * ky_M + kz * kz_M;
                    U_G_k += U_k_M_sq * G_k_M * k_dot_k_M;
                    U_k_sq += U_k_M_sq;
                }
                G_k[idx] = U_G_k / ((U_k_sq * U_k_sq) * k_sq);
                A_pm += U_G_k * U_G_k / ((U_k_sq * U_k_sq) * k_sq);
            }
        }
    }
    if (A_pm_out) *A_pm_out = A_pm;
    return 0;
}

/* ------------------------------------------------------------------ PM ---- */

int orc_pm(const double *pos, const double *Z, int n, const double *G_k, int Mx, int My, int Mz,
           double hx, double hy, double hz, double Lx, double Ly, double Lz,
           double *acc_pm, double *E_p, double *rho_out, double *U_out)
{
    size_t mv = (size_t)Mx * My * Mz;
    double V = Lx * Ly * Lz;
    double *rho = rho_out ? rho_out : malloc(mv * sizeof(double));
    cplx *rk = malloc(mv * sizeof(cplx));
    cplx *ek = malloc(mv * sizeof(cplx));
    double *E = malloc(3 * mv * sizeof(double));
    if (!rho || !rk || !ek || !E) return -2;
    orc_charge_assign(pos, Z, n, Mx, My, Mz, hx, hy, hz, rho);
    for (size_t i = 0; i < mv; ++i) { rk[i].re = rho[i]; rk[i].im = 0.0; }
    fft3d(rk, Mx, My, Mz, -1);
    int Ms[3] = {Mx, My, Mz};
    double Ls[3] = {Lx, Ly, Lz};
    /* natural index n <-> shifted index (n + M/2) % M <-> k = 2 pi (s - M/2)/L */
    csum U = {0.0, 0.0};
    for (int a = 0; a < 3; ++a) {
        if (a == 0) {  /* U_f, in the serial order */
            for (size_t i = 0; i < mv; ++i) {
                const int x = (int)(i % Mx), y = (int)((i / Mx) % My), z = (int)(i / ((size_t)Mx * My));
                const int sx = (x + Mx / 2) % Mx, sy = (y + My / 2) % My, sz = (z + Mz / 2) % Mz;
                const double g = G_k[((size_t)sz * My + sy) * Mx + sx];
                csum_add(&U, (rk[i].re * rk[i].re + rk[i].im * rk[i].im) * g);
            }
        }
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
        for (int zy = 0; zy < Mz * My; ++zy)
        for (int x = 0; x < Mx; ++x) {
            const int z = zy / My, y = zy % My;
            int nat[3] = {x, y, z};
            int sh[3];
            for (int d = 0; d < 3; ++d) sh[d] = (nat[d] + Ms[d] / 2) % Ms[d];
            size_t i = ((size_t)z * My + y) * Mx + x;
            double g = G_k[((size_t)sh[2] * My + sh[1]) * Mx + sh[0]];
            double pr = g * rk[i].re, pi = g * rk[i].im;
            double k = 2 * M_PI * (sh[a] - Ms[a] / 2) / Ls[a];
            /* E_k = -i k phi_k */
            ek[i].re = k * pi;
            ek[i].im = -k * pr;
        }
        fft3d(ek, Mx, My, Mz, +1);
        for (size_t i = 0; i < mv; ++i) E[a * mv + i] = ek[i].re / V;
    }
    orc_field_gather(E, E + mv, E + 2 * mv, pos, Z, n, Mx, My, Mz, hx, hy, hz, acc_pm, E_p);
    *U_out = 0.5 * csum_get(&U) / V;
    if (!rho_out) free(rho);
    free(rk); free(ek); free(E);
    return 0;
}

/* ------------------------------------------------------------ force_pot -- */

int orc_force_pot(const double *pos, const double *Z, int n, const double *Lv, double kappa,
                  double G, double rc, double af, double uf, const double *G_k, int M,
                  double *acc, double *acc_pp, double *acc_pm, double *U_parts)
{
    double h = Lv[0] / M;
    double U_pp = 0, U_pm = 0;
    int rc1 = orc_pp(pos, n, Lv, kappa, G, rc, acc_pp, &U_pp);
    if (rc1) return rc1;
    int rc2 = orc_pm(pos, Z, n, G_k, M, M, M, h, h, h, Lv[0], Lv[1], Lv[2], acc_pm, NULL, NULL, &U_pm);
    if (rc2) return rc2;
    for (int i = 0; i < 3 * n; ++i) acc[i] = af * (acc_pp[i] + acc_pm[i]);
    double U_self = -n * G / sqrt(M_PI);
    U_parts[0] = uf * (U_pp + U_pm + U_self);
    U_parts[1] = U_pp;
    U_parts[2] = U_pm;
    U_parts[3] = U_self;
    return 0;
}

/* ------------------------------------------------------------- MD loop --- */

/* nsteps Velocity-Verlet steps (NVE) with per-step K, U and n(q) along the
 * three axes for nq wave numbers.  energies[4*s..] = (E, K, U, T);
 * nqt[(s*nq + iq)*3 + a] complex (interleaved re, im). */
int orc_md_steps(double *pos, double *vel, double *acc, const double *Z, int n, int nsteps,
                 double dt, const double *Lv, double kappa, double G, double rc, double af,
                 double uf, double kf, const double *G_k, int M, const double *q, int nq,
                 double *energies, double *nqt)
{
    double *app = malloc((size_t)n * 3 * sizeof(double));
    double *apm = malloc((size_t)n * 3 * sizeof(double));
    if (!app || !apm) { free(app); free(apm); return -2; }
    for (int s = 0; s < nsteps; ++s) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
        for (int i = 0; i < 3 * n; ++i) {
            vel[i] += 0.5 * acc[i] * dt;
            pos[i] += vel[i] * dt;
            int d = i % 3;
            if (pos[i] > Lv[d]) pos[i] -= Lv[d];
            if (pos[i] < 0.0) pos[i] += Lv[d];
        }
        double Up[4];
        int rc1 = orc_force_pot(pos, Z, n, Lv, kappa, G, rc, af, uf, G_k, M, acc, app, apm, Up);
        if (rc1) { free(app); free(apm); return rc1; }
        double K = 0;
        for (int i = 0; i < 3 * n; ++i) {
            vel[i] += 0.5 * acc[i] * dt;
            K += vel[i] * vel[i];
        }
        K *= kf;
        if (energies) {
            energies[4 * s + 0] = K + Up[0];
            energies[4 * s + 1] = K;
            energies[4 * s + 2] = Up[0];
            energies[4 * s + 3] = K / kf / n;
        }
        if (nqt) {
#ifdef _OPENMP
#pragma omp parallel for collapse(2) schedule(static)
#endif
            for (int iq = 0; iq < nq; ++iq)
                for (int a = 0; a < 3; ++a) {
                    double sr = 0, si = 0;
                    for (int j = 0; j < n; ++j) {
                        double ph = -q[iq] * pos[3 * j + a];
                        sr += cos(ph);
                        si += sin(ph);
                    }
                    nqt[2 * (((size_t)s * nq + iq) * 3 + a)] = sr;
                    nqt[2 * (((size_t)s * nq + iq) * 3 + a) + 1] = si;
                }
        }
    }
    free(app);
    free(apm);
    return 0;
}
