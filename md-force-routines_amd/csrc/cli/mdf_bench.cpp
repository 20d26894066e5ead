// Standalone PP / PM timing through the C ABI, the counterpart of the
// reference's C driver (yukawa_p3m.c:419-560) and of the CUDA experiment's
// timing (lcl_g_2.cu:184-567).
//
//   mdf_bench N positions_file Gew rc [kappa [M [fp64|mixed [iters]]]]
//
// As the reference: N lines of x y z are read, the box is
// L = (4 pi / 3)^(1/3) N^(1/3), the mesh M^3 = 64^3, the assignment order 6,
// kappa 0 (Coulomb), unit charges; the same lines are printed (L, Gew, rc,
// the mesh spacings, PP and PM elapsed microseconds per call over iters = 10
// calls, the PP accelerations of particles 0..9, of particle 467 and of the
// last ten).  Times are HIP events on the library's stream around iters calls
// after one untimed call; the PM Green function is built once, untimed.
// A C caller of libmdforce with no Python or torch: device buffers come from
// hipMalloc here.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "mdforce.h"

#define CK(call)                                                                         \
    do {                                                                                 \
        int r_ = (call);                                                                 \
        if (r_ != MDF_OK) {                                                              \
            fprintf(stderr, "%s failed (%d): %s\n", #call, r_, mdf_last_error());        \
            return 1;                                                                    \
        }                                                                                \
    } while (0)
#define HK(call)                                                                         \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) {                                                          \
            fprintf(stderr, "%s: %s\n", #call, hipGetErrorString(e_));                   \
            return 1;                                                                    \
        }                                                                                \
    } while (0)

static int usage() {
    fprintf(stderr, "usage: mdf_bench N positions_file Gew rc [kappa [M [fp64|mixed [iters]]]]\n");
    return 2;
}

int main(int argc, char **argv) {
    if (argc < 5) return usage();
    const int N = atoi(argv[1]);
    const double Gew = atof(argv[3]);
    const double rc = atof(argv[4]);
    const double kappa = argc > 5 ? atof(argv[5]) : 0.0;
    const int M = argc > 6 ? atoi(argv[6]) : 64;
    const bool mixed = argc > 7 && strcmp(argv[7], "mixed") == 0;
    if (argc > 7 && !mixed && strcmp(argv[7], "fp64") != 0) return usage();
    const int iters = argc > 8 ? atoi(argv[8]) : 10;
    if (N < 1 || M < 8 || iters < 1) return usage();
    FILE *file = fopen(argv[2], "r");
    if (!file) {
        fprintf(stderr, "cannot open %s\n", argv[2]);
        return 1;
    }
    std::vector<double> pos(3 * (size_t)N);
    for (size_t i = 0; i < pos.size(); ++i)
        if (fscanf(file, "%lf", &pos[i]) != 1) {
            fprintf(stderr, "%s holds fewer than %d x y z lines\n", argv[2], N);
            fclose(file);
            return 1;
        }
    fclose(file);

    const double PI = 3.14159265358979323846;
    const double L = cbrt(4.0 * PI / 3.0) * cbrt((double)N);
    const double h = L / M;
    printf("L = %3.12lf\n", L);
    printf("Gew = %.4lf\n", Gew);
    printf("rc = %1.4lf\n", rc);
    printf("hx = %lf, hy = %lf, hz = %lf\n", h, h, h);

    mdf_p3m_params prm;
    memset(&prm, 0, sizeof prm);
    prm.n_max = N;
    prm.mesh[0] = prm.mesh[1] = prm.mesh[2] = M;
    prm.box[0] = prm.box[1] = prm.box[2] = L;
    prm.kappa = kappa;
    prm.alpha = Gew;
    prm.rc = rc;
    prm.order = 6;
    prm.precision = mixed ? MDF_PREC_MIXED : MDF_PREC_FP64;
    mdf_p3m_ctx *ctx = nullptr;
    CK(mdf_p3m_create(&prm, &ctx));

    hipStream_t s;
    HK(hipStreamCreate(&s));
    double *d_pos, *d_Z, *d_pp, *d_pm, *d_u;
    HK(hipMalloc(&d_pos, 3 * sizeof(double) * N));
    HK(hipMalloc(&d_Z, sizeof(double) * N));
    HK(hipMalloc(&d_pp, 3 * sizeof(double) * N));
    HK(hipMalloc(&d_pm, 3 * sizeof(double) * N));
    HK(hipMalloc(&d_u, 2 * sizeof(double)));
    HK(hipMemcpyAsync(d_pos, pos.data(), 3 * sizeof(double) * N, hipMemcpyHostToDevice, s));
    std::vector<double> ones(N, 1.0);
    HK(hipMemcpyAsync(d_Z, ones.data(), sizeof(double) * N, hipMemcpyHostToDevice, s));
    CK(mdf_p3m_build_green(ctx, 6, 3, 3, 3, nullptr, nullptr, s));

    hipEvent_t e0, e1;
    HK(hipEventCreate(&e0));
    HK(hipEventCreate(&e1));
    float ms = 0.0f;
    CK(mdf_pp_force(ctx, d_pos, N, d_pp, d_u, s));
    HK(hipEventRecord(e0, s));
    for (int it = 0; it < iters; ++it) CK(mdf_pp_force(ctx, d_pos, N, d_pp, d_u, s));
    HK(hipEventRecord(e1, s));
    HK(hipEventSynchronize(e1));
    HK(hipEventElapsedTime(&ms, e0, e1));
    printf("PP elapsed time = %lf micro-seconds\n", 1000.0 * ms / iters);
    CK(mdf_pm_force(ctx, d_pos, d_Z, N, d_pm, d_u + 1, nullptr, nullptr, nullptr, nullptr, s));
    HK(hipEventRecord(e0, s));
    for (int it = 0; it < iters; ++it)
        CK(mdf_pm_force(ctx, d_pos, d_Z, N, d_pm, d_u + 1, nullptr, nullptr, nullptr, nullptr, s));
    HK(hipEventRecord(e1, s));
    HK(hipEventSynchronize(e1));
    HK(hipEventElapsedTime(&ms, e0, e1));
    printf("PM elapsed time = %lf micro-seconds\n", 1000.0 * ms / iters);

    std::vector<double> acc(3 * (size_t)N);
    HK(hipMemcpy(acc.data(), d_pp, 3 * sizeof(double) * N, hipMemcpyDeviceToHost));
    for (int i = 0; i < 10 && i < N; ++i)
        printf("%d, %.8lf, %.8lf, %.8lf\n", i, acc[3 * i], acc[3 * i + 1], acc[3 * i + 2]);
    printf("----------------\n");
    const int chk = 467;
    if (chk < N) printf("%d, %.8lf, %.8lf, %.8lf\n", chk, acc[3 * chk], acc[3 * chk + 1], acc[3 * chk + 2]);
    printf("----------------\n");
    for (int i = N > 10 ? N - 10 : 0; i < N; ++i) {
        for (int j = 0; j < 3; ++j) {
            printf("%.8lf ", acc[3 * i + j]);
            if (j == 2) printf("\n");
        }
    }
    HK(hipFree(d_pos));
    HK(hipFree(d_Z));
    HK(hipFree(d_pp));
    HK(hipFree(d_pm));
    HK(hipFree(d_u));
    HK(hipEventDestroy(e0));
    HK(hipEventDestroy(e1));
    HK(hipStreamDestroy(s));
    CK(mdf_p3m_destroy(ctx));
    return 0;
}
