/*
 * mdforce — MI355X (gfx950) P3M Yukawa force engine, C ABI.
 *
 * Every function returns 0 on success and a negative code on error; it never
 * aborts.  mdf_last_error() returns a thread-local message for the last
 * failure.  Array pointers are DEVICE pointers owned by the caller (e.g.
 * torch tensors' data_ptr()); `stream` is a hipStream_t (NULL = default
 * stream).  Calls are asynchronous on that stream.  A context owns its own
 * scratch (cell lists, meshes, spectra, Green function) and allocates nothing
 * after mdf_p3m_create.  One context per stream; contexts are not re-entrant.
 *
 * Layouts (the reference's, yukawa_MD_p3m_ord6.py:107-118):
 *   pos, vel, acc : (n, 3) double, row-major (x, y, z per particle)
 *   Z             : (n,) double charges
 *   mesh          : (Mz, My, Mx) double, x fastest
 *   G_k           : (Mz, My, Mx) double in fftshift order (yukawa_gf_opt.py:10,94)
 */
#ifndef MDFORCE_H
#define MDFORCE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MDF_ABI_VERSION 1

/* Error codes. */
#define MDF_OK 0
#define MDF_EINVAL (-1)  /* bad argument */
#define MDF_ENOMEM (-2)  /* device allocation failed */
#define MDF_EHIP (-3)    /* HIP runtime error */
#define MDF_ESTATE (-4)  /* call out of order (e.g. no Green function set) */

/* Arithmetic of the force path. */
#define MDF_PREC_FP64 0  /* plain: FP64 everywhere */
#define MDF_PREC_MIXED 1 /* FP32 pair math / assignment / FFT, FP64 mesh sums and positions */

typedef struct mdf_p3m_ctx mdf_p3m_ctx;

typedef struct {
    int n_max;        /* largest particle count passed to this context */
    int mesh[3];      /* Mx, My, Mz (powers of two, 8..512) */
    double box[3];    /* Lx, Ly, Lz */
    double kappa;     /* Yukawa screening (kappa) */
    double alpha;     /* Ewald splitting parameter (the reference's G / G_ew) */
    double rc;        /* PP cutoff; needs rc <= L/3 per axis */
    int order;        /* charge-assignment order; 6 is the only one supported */
    int precision;    /* MDF_PREC_* */
} mdf_p3m_params;

/* Library identity. */
int mdf_abi_version(void);
const char *mdf_last_error(void);

/* Context lifetime.  Allocates every device buffer the force path needs. */
int mdf_p3m_create(const mdf_p3m_params *params, mdf_p3m_ctx **out);
int mdf_p3m_destroy(mdf_p3m_ctx *ctx);

/* Green function.
 * set_green: upload G_k (device pointer, full (Mz,My,Mx) array in the
 *   reference's fftshift order, as returned by gf_opt, yukawa_gf_opt.py:5-98).
 * build_green: compute the same optimal influence function on the device
 *   (yukawa_gf_opt.py:5-98) with assignment order p and alias ranges m_max;
 *   if G_k_out is not NULL the full shifted array is also written there, and
 *   if A_pm_out is not NULL the error sum A_pm (yukawa_gf_opt.py:96) is
 *   written to that device double. */
int mdf_p3m_set_green(mdf_p3m_ctx *ctx, const double *G_k, void *stream);
int mdf_p3m_build_green(mdf_p3m_ctx *ctx, int p, int mx_max, int my_max, int mz_max,
                        double *G_k_out, double *A_pm_out, void *stream);

/* Mesh error functional for the error estimate: Q_sum_out (device double) <-
 * sum over k != 0 of [sum_m |k_m|^2 G(k_m)^2 - A_k], with G the optimal
 * influence function of the same p and alias ranges and A_k the k term of
 * A_pm (yukawa_gf_opt.py:96, whose A_pm the reference returns unused).  For N
 * random unit charges in volume V the RMS mesh force error is
 * sqrt(N Q_sum) / V (Hockney-Eastwood).  Writes no Green function. */
int mdf_p3m_pm_error_sum(mdf_p3m_ctx *ctx, int p, int mx_max, int my_max, int mz_max, double *Q_sum_out,
                         void *stream);

/* Full P3M force and energy: replaces force_pot (yukawa_p3m_ord6.py:6-29).
 *   acc       <- af * (acc_pp + acc_pm)                       (n,3)
 *   energies  <- [U, U_pp, U_pm, U_self] (device, 4 doubles), U = uf*(U_pp+U_pm+U_self),
 *                U_self = -n*alpha/sqrt(pi) (yukawa_p3m_ord6.py:23-24)
 * Optional mirrors of the reference's scratch side effects (NULL to skip):
 *   acc_pp (n,3) = acc_s_r, acc_pm (n,3) = acc_fft, e_x/e_y/e_z (n,) = E_*_p,
 *   rho (Mz,My,Mx) = rho_r. */
int mdf_p3m_force(mdf_p3m_ctx *ctx, const double *pos, const double *Z, int n, double af,
                  double uf, double *acc, double *energies, double *acc_pp, double *acc_pm,
                  double *e_x, double *e_y, double *e_z, double *rho, void *stream);

/* Short-range part alone: replaces particle_particle (yukawa_pp.py:6).
 * acc_pp <- PP acceleration (n,3) (unit charges, as the reference);
 * u_pp   <- U_s_r (device double). */
int mdf_pp_force(mdf_p3m_ctx *ctx, const double *pos, int n, double *acc_pp, double *u_pp,
                 void *stream);

/* Long-range part alone: replaces particle_mesh_fft_r (yukawa_pm_ord6.py:440).
 * acc_pm <- Z*E_p (n,3); u_pm <- U_f (device double); e_x/e_y/e_z, rho optional. */
int mdf_pm_force(mdf_p3m_ctx *ctx, const double *pos, const double *Z, int n, double *acc_pm,
                 double *u_pm, double *e_x, double *e_y, double *e_z, double *rho, void *stream);

/* Velocity-Verlet pieces (velocity_verlet_p3m_ord6.py:5-26).
 * kick_drift: vel += 0.5*acc*dt; pos += vel*dt; if pbc, single wrap per
 *   component: pos > lmax -> pos - L, pos < lmin -> pos + L (:12-19).
 * kick: vel += 0.5*acc*dt; k_sum (device double, may be NULL) <- sum(vel^2). */
int mdf_verlet_kick_drift(mdf_p3m_ctx *ctx, double *pos, double *vel, const double *acc, int n,
                          double dt, int pbc, const double *lmin, const double *lmax,
                          void *stream);
int mdf_verlet_kick(mdf_p3m_ctx *ctx, double *vel, const double *acc, int n, double dt,
                    double *k_sum, void *stream);

/* Velocity rescale (thermostat_p3m_ord6.py:9-23): from k_sum = sum(vel^2)
 * (device), T = k_sum/n; fact = sqrt(T_d/T) if it <= 1999 else
 * sqrt((20 T_d/T - 1)/20); vel *= fact.  t_out (device, may be NULL) <- T. */
int mdf_thermostat_scale(mdf_p3m_ctx *ctx, double *vel, int n, const double *k_sum,
                         double T_desired, int it, double *t_out, void *stream);

/* Density modes n(q) = sum_j exp(-i q x_j) along x, y and z for nq wave
 * numbers q (device, nq doubles): yukawa_MD_p3m_ord6.py:214-219.
 * out: device complex (nq, 3) as interleaved (re, im) doubles. */
int mdf_density_modes(mdf_p3m_ctx *ctx, const double *pos, int n, const double *q, int nq,
                      double *out, void *stream);

/* nsteps NVE Velocity-Verlet steps (the production loop of
 * yukawa_MD_p3m_ord6.py:195-219) with kick/drift, force, kick, the FP64
 * K/U/momentum reductions and n(q,t) fused into the step's kernels.
 *   energies (device, nsteps*4): per step (E, K, U, T) with K = kf*sum(v^2),
 *     T = K/(kf*n).
 *   nqt (device complex, nsteps*nq*3) or NULL when nq == 0.
 *   q_step > 0: the wave numbers are the grid q_m = m*q_step, m = 1..nq (nq <= 64;
 *     q may be NULL), evaluated by recurrence from one sincos per particle and axis;
 *     otherwise q holds nq arbitrary wave numbers.
 * pos/vel/acc are updated in place; acc must hold the force of pos on entry. */
int mdf_md_run(mdf_p3m_ctx *ctx, double *pos, double *vel, double *acc, const double *Z, int n,
               int nsteps, double dt, double af, double uf, double kf, const double *q, int nq,
               double q_step, double *energies, double *nqt, void *stream);

/* Stage timing with HIP events recorded on the call's stream, for
 * measurement only.  mode 1: start (nout = event pool size, 0 = default),
 * mode 3: start with events at the PP stage's start and end only (two a
 * step; only the pp figure is then a stage time), mode 0: stop, mode 2:
 * synchronise and write per-stage milliseconds out[0..7] and interval counts
 * out[8..15] (nout >= 16) for the stages bin, pp, assign, fft, gather,
 * drift, modes, finalize; resets the totals. */
int mdf_p3m_profile(mdf_p3m_ctx *ctx, int mode, double *out, int nout);

/* Verlet pair lists for the PP stage (no reference counterpart: the reference
 * rebuilds its cell list and tests every candidate pair on every call,
 * yukawa_pp.py:44-112).  skin > 0: the context's cells become
 * w = L / floor(L / (rc + skin)) wide (rc + skin <= L/3 per axis), and the
 * lists take the list skin s = w - rc (the largest the cells allow; at most
 * the cap of mdf_p3m_set_list_skin): a step tests candidates only when some
 * particle has moved more than s / 2 since the last test ("build"), writing
 * every pair within rc + s to per-particle lists; the other steps evaluate
 * the listed pairs with the exact cutoff r < rc.  Forces and energies are
 * those of skin = 0 up to summation order.  mdf_p3m_force, mdf_pp_force, the
 * first step of each mdf_md_run and of each mdf_batch_run always build.
 * skin = 0 turns the lists off.  The context is rebuilt in place (the Green
 * function and an attached network carry over; stage profiling stops). */
int mdf_p3m_set_skin(mdf_p3m_ctx *ctx, double skin);
/* Cap on the list skin s (no reference counterpart): s = min(w - rc,
 * list_skin), 0 = no cap (the default).  A smaller skin lists fewer pairs and
 * builds more often.  Keeps the cells; reallocates the lists when s changes.
 * effective (host pointer, may be null): the list skin now in use (0 without
 * lists). */
int mdf_p3m_set_list_skin(mdf_p3m_ctx *ctx, double list_skin, double *effective);
/* host pointer, 4 ints: list builds and list-evaluation steps since the
 * skin was set, list entries per particle, reference cells along x. */
int mdf_p3m_list_stats(mdf_p3m_ctx *ctx, int *out4);

/* Diagnostics: number of PP tiles that needed more than one LDS pass and
 * particles whose stencil fell outside their LDS brick (both handled, only
 * slower), accumulated since create.  host pointer, 2 ints. */
int mdf_p3m_stats(mdf_p3m_ctx *ctx, int *out2);

/* Diagnostics (no reference counterpart): the PP launch configuration the
 * context chose.  host pointer, 4 ints: waves per PP workgroup, LDS tile
 * capacity in particles, pieces of the pair-term table, PP LDS bytes. */
int mdf_p3m_config(mdf_p3m_ctx *ctx, int *out4);

/* Launch choices (no reference counterpart; for the autotuner, which picks
 * them by timing).  pp_waves: waves per PP workgroup, 4, 8, 12 or 16;
 * assign_shape, gather_shape: PM block shapes in sub-cells per axis,
 * 1 = 8x8x8, 2 = 6x6x6, 3 = 4x4x4, 4 = 2x2x2, 5 = 1x2x2, 6 = 1x1x2,
 * 7 = 1x1x1, 8 = 6x6x4, 9 = 6x4x4, 10 = 4x4x4 reading the field mesh from
 * global memory instead of an LDS brick (assignment: 1..4).  0 leaves a choice to the context (its
 * occupancy and LDS rules).  A choice whose LDS does not fit, or whose PP
 * tile would hold fewer than 0.8 of the mean tile, is refused with
 * MDF_EINVAL and the context is unchanged.  The context is rebuilt in place
 * as by mdf_p3m_set_skin (skin, Green function and network carry over).
 * Forces do not depend on these choices beyond summation order. */
int mdf_p3m_set_launch(mdf_p3m_ctx *ctx, int pp_waves, int assign_shape, int gather_shape);
/* host pointer, 6 ints: PP waves, assignment and gather shapes (as above),
 * assignment and gather workgroups per launch, PP LDS tile capacity. */
int mdf_p3m_launch_config(mdf_p3m_ctx *ctx, int *out6);

/* Replica ensembles (no single reference counterpart: the reference runs
 * each seeded replica of yukawa_MD_p3m_ord6.py:195-219 as its own process
 * and averages S(q,w) afterwards, s_of_q_w_avg.py).  A batch owns one
 * context per replica (params[r]: its own box, mesh, kappa, alpha, rc,
 * precision) and n_streams worker streams (0 = 8).
 * mdf_batch_context: the replica's context, e.g. to set or build its Green
 *   function (every replica needs one before mdf_batch_run).
 * mdf_batch_run: nsteps NVE steps of every replica (the mdf_md_run step,
 *   n(q) on the grid q_m = m*2*pi/L_r, m = 1..nq, nq <= 64).  Arrays are
 *   host arrays of nrep device pointers: pos/vel/acc (n[r],3) updated in
 *   place; energies[r] (nsteps*4: E, K, U, T) may be NULL (or the whole
 *   array NULL); nqt[r] (nsteps*nq*3 complex) when nq > 0.  With
 *   graph_steps = K > 0 the steps of all replicas are captured into one
 *   hipGraph of K steps (recaptured when any argument changes) and replayed
 *   nsteps/K times, replicas on different worker streams overlapping; the
 *   remaining steps run eagerly on `stream`.  Calls on one batch must not
 *   overlap. */
typedef struct mdf_batch mdf_batch;
int mdf_batch_create(int nrep, const mdf_p3m_params *params, int n_streams, mdf_batch **out);
int mdf_batch_destroy(mdf_batch *batch);
int mdf_batch_context(mdf_batch *batch, int r, mdf_p3m_ctx **out);
/* Graph shape: 0 (default) = segmented, each stage of the step one launch for
 * all replicas (needs one precision in every replica and no attached
 * learned correction, else falls back to 1); 1 = each replica's own
 * launches on worker streams.  mdf_batch_mode returns the shape of the last captured graph
 * (0 or 1; -1 before any capture). */
int mdf_batch_set_mode(mdf_batch *batch, int mode);
int mdf_batch_mode(mdf_batch *batch);
int mdf_batch_run(mdf_batch *batch, double *const *pos, double *const *vel, double *const *acc,
                  const int *n, int nsteps, double dt, double af, double uf, double kf, int nq,
                  double *const *energies, double *const *nqt, int graph_steps, void *stream);

/* Learned short-range correction (no reference counterpart: a build-side
 * addition that lets the PP cutoff and the mesh shrink).  A pair network
 * (16 features -> 3 x 64 SiLU -> 1, BF16 weights and activations, FP32
 * accumulation on MFMA) predicts a radial force g_ij for every ordered pair
 * with r_ij < rc of the context; particle i gains sum_j g_ij (x_i - x_j)/r_ij.
 * Features: 12 Gaussians exp(-((r - mu0 - k dmu) inv_sigma)^2), the pair's
 * local density (mean of the two PP-cell counts over their mean, minus 1),
 * and the run constants f_kappa, f_alpha, f_rc.  g = out_scale * network.
 * Parameters: mdf_mlp_nparam() FP32 values, packed W1[16][64], b1[64],
 * W2[64][64], b2, W3[64][64], b3, w4[64], b4 (W[k][n]: input k, unit n). */
typedef struct {
    float mu0, dmu, inv_sigma;   /* radial basis */
    float f_kappa, f_alpha, f_rc; /* constant features of the setting */
    float out_scale;             /* output scale (force units) */
} mdf_mlp_meta;
typedef struct mdf_mlp mdf_mlp;
int mdf_mlp_nparam(void);
/* n_max particles, max_pairs ordered pairs (the pair list capacity). */
int mdf_mlp_create(int n_max, int max_pairs, const mdf_mlp_meta *meta, mdf_mlp **out);
int mdf_mlp_destroy(mdf_mlp *mlp);
int mdf_mlp_set_meta(mdf_mlp *mlp, const mdf_mlp_meta *meta);
/* Copy parameters in (device FP32, may be NULL) and/or out (device, may be
 * NULL); reset_adam != 0 clears the optimiser state. */
int mdf_mlp_params(mdf_mlp *mlp, const float *in, float *out, int reset_adam, void *stream);
/* Bins pos with ctx, builds the pair list (r < ctx rc) and runs the network.
 * Optional device outputs: pairs_out (1 + 6*npairs 32-bit words: the count,
 * then per pair i, j (sorted indices), r, u_x, u_y, u_z as float bits),
 * g_out (npairs floats), rho_out (n floats, sorted order), perm_out (n ints:
 * sorted -> caller index).  Synchronises; returns npairs (>= 0) or an error. */
int mdf_mlp_pairs(mdf_mlp *mlp, mdf_p3m_ctx *ctx, const double *pos, int n, int *pairs_out,
                  float *g_out, float *rho_out, int *perm_out, void *stream);
/* f_out (n,3) <- the network's correction force for pos (caller order). */
int mdf_mlp_correction(mdf_mlp *mlp, mdf_p3m_ctx *ctx, const double *pos, int n, double *f_out,
                       void *stream);
/* Loss (1/n) sum_i |F_i - target_i|^2 of the correction F against target
 * (n,3, caller order) and its gradient (device FP32, nparam) from the BF16
 * MFMA backward pass.  build != 0 bins pos and rebuilds the pair list first. */
int mdf_mlp_grad(mdf_mlp *mlp, mdf_p3m_ctx *ctx, const double *pos, const double *target, int n,
                 float *grad_out, double *loss_out, int build, void *stream);
/* One Adam step on the last gradient. */
int mdf_mlp_adam(mdf_mlp *mlp, float lr, float beta1, float beta2, float eps, void *stream);
/* host out2: [pair count of the last list over n particles, overflow flag] */
int mdf_mlp_stats(mdf_mlp *mlp, int n, int *out2);
/* Attach (or detach with NULL) a network to a context: mdf_p3m_force and
 * mdf_md_run then add its correction to the PP force of every evaluation. */
int mdf_p3m_attach_mlp(mdf_p3m_ctx *ctx, mdf_mlp *mlp);

/* ---- S(q, omega) from the density modes n(q, t) ------------------------
 * Replaces n_of_q_w (s_of_q_w.py:3-20) and s_of_q_w_avg_mat
 * (s_of_q_w_avg.py:5-20): per run and q index i < nq, the three axis series
 * n(t) = nqt[run, t, i * q_stride, axis] are transformed along t (FP64
 * mixed-radix FFT, any npd >= 1) and S = |dt * FFT(n)|^2 / (N * T), with T the
 * total time npd * dt of the reference's scripts.
 *   nqt : device complex128 (nrun, npd, nq_total, 3), the reference's n_q_t
 *         of each run stacked (mdf_md_run's nqt layout); q_stride 3 takes the
 *         columns i * 3 as s_of_q_w_avg_mat does, 1 takes every column
 *   out : device double, (nrun, nq, npd) means over the axes when mean_axes
 *         != 0 (s_of_q_w_avg_mat's sqw_mat), else (nrun, nq, npd, 3) (each
 *         series as n_of_q_w returns it)
 *   work: device scratch of at least mdf_sqw_workspace bytes
 * Asynchronous on stream; allocates nothing. */
int mdf_sqw_workspace(int nrun, int npd, int nq, size_t *bytes);
int mdf_sqw(const double *nqt, int nrun, int npd, int nq_total, int q_stride, int nq, double N, double dt,
            double T, int mean_axes, double *out, void *work, size_t work_bytes, void *stream);

/* Allocation-free S(q, omega) (no counterpart in the reference): mdf_sqw for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results as mdf_sqw without the workspace: the stages run in a static device
 * array of MDF_SQW_STATIC_BYTES that the library holds on each GPU it runs on
 * until the process ends, so a call takes at most MDF_SQW_STATIC_BYTES / 32
 * series samples (nrun * nq * 3 * npd); larger calls return MDF_EINVAL.
 * Calls no hipMalloc/hipFree.  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (other streams, other threads) race. */
#define MDF_SQW_STATIC_BYTES (128u << 20)
int mdf_sqw_noalloc(const double *nqt, int nrun, int npd, int nq_total, int q_stride, int nq, double N,
                    double dt, double T, int mean_axes, double *out, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* MDFORCE_H */
