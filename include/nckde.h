/*
 * nckde.h -- C-ABI of the MI355X (gfx950) kernel-density engine.
 *
 * One nckde handle is one NcmStatsDistKDE or NcmStatsDistVKDE object with a
 * Gauss or Student-t kernel.  Its prepared state (whitened sample, per-kernel
 * Cholesky factors, log-normalisations, weights, evaluation factors) lives in
 * HBM and persists until the next prepare or reset.  All signatures use plain
 * pointers and sizes; host buffers are row-major doubles unless stated.
 *
 * Reference interfaces (NumCosmo, paths relative to its source tree):
 *   NcmStatsDistClass vtable ....... numcosmo/ncm/stats/ncm_stats_dist.h:44-65
 *   public NcmStatsDist API ........ numcosmo/ncm/stats/ncm_stats_dist.h:87-141
 *   NcmStatsDistKDE API ............ numcosmo/ncm/stats/ncm_stats_dist_kde.h:74-86
 *   NcmStatsDistVKDE API ........... numcosmo/ncm/stats/ncm_stats_dist_vkde.h:53-62
 *   NcmStatsDistKernel{Gauss,ST} ... numcosmo/ncm/stats/ncm_stats_dist_kernel.h:57-90
 *
 * Errors: every function returns NCKDE_OK (0) or a negative status; the
 * message of the last failure on the calling thread is nckde_last_error().
 * Where the reference aborts with g_error, the same text is reported here.
 *
 * Threading: calls on one handle may come from several host threads; an
 * internal lock serialises them.  Host-buffer calls return after their work
 * is complete.  nckde_eval_m2lnp_device is asynchronous on the given stream;
 * a later call on a different stream first waits for it.
 */
#ifndef NCKDE_H
#define NCKDE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nckde nckde;

/* NcmStatsDistKDE / NcmStatsDistVKDE */
enum { NCKDE_KDE = 0, NCKDE_VKDE = 1 };
/* NcmStatsDistKernelGauss / NcmStatsDistKernelST */
enum { NCKDE_KERNEL_GAUSS = 0, NCKDE_KERNEL_ST = 1 };
/* NcmStatsDistCV (ncm_stats_dist.h:77-85) */
enum { NCKDE_CV_NONE = 0, NCKDE_CV_SPLIT = 1, NCKDE_CV_SPLIT_NOFIT = 2, NCKDE_CV_LOO = 3 };
/* NcmStatsDistKDECovType (ncm_stats_dist_kde.h:64-72): sample, fixed, robust diagonal (Qn scales) and
 * robust OGK covariances (ncm_stats_vec.c:1834-2064) */
enum { NCKDE_COV_SAMPLE = 0, NCKDE_COV_FIXED = 1, NCKDE_COV_ROBUST_DIAG = 2, NCKDE_COV_ROBUST = 3 };
/* arithmetic of the pairwise distance and kernel sums in eval: FP64; MIXED
 * (FP16 hi/lo split products on the matrix cores, FP32 sums); FP32 (the
 * FP32 rung of the precision ladder: single-precision distances and sums) */
enum { NCKDE_PREC_FP64 = 0, NCKDE_PREC_MIXED = 1, NCKDE_PREC_FP32 = 2 };
/* NcmNNLSUMethod (ncm_nnls.h:51-60): the unconstrained solve of the active-set
 * NNLS.  NORMAL: normal equations by Cholesky, else the symmetric solve, else
 * QR; NORMAL_LU: symmetric solve, else QR; QR: least squares on A_P. */
enum { NCKDE_NNLS_NORMAL = 0, NCKDE_NNLS_NORMAL_LU = 1, NCKDE_NNLS_QR = 2 };

enum
{
  NCKDE_OK           = 0,
  NCKDE_EINVAL       = -1,  /* bad argument */
  NCKDE_EHIP         = -2,  /* HIP runtime failure (no device, launch error) */
  NCKDE_ESTATE       = -3,  /* call out of order (e.g. eval before prepare) */
  NCKDE_ENOMEM       = -4,  /* device allocation failed */
  NCKDE_ESAMPLE      = -5,  /* sample too small (reference g_error) */
  NCKDE_EUNSUPPORTED = -6,  /* option outside the implemented path */
  NCKDE_ENUMERIC     = -7   /* numerical failure (non-PD fixed covariance, ...) */
};

/* Message of the last failing call on this thread. */
const char *nckde_last_error (void);

/* ---- lifetime ---------------------------------------------------------- */

/* ncm_stats_dist_kde_new / ncm_stats_dist_vkde_new (ncm_stats_dist_kde.h:74,
 * ncm_stats_dist_vkde.h:53) with ncm_stats_dist_kernel_gauss_new /
 * ncm_stats_dist_kernel_st_new (ncm_stats_dist_kernel_gauss.h:41,
 * ncm_stats_dist_kernel_st.h:41).  nu is ignored for the Gauss kernel. */
int nckde_create (nckde **out, int kind, int kernel, double nu, unsigned dim, int cv_type);

/* ncm_stats_dist_free (ncm_stats_dist.h:88) */
int nckde_destroy (nckde *h);

/* ---- properties -------------------------------------------------------- */

/* ncm_stats_dist_set_over_smooth (ncm_stats_dist.h:100) */
int nckde_set_over_smooth (nckde *h, double over_smooth);
/* ncm_stats_dist_set_shrink (ncm_stats_dist.h:106) */
int nckde_set_shrink (nckde *h, double shrink);
/* ncm_stats_dist_vkde_set_local_frac (ncm_stats_dist_vkde.h:58) */
int nckde_set_local_frac (nckde *h, double local_frac);
/* ncm_stats_dist_vkde_set_use_rot_href (ncm_stats_dist_vkde.h:61) */
int nckde_set_use_rot_href (nckde *h, int use_rot_href);
/* ncm_stats_dist_kde_set_cov_type (ncm_stats_dist_kde.h:82) */
int nckde_set_cov_type (nckde *h, int cov_type);
/* ncm_stats_dist_kde_set_cov_fixed (ncm_stats_dist_kde.h:85); cov is d x d */
int nckde_set_cov_fixed (nckde *h, const double *cov);
/* ncm_stats_dist_kde_set_nearPD_maxiter (ncm_stats_dist_kde.h:79) */
int nckde_set_nearPD_maxiter (nckde *h, unsigned maxiter);
/* ncm_stats_dist_set_use_threads / get (ncm_stats_dist.h:115-116): stored;
 * the device path does not depend on it */
int nckde_set_use_threads (nckde *h, int use_threads);
int nckde_get_use_threads (nckde *h, int *use_threads);
/* ncm_stats_dist_get_shrink (ncm_stats_dist.h:107) */
int nckde_get_shrink (nckde *h, double *shrink);
/* ncm_stats_dist_vkde_get_local_frac (ncm_stats_dist_vkde.h:59) */
int nckde_get_local_frac (nckde *h, double *local_frac);
/* ncm_stats_dist_vkde_get_use_rot_href (ncm_stats_dist_vkde.h:62) */
int nckde_get_use_rot_href (nckde *h, int *use_rot_href);
/* ncm_stats_dist_kde_get_nearPD_maxiter (ncm_stats_dist_kde.h:80) */
int nckde_get_nearPD_maxiter (nckde *h, unsigned *maxiter);
/* ncm_stats_dist_kde_get_cov_type (ncm_stats_dist_kde.h:83) */
int nckde_get_cov_type (nckde *h, int *cov_type);
/* ncm_stats_dist_kde_peek_cov_fixed (ncm_stats_dist_kde.h:86): copies the d x d
 * matrix; NCKDE_ESTATE when none was set (the reference returns NULL) */
int nckde_get_cov_fixed (nckde *h, double *cov);
/* ncm_stats_dist_get_over_smooth (ncm_stats_dist.h:101); the CV_SPLIT,
 * CV_SPLIT_NOFIT and CV_LOO fits change it */
int nckde_get_over_smooth (nckde *h, double *over_smooth);
/* ncm_stats_dist_set_split_frac / get (ncm_stats_dist.h:103-104); the
 * reference asserts 0.01 <= split_frac <= 1, this returns NCKDE_EINVAL */
int nckde_set_split_frac (nckde *h, double split_frac);
int nckde_get_split_frac (nckde *h, double *split_frac);
/* ncm_stats_dist_set_print_fit / get (ncm_stats_dist.h:109-110): one stdout
 * line per objective evaluation of the cross-validation fits */
int nckde_set_print_fit (nckde *h, int print_fit);
int nckde_get_print_fit (nckde *h, int *print_fit);
/* ncm_stats_dist_set_cv_type / get (ncm_stats_dist.h:112-113) */
int nckde_set_cv_type (nckde *h, int cv_type);
int nckde_get_cv_type (nckde *h, int *cv_type);
/* No reference counterpart (measurement hook): objective evaluations (SPLIT:
 * IM + NNLS fits) and wall seconds of the last cross-validation fit */
int nckde_get_cv_stats (nckde *h, unsigned *n_objective, double *seconds);
/* No reference counterpart: the check of the mixed-precision factors made
 * when they are built (at the first mixed eval after a prepare).  Two probe
 * points 1.5 kernel widths from each of the 64 kernels most exposed to the
 * rounding of the split products (narrow kernels far from the sample mean)
 * are evaluated in both precisions; *max_rel_err is the largest
 * |m2lnp_mixed - m2lnp_fp64| / max(|m2lnp_fp64|, 1) found, and *fallback is
 * 1 when it exceeded the handle's tolerance (5e-7 unless set with
 * nckde_set_mixed_tolerance), in which case NCKDE_PREC_MIXED evals of this
 * handle run the FP64 path until the next prepare.  Builds the factors if
 * needed; either output may be NULL. */
int nckde_get_mixed_check (nckde *h, double *max_rel_err, int *fallback);
/* No reference counterpart (tuning hook): force the number of kernel-range
 * splits of this handle's evals (each split is one grid row of blocks whose
 * partial sums are merged in a fixed order); 0 restores the choice from the
 * tuned table (numcosmo_amd/tune.py), else the default rule.  Results depend
 * on the split count only in the last bits. */
int nckde_set_eval_split (nckde *h, unsigned nsplit);
/* No reference counterpart: replay of captured evals.  With on != 0, an eval
 * repeated with the same shape (points, precision, linear/log form), the
 * same device buffers and the same non-null stream is captured into a HIP
 * graph at its second call and replayed as one graph launch afterwards
 * (host-buffer evals reuse the handle's point and output buffers, so
 * repeated evals of m points qualify).  A prepare, a change of the weights
 * or of the bandwidth, or a factor rebuild invalidates the graphs.  Results
 * are bit-identical to the eager launches.  Off by default. */
int nckde_set_graphs (nckde *h, int on);
/* No reference counterpart: the probe error above which this handle's mixed
 * evals run in FP64 (default 5e-7, half the 1e-6 bar).  A caller with a
 * looser bar (the APES walker holds its transition densities to 3e-6) may
 * raise it; an existing probe result is judged again. */
int nckde_set_mixed_tolerance (nckde *h, double tol);
/* ncm_nnls_set_umethod / get (ncm_nnls.h:66-67) of the object's weight fit,
 * which the reference fixes to NCM_NNLS_UMETHOD_NORMAL (ncm_stats_dist.c:861);
 * DGELSD and GSL give NCKDE_EUNSUPPORTED */
int nckde_set_nnls_umethod (nckde *h, int umethod);
int nckde_get_nnls_umethod (nckde *h, int *umethod);
/* No reference counterpart: at most max_steps refinement steps after a
 * Cholesky solve of the normal equations whose pivots spread by more than
 * 1e10 (squared ratio of the extreme diagonal entries of the factor):
 * delta = (A_P^T A_P)^{-1} A_P^T (f - A_P x), x += delta, a step kept only if
 * the residual norm does not grow.  The default is 0 (no refinement, as the
 * reference's solver); at most 16. */
int nckde_set_nnls_refine (nckde *h, unsigned max_steps);
int nckde_get_nnls_refine (nckde *h, unsigned *max_steps);
/* No reference counterpart (diagnostic): unconstrained solves of the last
 * prepare_interp by rung -- counts[0] Cholesky, [1] symmetric solve,
 * [2] QR -- and counts[3] refinement steps taken */
int nckde_get_nnls_rungs (nckde *h, unsigned counts[4]);

/* ---- sample ------------------------------------------------------------ */

/* ncm_stats_dist_add_obs (ncm_stats_dist.h:130): x (d doubles) is copied */
int nckde_add_obs (nckde *h, const double *x);
/* n rows of a row-major n x d array with leading dimension ld (>= d) */
int nckde_add_obs_batch (nckde *h, const double *X, size_t n, size_t ld);
/* ncm_stats_dist_reset (ncm_stats_dist.h:141) */
int nckde_reset (nckde *h);

/* ---- prepare ----------------------------------------------------------- */

/* ncm_stats_dist_prepare (ncm_stats_dist.h:119) -> class prepare
 * (ncm_stats_dist.c:703-789): global covariance, whitening, kNN local
 * covariances, batched FP64 Cholesky, uniform weights. */
int nckde_prepare (nckde *h);

/* ncm_stats_dist_prepare_interp (ncm_stats_dist.h:120) -> class
 * prepare_interp (ncm_stats_dist.c:878-1094): interpolation matrix and
 * active-set NNLS weight fit on the device.  m2lnp has n = sample size
 * entries.  May drop sample points (the reference's m2lnp range cut). */
int nckde_prepare_interp (nckde *h, const double *m2lnp, size_t n);

/* class compute_IM (ncm_stats_dist.h:54): IM is n_obs x n_kernels,
 * IM[j][i] = K_i(theta_j) / (u_i h^d).  Valid after prepare. */
int nckde_compute_IM (nckde *h, double *IM);

/* ---- evaluation -------------------------------------------------------- */

/* ncm_stats_dist_eval_m2lnp (ncm_stats_dist.h:123) over m points (rows of
 * X, leading dimension ld), out[m].  precision: NCKDE_PREC_FP64 or
 * NCKDE_PREC_MIXED (FP16-split MFMA distances, FP32 kernel sums). */
int nckde_eval_m2lnp (nckde *h, const double *X, size_t m, size_t ld, double *out, int precision);

/* ncm_stats_dist_eval (ncm_stats_dist.h:122) over m points. */
int nckde_eval (nckde *h, const double *X, size_t m, size_t ld, double *out, int precision);

/* Device-pointer form of nckde_eval_m2lnp for callers whose points already
 * live in HBM: dX is m x d row-major (device), dout m doubles (device),
 * stream is a hipStream_t (NULL: the handle's stream).  Asynchronous. */
int nckde_eval_m2lnp_device (nckde *h, const double *dX, size_t m, double *dout, int precision, void *stream);

/* ---- sampling ---------------------------------------------------------- */

/* ncm_stats_dist_kernel_choose (ncm_stats_dist.h:125) for m host-supplied
 * uniforms u[m] in [0,1) (gsl_ran_flat(0,1) returns u itself); idx[m]. */
int nckde_kernel_choose (nckde *h, const double *u, size_t m, uint32_t *idx);

/* ncm_stats_dist_sample (ncm_stats_dist.h:126) for m draws: u[m] uniforms,
 * z[m*d] standard normals, chi2[m] chi-squared(nu) draws (Student-t only;
 * may be NULL for Gauss).  out[m*d] row-major; idx[m] (may be NULL) gets the
 * chosen kernels. */
int nckde_sample (nckde *h, const double *u, const double *z, const double *chi2,
                  size_t m, double *out, uint32_t *idx);

/* ---- peeks (valid after prepare) --------------------------------------- */

/* ncm_stats_dist_get_dim (ncm_stats_dist.h:95) */
int nckde_get_dim (nckde *h, unsigned *dim);
/* ncm_stats_dist_get_sample_size (ncm_stats_dist.h:96) */
int nckde_get_sample_size (nckde *h, size_t *n);
/* ncm_stats_dist_get_n_kernels (ncm_stats_dist.h:97) */
int nckde_get_n_kernels (nckde *h, size_t *n);
/* number of points added (the reference's "N" property) */
int nckde_get_n_samples (nckde *h, size_t *n);
/* ncm_stats_dist_get_href (ncm_stats_dist.h:98) */
int nckde_get_href (nckde *h, double *href);
/* ncm_stats_dist_get_rnorm (ncm_stats_dist.h:128): squared NNLS residual */
int nckde_get_rnorm (nckde *h, double *rnorm);
/* ncm_stats_dist_get_lnnorm (ncm_stats_dist.h:136): includes d ln(href) */
int nckde_get_lnnorm (nckde *h, size_t i, double *lnnorm);
/* ncm_stats_dist_peek_weights (ncm_stats_dist.h:137): n_kernels doubles */
int nckde_get_weights (nckde *h, double *w);
/* ncm_stats_dist_peek_cov_decomp (ncm_stats_dist.h:133): d x d, U in the
 * upper triangle (cov = U^T U); the strict lower triangle holds what LAPACK
 * leaves there (the covariance's lower triangle). */
int nckde_get_cov_decomp (nckde *h, size_t i, double *U);
/* ncm_stats_dist_get_Ki (ncm_stats_dist.h:139, ncm_stats_dist.c:1810-1828):
 * kernel i's centre y_i[d] (its sample row), covariance cov_i[d*d] =
 * href^2 U_i^T U_i, normalisation n_i = exp(lnnorm_i) (lnnorm as
 * nckde_get_lnnorm) and weight w_i.  Outputs are the caller's buffers. */
int nckde_get_Ki (nckde *h, size_t i, double *y_i, double *cov_i, double *n_i, double *w_i);
/* ncm_stats_dist_peek_full_cov (ncm_stats_dist.h:135) */
int nckde_get_full_cov (nckde *h, double *cov);
/* ncm_stats_dist_peek_full_cov_decomp (ncm_stats_dist.h:134) */
int nckde_get_full_cov_decomp (nckde *h, double *U);
/* ncm_stats_dist_peek_sample_array (ncm_stats_dist.h:132): n_samples x d */
int nckde_get_sample (nckde *h, double *X);
/* whitened sample rows y_j (U^T y_j = theta_j), n_obs x d */
int nckde_get_invUsample (nckde *h, double *Y);
/* VKDE: number of neighbours k and the n_kernels x k neighbour lists in
 * (squared distance, index) order */
int nckde_get_knn (nckde *h, size_t *k, int32_t *idx);

/* No reference counterpart (diagnostic hook for checking a fit): the last
 * NNLS fit's solution x (n_kernels, before shrink and normalisation), its
 * gradient g = A^T (f - A x) (n_kernels) and residual f - A x (sample size),
 * with A = diag(1/f) IM and f_j = exp(-(m2lnp_j - min m2lnp)/2), as the
 * reference's ncm_nnls_solve leaves them (ncm_nnls.c:767-871).  Any output
 * may be NULL.  NCKDE_ESTATE when the last prepare fitted no NNLS. */
int nckde_get_nnls_state (nckde *h, double *x, double *g, double *r);

/* No reference counterpart: the evals of several handles at once, results
 * bit-identical to one nckde_eval_m2lnp per handle.  Evaluation i takes
 * m[i] host points X[i] (row stride ld[i] doubles; ld == NULL means d) into
 * out[i].  In mixed precision the evals of handles with the same MFMA
 * dimension and kernel type run as one launch per pass over a segment table
 * (the served stream's many small requests fill the GPU together); other
 * evals (other precisions, and d > 16 handles of 1024 kernels or more, whose
 * evals start with a pilot pass) run one by one.  A handle may appear once. */
int nckde_eval_m2lnp_multi (nckde *const *hs, const double *const *X, const size_t *m, const size_t *ld,
                            double *const *out, size_t count, int precision);

/* ---- allocation-free eval ---------------------------------------------- */

/* No reference counterpart.  For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): nckde_eval_m2lnp_device with every scratch buffer (point work
 * and split partials) taken from a fixed device array of
 * nckde_noalloc_scratch_bytes() bytes (64 MiB) that each GPU holds for the
 * life of the process; the call allocates nothing.  The caller passes device
 * points and output as for nckde_eval_m2lnp_device, on the handle's device.
 * Limits: NCKDE_EINVAL when m points need more scratch than the array holds
 * (about 6e5 points in mixed precision at d <= 16, 4e6 in FP64; evaluate in
 * chunks); NCKDE_ESTATE for NCKDE_PREC_MIXED until the mixed factors are
 * built (an allocating mixed eval or nckde_get_mixed_check after the
 * prepare builds them).  All calls on one GPU share the array, whatever the
 * handle: calls that overlap in time on one GPU race, so order them (one
 * stream, or a synchronisation between them). */
int nckde_eval_m2lnp_device_noalloc (nckde *h, const double *dX, size_t m, double *dout, int precision,
                                     void *stream);
size_t nckde_noalloc_scratch_bytes (void);

/* ---- serialisation ------------------------------------------------------
 * The reference serialises an object's configuration with NcmSerialize
 * (ncm_serialize.h; exercised by test_ncm_stats_dist.c:955-957: a
 * serialised-then-restored object fed the same sample gives the same fit).
 * nckde_serialize writes the configuration and the sample, and with
 * with_state != 0 also the prepared state (factors, weights, neighbour
 * lists), so that a served session can be checkpointed and resumed without
 * a new prepare.  buf == NULL: only *len (the size needed) is set.  The
 * format is native-endian binary, versioned by its 8-byte magic. */
int nckde_serialize (nckde *h, int with_state, void *buf, size_t cap, size_t *len);
int nckde_deserialize (nckde **out, const void *buf, size_t len);

/* ---- measurement --------------------------------------------------------
 * No counterpart in the reference.  With profiling on, every eval records HIP
 * events around its main evaluation kernel (k_eval_f64, or the first pass of
 * the mixed-precision kernel); nckde_get_kernel_ms waits for the last one and
 * returns its duration. */
int nckde_set_profiling (nckde *h, int on);
int nckde_get_kernel_ms (nckde *h, float *ms);

/* ---- APES ensemble sampler ----------------------------------------------
 *
 * NcmFitESMCMCWalkerAPES (numcosmo/ncm/fit/ncm_fit_esmcmc_walker_apes.h:
 * 51-105, .c:509-919) with the ensemble loop of NcmFitESMCMC
 * (ncm_fit_esmcmc.c:2136-2288).  Each half-ensemble is moved with a VKDE of
 * the other half: one prepare_interp, one batched sample and one batched eval
 * of both transition densities per half.  The likelihood is a host callback
 * over a batch of m points (the reference calls NcmFit per walker).  The
 * jump uniforms, random-walk decisions and steps come from a seeded host
 * generator in the reference's per-walker order; the variates of the KDE
 * draws come from a counter-based device generator (Philox4x32-10) keyed by
 * the same seed.  Neither stream is GSL's mt19937.
 */
typedef struct nckde_apes nckde_apes;

/* NcmFitESMCMCWalkerAPESMethod (ncm_fit_esmcmc_walker_apes.h:51-57) */
enum { NCKDE_APES_METHOD_KDE = 0, NCKDE_APES_METHOD_VKDE = 1 };
/* NcmFitESMCMCWalkerAPESKType (ncm_fit_esmcmc_walker_apes.h:68-75) */
enum { NCKDE_APES_KTYPE_CAUCHY = 0, NCKDE_APES_KTYPE_ST3 = 1, NCKDE_APES_KTYPE_GAUSS = 2 };

/* m2lnL[j] = -2 ln L(theta[j*d .. j*d+d-1]) for j < m; returns 0 on success
 * (non-finite values reject the proposal, as ncm_fit_esmcmc.c:2183-2200). */
typedef int (*nckde_m2lnL_batch_fn) (const double *theta, size_t m, unsigned d, double *m2lnL, void *user);

/* ncm_fit_esmcmc_walker_apes_new_full (ncm_fit_esmcmc_walker_apes.h:78) */
int nckde_apes_create (nckde_apes **out, unsigned nwalkers, unsigned dim, int method, int ktype, uint64_t seed);
int nckde_apes_destroy (nckde_apes *a);
/* ncm_fit_esmcmc_walker_apes_set_over_smooth (ncm_fit_esmcmc_walker_apes.h:85) */
int nckde_apes_set_over_smooth (nckde_apes *a, double over_smooth);
/* ncm_fit_esmcmc_walker_apes_set_shrink (ncm_fit_esmcmc_walker_apes.h:86) */
int nckde_apes_set_shrink (nckde_apes *a, double shrink);
/* ncm_fit_esmcmc_walker_apes_set_local_frac (ncm_fit_esmcmc_walker_apes.h:105) */
int nckde_apes_set_local_frac (nckde_apes *a, double local_frac);
/* ncm_fit_esmcmc_walker_apes_set_random_walk_prob (ncm_fit_esmcmc_walker_apes.h:87) */
int nckde_apes_set_random_walk_prob (nckde_apes *a, double prob);
/* ncm_fit_esmcmc_walker_apes_use_interp (ncm_fit_esmcmc_walker_apes.h:97) */
int nckde_apes_set_use_interp (nckde_apes *a, int use_interp);
/* ncm_fit_esmcmc_walker_apes_set_exploration (ncm_fit_esmcmc_walker_apes.h:110) */
int nckde_apes_set_exploration (nckde_apes *a, unsigned steps);
/* eval arithmetic of the transition densities (NCKDE_PREC_*) */
int nckde_apes_set_precision (nckde_apes *a, int precision);
/* parameter bounds (ncm_mset_fparam_valid_bounds); default (-inf, inf) */
int nckde_apes_set_bounds (nckde_apes *a, const double *lb, const double *ub);
/* initial walkers: theta is nwalkers x dim, m2lnL their -2 ln L */
int nckde_apes_init (nckde_apes *a, const double *theta, const double *m2lnL);
/* ncm_fit_esmcmc_run (ncm_fit_esmcmc.h:78): nsteps ensemble steps; chain
 * (may be NULL) receives nsteps x nwalkers rows of (m2lnL, theta) */
int nckde_apes_run (nckde_apes *a, unsigned nsteps, nckde_m2lnL_batch_fn like, void *user, double *chain);
/* The resumed run of ncm_fit_esmcmc_run (ncm_fit_esmcmc.c:2244-2286): as
 * nckde_apes_run, but the first step moves only walkers [k_first, nwalkers),
 * the rest of an ensemble whose first k_first walkers a catalog already holds
 * at that step.  chain gets nsteps x nwalkers rows as in nckde_apes_run; the
 * first step's rows below k_first are the unmoved walkers. */
int nckde_apes_run_from (nckde_apes *a, unsigned k_first, unsigned nsteps, nckde_m2lnL_batch_fn like, void *user,
                         double *chain);
int nckde_apes_get_state (nckde_apes *a, double *theta, double *m2lnL);
/* acceptance counters and seconds spent in (prepare, sample, eval, likelihood) */
int nckde_apes_get_stats (nckde_apes *a, uint64_t *accepted, uint64_t *proposed, double *seconds);
/* No reference counterpart (diagnostic hook): from the next nckde_apes_run
 * on, one row of 2 dim + 9 doubles per walker and half-step goes to buf,
 * up to max_rows rows: half (0, 1), walker, exploration active (0, 1),
 * m2lnL, m2lnL*, m2lnp*, m2lnp_cur, jump, accepted (0, 1), theta[dim],
 * theta*[dim].  buf = NULL stops the record.  The buffer stays the
 * caller's. */
int nckde_apes_set_trace (nckde_apes *a, double *buf, size_t max_rows);
int nckde_apes_get_trace_rows (nckde_apes *a, size_t *rows);

/* ---- served requests (continuous batching) -------------------------------
 *
 * No counterpart in the reference, which evaluates one object per caller
 * thread.  A server runs `slots` worker threads; each admitted request runs
 * on its session's own HIP stream, so requests on different sessions overlap
 * on the device.  Admission is in arrival order with backfill: a request is
 * admitted when a slot is free, its session has no request running or
 * waiting ahead of it (per-session order is arrival order, so results do not
 * depend on the mix), and the device bytes of the requests running plus its
 * own stay within hbm_budget (a request larger than the budget runs alone).
 * A worker that takes a mixed-precision eval also takes the waiting mixed
 * evals of other idle sessions with the same MFMA dimension and kernel type
 * (up to 16, within the budget) and runs them as one fused launch per pass
 * (nckde_eval_m2lnp_multi), with results bit-identical to separate runs.
 * A session is an nckde handle; the caller keeps every buffer of a request
 * alive until nckde_server_wait returns for it.
 */
typedef struct nckde_server nckde_server;

enum { NCKDE_REQ_EVAL_M2LNP = 0, NCKDE_REQ_SAMPLE = 1, NCKDE_REQ_PREPARE_INTERP = 2 };

typedef struct
{
  int op;                   /* NCKDE_REQ_* */
  nckde *h;                 /* session */
  size_t m;                 /* EVAL: points; SAMPLE: draws; PREPARE_INTERP: sample size */
  const double *X;          /* EVAL: m x d points; PREPARE_INTERP: m x d new sample (NULL: keep the session's) */
  const double *m2lnp;      /* PREPARE_INTERP: m values of -2 ln p at the sample */
  const double *u;          /* SAMPLE: as nckde_sample */
  const double *z;
  const double *chi2;
  double *out;              /* EVAL: m values; SAMPLE: m x d; PREPARE_INTERP (may be NULL): the
                               n_kernels fitted weights, then NaN up to m (the range cut drops points) */
  uint32_t *idx;            /* SAMPLE: chosen kernels (may be NULL) */
  int precision;            /* EVAL: NCKDE_PREC_* */
} nckde_request;

typedef struct
{
  int status;               /* NCKDE_OK or the failing call's status */
  double t_submit;          /* seconds on the server clock */
  double t_start;           /* admitted */
  double t_end;             /* results in the caller's buffers */
  size_t bytes;             /* device bytes charged at admission */
  char message[256];        /* nckde_last_error of a failing request */
  unsigned batch;           /* requests that ran in one fused launch with this one (1: alone) */
} nckde_request_info;

int nckde_server_create (nckde_server **out, unsigned slots, size_t hbm_budget);
/* waits for every submitted request, then stops the workers */
int nckde_server_destroy (nckde_server *s);
/* queue a request; returns at once with its ticket */
int nckde_server_submit (nckde_server *s, const nckde_request *req, uint64_t *ticket);
/* block until the request is done; each ticket is waited once (a second wait
 * on it, also from another thread, returns NCKDE_EINVAL) */
int nckde_server_wait (nckde_server *s, uint64_t ticket, nckde_request_info *info);
/* seconds since nckde_server_create (the clock of nckde_request_info) */
double nckde_server_clock (nckde_server *s);
/* device bytes a request would be charged at admission */
int nckde_request_bytes (const nckde_request *req, size_t *bytes);

#ifdef __cplusplus
}
#endif

#endif /* NCKDE_H */
