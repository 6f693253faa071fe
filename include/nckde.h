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
/* NcmStatsDistCV (ncm_stats_dist.h:77-85); only CV_NONE is implemented */
enum { NCKDE_CV_NONE = 0, NCKDE_CV_SPLIT = 1, NCKDE_CV_SPLIT_NOFIT = 2, NCKDE_CV_LOO = 3 };
/* NcmStatsDistKDECovType (ncm_stats_dist_kde.h:64-72); SAMPLE and FIXED */
enum { NCKDE_COV_SAMPLE = 0, NCKDE_COV_FIXED = 1, NCKDE_COV_ROBUST_DIAG = 2, NCKDE_COV_ROBUST = 3 };
/* arithmetic of the pairwise distance and kernel sums in eval */
enum { NCKDE_PREC_FP64 = 0, NCKDE_PREC_MIXED = 1 };

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
/* ncm_stats_dist_set_use_threads (ncm_stats_dist.h:115): accepted, no effect */
int nckde_set_use_threads (nckde *h, int use_threads);

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

#ifdef __cplusplus
}
#endif

#endif /* NCKDE_H */
