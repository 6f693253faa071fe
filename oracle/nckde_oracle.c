/*
 * nckde_oracle.c -- FP64 CPU restatement of NumCosmo's kernel-density engine.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP engine.  Loaded
 * only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg;
 * never by the product library.
 *
 * Built with -ffp-contract=off so that every sum and product rounds exactly
 * as written (the reference is built without FMA contraction on x86-64).
 * Reference third-party arithmetic (GSL 2.8, OpenBLAS 0.3.34, LAPACK 3.11)
 * is restated from the published algorithms: dpotrf as the textbook
 * dot-product Cholesky, dsyevr as cyclic Jacobi, dsysv as Gaussian
 * elimination with partial pivoting, dgels as Householder QR.
 *
 * Parity pinning (what fixes each function, see tests/test_oracle.py):
 *   kernels, lnnorm, sum0/sum1 ... closed forms (reference KATs
 *       tests/c/ncm/stats/test_ncm_stats_dist_kernel.c:183-420)
 *   Cholesky/lndet ............... numpy.linalg.cholesky / slogdet
 *   nearPD ....................... Higham invariants (SPD, diagonal kept)
 *   Welford covariance ........... numpy.cov (test_ncm_stats_vec.c:320-353)
 *   kNN .......................... brute-force sort on (dist2, index)
 *   NNLS ......................... KKT conditions + scipy.optimize.nnls
 *   eval_m2lnp / eval ............ brute-force logsumexp over the factors;
 *                                  Gaussian KDE with a fixed covariance
 *                                  reduces to a Gaussian mixture
 *   kernel_choose ................ serial prefix sum + bisection (exact)
 *   nearPD iterate order, NNLS active-set trajectory and BLAS summation
 *   order are tolerance-only ("parity unpinned" beyond tolerance).
 */
#include "nckde_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define OK_LN2PI 1.8378770664093454835606594728112353 /* ncm_c.h:434-437 */
#define OK_LNPI  1.1447298858494001741434273513530587 /* ncm_c.h:440-443 */
#define OK_LOG_DBL_EPSILON (-3.6043653389117154e+01)  /* GSL_LOG_DBL_EPSILON */

static __thread char ok_errbuf[512];

const char *
ok_last_error (void)
{
  return ok_errbuf;
}

static void
ok_set_error (const char *msg)
{
  snprintf (ok_errbuf, sizeof (ok_errbuf), "%s", msg);
}

static void
ok_set_error_few (int n)
{
  snprintf (ok_errbuf, sizeof (ok_errbuf),
            "ncm_stats_vec_compute_cov_robust_diag: too few points to estimate the covariance [%d].", n);
}

/* ======================================================================= */
/* Kernel family                                                           */
/* ======================================================================= */

/* ncm_stats_dist_kernel_gauss.c:193-199 and ncm_stats_dist_kernel_st.c:225-239 */
double
ok_kernel_rot_bandwidth (int kernel, double nu, int d_, double n)
{
  const unsigned d = (unsigned) d_;

  if (kernel == OK_GAUSS)
    return pow (4.0 / (n * (d + 2.0)), 1.0 / (d + 4.0));

  {
    const double nu3 = (nu >= 3.0) ? nu : 3.0;

    return pow (
      16.0 * ((nu3 - 2) * (nu3 - 2)) * (1.0 + d + nu3) * (3.0 + d + nu3) /
      ((2.0 + d) * (d + nu3) * (2.0 + d + nu3) * (d + 2.0 * nu3) * (2.0 + d + 2.0 * nu3) * n),
      1.0 / (d + 4.0));
  }
}

/* ncm_matrix.c:1157-1186 (frexp-guarded log-determinant of U^T U) */
double
ok_cholesky_lndet (int n, const double *U, int lda)
{
  const double lb = 1.0e-200;
  const double ub = 1.0e+200;
  double detL     = 1.0;
  long exponent   = 0;
  int i;

  for (i = 0; i < n; i++)
  {
    const double Lii   = fabs (U[(size_t) i * lda + i]);
    const double ndetL = detL * Lii;

    if ((ndetL < lb) || (ndetL > ub))
    {
      int exponent_i = 0;

      detL      = frexp (ndetL, &exponent_i);
      exponent += exponent_i;
    }
    else
    {
      detL = ndetL;
    }
  }

  return 2.0 * (log (detL) + exponent * M_LN2);
}

/* ncm_stats_dist_kernel_gauss.c:201-207, ncm_stats_dist_kernel_st.c:241-253 */
double
ok_kernel_lnnorm (int kernel, double nu, int d_, const double *U)
{
  const unsigned d = (unsigned) d_;

  if (kernel == OK_GAUSS)
    return 0.5 * (d * OK_LN2PI + ok_cholesky_lndet (d_, U, d_));

  {
    const double lg_lnnorm   = lgamma (nu / 2.0) - lgamma ((nu + d) / 2.0);
    const double chol_lnnorm = 0.5 * ok_cholesky_lndet (d_, U, d_);
    const double nc_lnnorm   = (d / 2.0) * (OK_LNPI + log (nu));

    return lg_lnnorm + nc_lnnorm + chol_lnnorm;
  }
}

/* ncm_stats_dist_kernel_gauss.c:209-213, ncm_stats_dist_kernel_st.c:255-263 */
double
ok_kernel_eval_unnorm (int kernel, double nu, int d_, double chi2)
{
  const unsigned d = (unsigned) d_;

  if (kernel == OK_GAUSS)
    return exp (-0.5 * chi2);

  return pow (1.0 + chi2 / nu, -0.5 * (nu + d));
}

static inline double
ok_lnt0 (int kernel, double kappa, double nu, double chi2_i, double lnu_i, double w_i)
{
  if (kernel == OK_GAUSS)
    return -0.5 * chi2_i - lnu_i + log (w_i);     /* kernel_gauss.c:269 */

  return kappa * log1p (chi2_i / nu) - lnu_i + log (w_i); /* kernel_st.c:320 */
}

/* ncm_stats_dist_kernel_gauss.c:246-290, ncm_stats_dist_kernel_st.c:295-341:
 * gamma = first maximum of lnt_i, lambda = sum over the other terms of
 * exp(lnt_i - gamma), in index order. */
void
ok_kernel_sum0 (int kernel, double nu, int d_, int n, const double *chi2,
                const double *w, const double *lnnorms, double *gamma, double *lambda)
{
  const unsigned d    = (unsigned) d_;
  const double kappa  = -0.5 * (nu + d);
  double *lnK         = malloc (sizeof (double) * (size_t) (n > 0 ? n : 1));
  double lnt_max      = -INFINITY;
  int i, i_max = 0;

  for (i = 0; i < n; i++)
  {
    const double lnt_i = ok_lnt0 (kernel, kappa, nu, chi2[i], lnnorms[i], w[i]);

    if (lnt_i > lnt_max)
    {
      i_max   = i;
      lnt_max = lnt_i;
    }

    lnK[i] = lnt_i;
  }

  lambda[0] = 0.0;

  for (i = 0; i < i_max; i++)
    lambda[0] += exp (lnK[i] - lnt_max);

  for (i = i_max + 1; i < n; i++)
    lambda[0] += exp (lnK[i] - lnt_max);

  gamma[0] = lnt_max;
  free (lnK);
}

/* ncm_stats_dist_kernel_gauss.c:292-333, ncm_stats_dist_kernel_st.c:343-386 */
void
ok_kernel_sum1 (int kernel, double nu, int d_, int n, const double *chi2,
                const double *w, double lnnorm, double *gamma, double *lambda)
{
  const unsigned d   = (unsigned) d_;
  const double kappa = -0.5 * (nu + d);
  double *lnK        = malloc (sizeof (double) * (size_t) (n > 0 ? n : 1));
  double lnt_max     = -INFINITY;
  int i, i_max = 0;

  for (i = 0; i < n; i++)
  {
    double lnt_i;

    if (kernel == OK_GAUSS)
      lnt_i = -0.5 * chi2[i] + log (w[i]);
    else
      lnt_i = kappa * log1p (chi2[i] / nu) + log (w[i]);

    if (lnt_i > lnt_max)
    {
      i_max   = i;
      lnt_max = lnt_i;
    }

    lnK[i] = lnt_i;
  }

  lambda[0] = 0.0;

  for (i = 0; i < i_max; i++)
    lambda[0] += exp (lnK[i] - lnt_max);

  for (i = i_max + 1; i < n; i++)
    lambda[0] += exp (lnK[i] - lnt_max);

  gamma[0] = lnt_max - lnnorm;
  free (lnK);
}

/* ======================================================================= */
/* Dense algebra                                                           */
/* ======================================================================= */

/*
 * ncm_matrix.c:1123-1129 -> ncm_lapack.c:215-237: dpotrf on the row-major
 * upper triangle (column-major 'L'), so A = U^T U.  Only the upper triangle
 * is read and written; the strict lower triangle is left as it was, as
 * LAPACK does.  Returns 0, or j+1 when the leading minor of order j+1 is not
 * positive definite (the LAPACK info convention).
 */
int
ok_cholesky_upper (int d, double *A, int lda)
{
  int j, i, k;

  for (j = 0; j < d; j++)
  {
    double s = A[(size_t) j * lda + j];

    for (k = 0; k < j; k++)
    {
      const double ukj = A[(size_t) k * lda + j];

      s = s - ukj * ukj;
    }

    if (!(s > 0.0))
    {
      A[(size_t) j * lda + j] = s;
      return j + 1;
    }

    s                       = sqrt (s);
    A[(size_t) j * lda + j] = s;

    for (i = j + 1; i < d; i++)
    {
      double t = A[(size_t) j * lda + i];

      for (k = 0; k < j; k++)
        t = t - A[(size_t) k * lda + j] * A[(size_t) k * lda + i];

      A[(size_t) j * lda + i] = t / s;
    }
  }

  return 0;
}

/*
 * Symmetric eigen-decomposition (stand-in for LAPACK dsyevr, range 'A').
 * Cyclic Jacobi on the symmetric matrix whose upper triangle is A.
 * Output: eigenvalues ascending in eva[d]; eigenvector i in row i of eve.
 */
int
ok_jacobi_eigen (int d, const double *A, double *eva, double *eve)
{
  double *a = malloc (sizeof (double) * d * d);
  double *v = malloc (sizeof (double) * d * d);
  int *ord  = malloc (sizeof (int) * d);
  int i, j, p, q, sweep;

  for (i = 0; i < d; i++)
    for (j = 0; j < d; j++)
    {
      a[i * d + j] = (j >= i) ? A[i * d + j] : A[j * d + i];
      v[i * d + j] = (i == j) ? 1.0 : 0.0;
    }

  for (sweep = 0; sweep < 100; sweep++)
  {
    double off = 0.0;

    for (p = 0; p < d; p++)
      for (q = p + 1; q < d; q++)
        off += a[p * d + q] * a[p * d + q];

    if (off == 0.0)
      break;

    for (p = 0; p < d; p++)
    {
      for (q = p + 1; q < d; q++)
      {
        const double apq = a[p * d + q];
        const double app = a[p * d + p];
        const double aqq = a[q * d + q];

        if (fabs (apq) <= DBL_EPSILON * 1.0e-3 * sqrt (fabs (app * aqq)) || apq == 0.0)
        {
          a[p * d + q] = 0.0;
          a[q * d + p] = 0.0;
          continue;
        }

        {
          const double theta = (aqq - app) / (2.0 * apq);
          const double t     = (theta >= 0.0 ? 1.0 : -1.0) / (fabs (theta) + sqrt (theta * theta + 1.0));
          const double c     = 1.0 / sqrt (t * t + 1.0);
          const double s     = t * c;
          int k;

          for (k = 0; k < d; k++)
          {
            const double akp = a[k * d + p];
            const double akq = a[k * d + q];

            a[k * d + p] = c * akp - s * akq;
            a[k * d + q] = s * akp + c * akq;
          }

          for (k = 0; k < d; k++)
          {
            const double apk = a[p * d + k];
            const double aqk = a[q * d + k];

            a[p * d + k] = c * apk - s * aqk;
            a[q * d + k] = s * apk + c * aqk;
          }

          for (k = 0; k < d; k++)
          {
            const double vkp = v[k * d + p];
            const double vkq = v[k * d + q];

            v[k * d + p] = c * vkp - s * vkq;
            v[k * d + q] = s * vkp + c * vkq;
          }
        }
      }
    }
  }

  for (i = 0; i < d; i++)
    ord[i] = i;

  /* insertion sort, ascending eigenvalue */
  for (i = 1; i < d; i++)
  {
    const int oi = ord[i];

    j = i - 1;

    while (j >= 0 && a[ord[j] * d + ord[j]] > a[oi * d + oi])
    {
      ord[j + 1] = ord[j];
      j--;
    }

    ord[j + 1] = oi;
  }

  for (i = 0; i < d; i++)
  {
    eva[i] = a[ord[i] * d + ord[i]];

    for (j = 0; j < d; j++)
      eve[i * d + j] = v[j * d + ord[i]];
  }

  free (a);
  free (v);
  free (ord);

  return 0;
}

/*
 * ncm_matrix.c:1247-1344: nearest positive-definite matrix (Higham 2002,
 * alternating projections with Dykstra's correction) on the upper triangle;
 * negative eigenvalues are raised to min_pos_ev * DBL_EPSILON; the diagonal
 * is restored after each projection; stops once the Cholesky succeeds or
 * after maxiter + 2 projections.  On exit A holds the Cholesky factor (upper
 * triangle) and the return value is the last Cholesky info.
 */
int
ok_nearPD (int n, double *cm, int maxiter)
{
  double *eva  = malloc (sizeof (double) * n);
  double *diag = malloc (sizeof (double) * n);
  double *eve  = malloc (sizeof (double) * n * n);
  double *D_S  = calloc ((size_t) n * n, sizeof (double));
  double *R    = malloc (sizeof (double) * n * n);
  int ret      = 0;
  int iter     = 0;
  int i, j, k;

  for (i = 0; i < n; i++)
    diag[i] = cm[i * n + i];

  while (1)
  {
    double min_pos_ev = INFINITY;

    for (i = 0; i < n * n; i++)
      cm[i] = cm[i] - D_S[i];

    memcpy (R, cm, sizeof (double) * n * n);

    ok_jacobi_eigen (n, cm, eva, eve);

    for (i = 0; i < n; i++)
      if (eva[i] > 0.0)
        min_pos_ev = fmin (eva[i], min_pos_ev);

    for (i = 0; i < n; i++)
    {
      const double s = (eva[i] < 0.0) ? sqrt (min_pos_ev * DBL_EPSILON) : sqrt (eva[i]);

      if (eva[i] < 0.0)
        eva[i] = min_pos_ev * DBL_EPSILON;

      for (j = 0; j < n; j++)
        eve[i * n + j] = s * eve[i * n + j];
    }

    /* dsyrk, row-major, upper, trans: C = eve^T eve on the upper triangle */
    for (i = 0; i < n; i++)
      for (j = i; j < n; j++)
      {
        double s = 0.0;

        for (k = 0; k < n; k++)
          s = s + eve[k * n + i] * eve[k * n + j];

        cm[i * n + j] = s;
      }

    for (i = 0; i < n * n; i++)
      D_S[i] = cm[i] - R[i];

    for (i = 0; i < n; i++)
      cm[i * n + i] = diag[i];

    memcpy (R, cm, sizeof (double) * n * n);

    if ((ret = ok_cholesky_upper (n, R, n)) == 0)
      break;

    if (iter > maxiter)
      break;

    iter++;
  }

  memcpy (cm, R, sizeof (double) * n * n);

  free (eva);
  free (diag);
  free (eve);
  free (D_S);
  free (R);

  return ret;
}

/* ncm_stats_dist_vkde.c:337-360 (identical to ncm_stats_dist_kde.c:344-367):
 * Cholesky, else nearPD, else the Cholesky of the diagonal. */
void
ok_cholesky_decomp_fallback (int d, const double *cov, double *U, int maxiter)
{
  memcpy (U, cov, sizeof (double) * d * d);

  if (ok_cholesky_upper (d, U, d) != 0)
  {
    memcpy (U, cov, sizeof (double) * d * d);

    if (ok_nearPD (d, U, maxiter) != 0)
    {
      int i;

      memset (U, 0, sizeof (double) * d * d);

      for (i = 0; i < d; i++)
        U[i * d + i] = cov[i * d + i];

      if (ok_cholesky_upper (d, U, d) != 0)
        ok_set_error ("_cholesky_decomp: diagonal covariance is not positive.");
    }
  }
}

/* ======================================================================= */
/* NcmStatsVec: weighted one-pass covariance                               */
/* ======================================================================= */

/*
 * ncm_stats_vec.c:510-565 (update, weight 1) and :2376-2399 (get_cov_matrix):
 * West/Welford update in insertion order; off-diagonal from the co-moment,
 * diagonal from the variance accumulator, scaled by
 * bias_wt = 1 / (W - W2 / W).
 */
void
ok_cov_welford (int d, int nrows, const double *X, const int *rows, double *cov)
{
  double *mean  = calloc (d, sizeof (double));
  double *var   = calloc (d, sizeof (double));
  double *C     = calloc ((size_t) d * d, sizeof (double));
  double weight = 0.0, weight2 = 0.0, bias_wt = 0.0;
  const double w = 1.0;
  int t, i, j;

  for (t = 0; t < nrows; t++)
  {
    const double *x        = X + (size_t) (rows ? rows[t] : t) * d;
    const double curweight = weight + w;

    for (i = 0; i < d; i++)
    {
      double mean_i        = mean[i];
      const double x_i     = x[i];
      const double delta_i = x_i - mean_i;
      const double R_i     = delta_i * w / curweight;
      const double vv      = var[i];
      const double dvar    = weight * delta_i * R_i;

      mean_i += R_i;
      mean[i] = mean_i;
      var[i]  = vv + dvar;

      for (j = i + 1; j < d; j++)
      {
        const double x_j    = x[j];
        const double mean_j = mean[j];
        const double dC_ij  = w * (x_i - mean_i) * (x_j - mean_j);
        const double C_ij   = C[i * d + j] + dC_ij;

        C[i * d + j] = C_ij;
        C[j * d + i] = C_ij;
      }
    }

    weight   = curweight;
    weight2 += w * w;
    bias_wt  = 1.0 / (weight - weight2 / weight);
  }

  for (i = 0; i < d; i++)
    for (j = 0; j < d; j++)
      cov[i * d + j] = ((i == j) ? var[i] : C[i * d + j]) * bias_wt;

  free (mean);
  free (var);
  free (C);
}

/* ======================================================================= */
/* Exact kNN                                                               */
/* ======================================================================= */

typedef struct
{
  double d2;
  int j;
} ok_pair;

static inline int
ok_pair_lt (const ok_pair *a, const ok_pair *b)
{
  /* rb_knn_list.c:31-40: order by distance, then by index */
  if (a->d2 < b->d2)
    return 1;

  if (a->d2 > b->d2)
    return 0;

  return a->j < b->j;
}

static int
ok_pair_cmp (const void *pa, const void *pb)
{
  const ok_pair *a = pa, *b = pb;

  if (ok_pair_lt (a, b))
    return -1;

  if (ok_pair_lt (b, a))
    return 1;

  return 0;
}

static void
ok_select (ok_pair *p, int n, int k)
{
  /* quickselect: afterwards p[0..k-1] hold the k smallest keys (any order);
   * keys are distinct, so the selected set does not depend on pivots */
  int lo = 0, hi = n - 1;

  while (hi > lo)
  {
    const ok_pair piv = p[lo + (hi - lo) / 2];
    int i = lo, j = hi;

    while (i <= j)
    {
      while (ok_pair_lt (&p[i], &piv))
        i++;

      while (ok_pair_lt (&piv, &p[j]))
        j--;

      if (i <= j)
      {
        const ok_pair t = p[i];

        p[i] = p[j];
        p[j] = t;
        i++;
        j--;
      }
    }

    if (k - 1 <= j)
      hi = j;
    else if (k - 1 >= i)
      lo = i;
    else
      break;
  }
}

/*
 * kdtree.c:27-43 (squared distance summed over coordinates in order),
 * kdtree.c:192-320 (exact k-NN search) and rb_knn_list.c:31-40 ((dist2,
 * index) ordering).  The tree search returns exactly the k smallest
 * (dist2, index) keys, in ascending order; this is a brute-force restatement.
 */
void
ok_knn (int n, int d, const double *Y, int q, int k, int *out_idx, double *out_d2)
{
  ok_pair *p       = malloc (sizeof (ok_pair) * n);
  const double *yq = Y + (size_t) q * d;
  int j, l;

  for (j = 0; j < n; j++)
  {
    const double *yj = Y + (size_t) j * d;
    double dist      = 0.0;

    for (l = 0; l < d; l++)
    {
      const double t = yj[l] - yq[l];

      dist = dist + t * t;
    }

    p[j].d2 = dist;
    p[j].j  = j;
  }

  if (k < n)
    ok_select (p, n, k);

  qsort (p, (k < n) ? k : n, sizeof (ok_pair), ok_pair_cmp);

  for (l = 0; l < k && l < n; l++)
  {
    out_idx[l] = p[l].j;

    if (out_d2)
      out_d2[l] = p[l].d2;
  }

  free (p);
}

/* ======================================================================= */
/* Active-set NNLS                                                         */
/* ======================================================================= */

/* Cholesky solve of an SPD system (dposv, ncm_matrix.c:1198-1208).
 * S is |P|x|P| row-major (upper triangle used), b overwritten by x. */
static int
ok_chol_solve (int n, double *S, double *b)
{
  int i, k;
  int ret = ok_cholesky_upper (n, S, n);

  if (ret != 0)
    return ret;

  /* U^T y = b */
  for (i = 0; i < n; i++)
  {
    double t = b[i];

    for (k = 0; k < i; k++)
      t = t - S[(size_t) k * n + i] * b[k];

    b[i] = t / S[(size_t) i * n + i];
  }

  /* U x = y */
  for (i = n - 1; i >= 0; i--)
  {
    double t = b[i];

    for (k = i + 1; k < n; k++)
      t = t - S[(size_t) i * n + k] * b[k];

    b[i] = t / S[(size_t) i * n + i];
  }

  return 0;
}

/* Symmetric solve fallback (stands in for dsysv, ncm_nnls.c:574-605):
 * Gaussian elimination with partial pivoting on the full symmetric matrix.
 * Returns >0 when a pivot is exactly zero. */
static int
ok_lu_solve (int n, double *S, double *b)
{
  int i, j, k;

  for (i = 0; i < n; i++)
    for (j = 0; j < i; j++)
      S[(size_t) i * n + j] = S[(size_t) j * n + i];

  for (k = 0; k < n; k++)
  {
    int piv      = k;
    double amax  = fabs (S[(size_t) k * n + k]);

    for (i = k + 1; i < n; i++)
      if (fabs (S[(size_t) i * n + k]) > amax)
      {
        amax = fabs (S[(size_t) i * n + k]);
        piv  = i;
      }

    if (amax == 0.0)
      return k + 1;

    if (piv != k)
    {
      for (j = 0; j < n; j++)
      {
        const double t = S[(size_t) k * n + j];

        S[(size_t) k * n + j]   = S[(size_t) piv * n + j];
        S[(size_t) piv * n + j] = t;
      }

      {
        const double t = b[k];

        b[k]   = b[piv];
        b[piv] = t;
      }
    }

    for (i = k + 1; i < n; i++)
    {
      const double l = S[(size_t) i * n + k] / S[(size_t) k * n + k];

      for (j = k + 1; j < n; j++)
        S[(size_t) i * n + j] = S[(size_t) i * n + j] - l * S[(size_t) k * n + j];

      b[i] = b[i] - l * b[k];
    }
  }

  for (i = n - 1; i >= 0; i--)
  {
    double t = b[i];

    for (j = i + 1; j < n; j++)
      t = t - S[(size_t) i * n + j] * b[j];

    b[i] = t / S[(size_t) i * n + i];
  }

  return 0;
}

/* Minimum-norm solution of the underdetermined A_P x = f (nrows < np), the
 * other half of dgels 'N' (LAPACK dgels: LQ factorisation of A_P).  Done as
 * the Householder QR of B = A_P^T (np x nrows, B = Q R): A_P = R^T Q^T, so
 * x = Q [R^{-T} f; 0].  Reflector signs as dlarfg (beta = -sign(a) |.|). */
static void
ok_lq_minnorm (int nrows, int ncols, const double *A, const int *P, int np, const double *f, double *x)
{
  double *B   = malloc (sizeof (double) * (size_t) np * nrows);
  double *v0  = malloc (sizeof (double) * nrows);
  double *vtv = malloc (sizeof (double) * nrows);
  int i, j, k;

  for (j = 0; j < np; j++)
    for (i = 0; i < nrows; i++)
      B[(size_t) j * nrows + i] = A[(size_t) i * ncols + P[j]];

  for (k = 0; k < nrows; k++)
  {
    double nrm = 0.0, alpha, t = 0.0;

    for (j = k; j < np; j++)
      nrm = nrm + B[(size_t) j * nrows + k] * B[(size_t) j * nrows + k];

    nrm    = sqrt (nrm);
    vtv[k] = 0.0;
    v0[k]  = 0.0;

    if (nrm == 0.0)
      continue;

    alpha = (B[(size_t) k * nrows + k] >= 0.0) ? -nrm : nrm;
    B[(size_t) k * nrows + k] -= alpha;

    for (j = k; j < np; j++)
      t = t + B[(size_t) j * nrows + k] * B[(size_t) j * nrows + k];

    vtv[k] = t;

    for (i = k + 1; i < nrows; i++)
    {
      double sd = 0.0;

      for (j = k; j < np; j++)
        sd = sd + B[(size_t) j * nrows + k] * B[(size_t) j * nrows + i];

      sd = 2.0 * sd / t;

      for (j = k; j < np; j++)
        B[(size_t) j * nrows + i] -= sd * B[(size_t) j * nrows + k];
    }

    v0[k]                     = B[(size_t) k * nrows + k];
    B[(size_t) k * nrows + k] = alpha;
  }

  /* R^T y = f (forward), y into x[0..nrows-1], zeros below */
  for (k = 0; k < nrows; k++)
  {
    double t = f[k];

    for (i = 0; i < k; i++)
      t = t - B[(size_t) i * nrows + k] * x[i];

    x[k] = (B[(size_t) k * nrows + k] != 0.0) ? t / B[(size_t) k * nrows + k] : 0.0;
  }

  for (j = nrows; j < np; j++)
    x[j] = 0.0;

  /* x = H_0 H_1 ... H_{nrows-1} [y; 0] */
  for (k = nrows - 1; k >= 0; k--)
  {
    double sd;

    if (vtv[k] == 0.0)
      continue;

    sd = v0[k] * x[k];

    for (j = k + 1; j < np; j++)
      sd = sd + B[(size_t) j * nrows + k] * x[j];

    sd = 2.0 * sd / vtv[k];

    x[k] -= sd * v0[k];

    for (j = k + 1; j < np; j++)
      x[j] -= sd * B[(size_t) j * nrows + k];
  }

  free (B);
  free (v0);
  free (vtv);
}

/* Least squares min |A_P x - f| by Householder QR (stands in for dgels,
 * ncm_nnls.c:608-640).  A is nrows x ncols row-major, P the column list;
 * nrows < np goes to the minimum-norm solution as dgels does.
 * A column that is exactly zero below the current row (an exactly
 * dependent column, e.g. a duplicate) is skipped with zero weight and the
 * next reflector starts on the same row, so the others solve the reduced
 * problem; dgels would stop there with info > 0 (a zero diagonal of R). */
static void
ok_qr_lstsq (int nrows, int ncols, const double *A, const int *P, int np, const double *f, double *x)
{
  double *Q, *rh;
  int *prow;
  int i, j, k, r = 0;

  if (nrows < np)
  {
    ok_lq_minnorm (nrows, ncols, A, P, np, f, x);
    return;
  }

  Q    = malloc (sizeof (double) * (size_t) nrows * np);
  rh   = malloc (sizeof (double) * nrows);
  prow = malloc (sizeof (int) * np);

  for (i = 0; i < nrows; i++)
    for (j = 0; j < np; j++)
      Q[(size_t) i * np + j] = A[(size_t) i * ncols + P[j]];

  memcpy (rh, f, sizeof (double) * nrows);

  for (k = 0; k < np; k++)
  {
    double nrm = 0.0, alpha, vtv = 0.0;

    prow[k] = -1;

    if (r == nrows)
      continue;

    for (i = r; i < nrows; i++)
      nrm = nrm + Q[(size_t) i * np + k] * Q[(size_t) i * np + k];

    if (nrm == 0.0)
      continue;

    nrm   = sqrt (nrm);
    alpha = (Q[(size_t) r * np + k] >= 0.0) ? -nrm : nrm;

    Q[(size_t) r * np + k] -= alpha;

    for (i = r; i < nrows; i++)
      vtv = vtv + Q[(size_t) i * np + k] * Q[(size_t) i * np + k];

    for (j = k + 1; j < np; j++)
    {
      double sd = 0.0;

      for (i = r; i < nrows; i++)
        sd = sd + Q[(size_t) i * np + k] * Q[(size_t) i * np + j];

      sd = 2.0 * sd / vtv;

      for (i = r; i < nrows; i++)
        Q[(size_t) i * np + j] -= sd * Q[(size_t) i * np + k];
    }

    {
      double sd = 0.0;

      for (i = r; i < nrows; i++)
        sd = sd + Q[(size_t) i * np + k] * rh[i];

      sd = 2.0 * sd / vtv;

      for (i = r; i < nrows; i++)
        rh[i] -= sd * Q[(size_t) i * np + k];
    }

    Q[(size_t) r * np + k] = alpha;
    prow[k]                = r++;
  }

  for (k = np - 1; k >= 0; k--)
  {
    const int row = prow[k];
    double t;

    if (row < 0)
    {
      x[k] = 0.0;
      continue;
    }

    t = rh[row];

    for (j = k + 1; j < np; j++)
      if (prow[j] >= 0)
        t = t - Q[(size_t) row * np + j] * x[j];

    x[k] = t / Q[(size_t) row * np + k];
  }

  free (Q);
  free (rh);
  free (prow);
}

typedef struct
{
  int nrows, ncols;
  int umethod; /* OK_NNLS_NORMAL, OK_NNLS_NORMAL_LU or OK_NNLS_QR */
  const double *A;
  const double *f;
  double *M;   /* A^T A, full */
  double *b;   /* A^T f */
  double *S;   /* scratch |P|^2 */
  double *bs;  /* scratch |P| */
  int *P;      /* sorted index list */
} ok_nnls_ws;

/* Solves that ended on each rung (Cholesky, symmetric solve, QR) since the
 * last ok_nnls_rung_counts (..., 1): a test can show that a fit took a
 * fallback rung. */
static __thread unsigned ok_rungs[3];

void
ok_nnls_rung_counts (unsigned *out, int reset)
{
  memcpy (out, ok_rungs, sizeof (ok_rungs));

  if (reset)
    memset (ok_rungs, 0, sizeof (ok_rungs));
}

static void
ok_gather_normal (ok_nnls_ws *ws, int np)
{
  int a, c;

  for (a = 0; a < np; a++)
  {
    for (c = 0; c < np; c++)
      ws->S[(size_t) a * np + c] = ws->M[(size_t) ws->P[a] * ws->ncols + ws->P[c]];

    ws->bs[a] = ws->b[ws->P[a]];
  }
}

/* ncm_nnls.c:671-706 with :655-666 -> :574-605 -> :608-640.
 * NORMAL: normal equations by Cholesky, else the symmetric-indefinite
 * solve, else least squares by QR; NORMAL_LU starts at the symmetric solve;
 * QR solves min |A_P x - f| directly.
 * Writes the solution on the sorted index list into sol[0..np-1]. */
static void
ok_nnls_unconstrained (ok_nnls_ws *ws, const char *inP, double *sol)
{
  int np = 0, a;

  for (a = 0; a < ws->ncols; a++)
    if (inP[a])
      ws->P[np++] = a;

  if (ws->umethod == OK_NNLS_QR)
  {
    ok_qr_lstsq (ws->nrows, ws->ncols, ws->A, ws->P, np, ws->f, ws->bs);
    ok_rungs[2]++;
  }
  else
  {
    int chol_fail = 1;

    if (ws->umethod == OK_NNLS_NORMAL)
    {
      ok_gather_normal (ws, np);
      chol_fail = (ok_chol_solve (np, ws->S, ws->bs) > 0);

      if (!chol_fail)
        ok_rungs[0]++;
    }

    if (chol_fail)
    {
      ok_gather_normal (ws, np);

      if (ok_lu_solve (np, ws->S, ws->bs) > 0)
      {
        ok_qr_lstsq (ws->nrows, ws->ncols, ws->A, ws->P, np, ws->f, ws->bs);
        ok_rungs[2]++;
      }
      else
      {
        ok_rungs[1]++;
      }
    }
  }

  memcpy (sol, ws->bs, sizeof (double) * np);
}

/* min |A x - f| over all columns by the Householder QR rung (dgels). */
void
ok_lstsq_qr (int nrows, int ncols, const double *A, const double *f, double *x)
{
  int *P = malloc (sizeof (int) * ncols);
  int j;

  for (j = 0; j < ncols; j++)
    P[j] = j;

  ok_qr_lstsq (nrows, ncols, A, P, ncols, f, x);
  free (P);
}

/* ncm_nnls.c:728-751: solve on P, drop indices whose solution is < 1e-300
 * (at most max_remove per pass, smallest first, ncm_iset.c:919-979) and
 * repeat until the solution is feasible. */
static void
ok_nnls_feasible (ok_nnls_ws *ws, char *inP, double *x, int max_remove)
{
  const int n  = ws->ncols;
  double *sol  = malloc (sizeof (double) * n);
  int *inv     = malloc (sizeof (int) * n);
  int ninv, a, np;

  while (1)
  {
    ok_nnls_unconstrained (ws, inP, sol);

    memset (x, 0, sizeof (double) * n);
    np = 0;

    for (a = 0; a < n; a++)
      if (inP[a])
        x[a] = sol[np++];

    ninv = 0;

    for (a = 0; a < n; a++)
      if (inP[a] && (x[a] < 1.0e-300))
        inv[ninv++] = a;

    if (ninv == 0)
      break;

    if (max_remove >= ninv)
    {
      for (a = 0; a < ninv; a++)
        inP[inv[a]] = 0;
    }
    else
    {
      /* remove the max_remove smallest x (ties by index) */
      ok_pair *pp = malloc (sizeof (ok_pair) * ninv);

      for (a = 0; a < ninv; a++)
      {
        pp[a].d2 = x[inv[a]];
        pp[a].j  = inv[a];
      }

      qsort (pp, ninv, sizeof (ok_pair), ok_pair_cmp);

      for (a = 0; a < max_remove; a++)
        inP[pp[a].j] = 0;

      free (pp);
    }
  }

  free (sol);
  free (inv);
}

static double
ok_nnls_residuals (ok_nnls_ws *ws, const double *x, double *r)
{
  double s = 0.0;
  int i, j;

  for (i = 0; i < ws->nrows; i++)
  {
    double t = 0.0;

    for (j = 0; j < ws->ncols; j++)
      t = t + ws->A[(size_t) i * ws->ncols + j] * x[j];

    r[i] = ws->f[i] - t;
    s    = s + r[i] * r[i];
  }

  return sqrt (s);
}

static void
ok_nnls_mgrad (ok_nnls_ws *ws, const double *r, double *g)
{
  int i, j;

  for (j = 0; j < ws->ncols; j++)
    g[j] = 0.0;

  for (i = 0; i < ws->nrows; i++)
    for (j = 0; j < ws->ncols; j++)
      g[j] = g[j] + ws->A[(size_t) i * ws->ncols + j] * r[i];
}

/*
 * ncm_nnls.c:767-871 (umethod NORMAL, reltol DBL_EPSILON): active-set NNLS
 * on the normal equations.  Returns the residual norm |f - A x|.
 */
double
ok_nnls (int nrows, int ncols, const double *A, const double *f, double *x)
{
  return ok_nnls_m (nrows, ncols, A, f, x, OK_NNLS_NORMAL);
}

double
ok_nnls_m (int nrows, int ncols, const double *A, const double *f, double *x, int umethod)
{
  const double reltol = DBL_EPSILON;
  ok_nnls_ws ws;
  char *P     = malloc (ncols);
  char *Ptry  = malloc (ncols);
  double *xt  = malloc (sizeof (double) * ncols);
  double *r   = malloc (sizeof (double) * nrows);
  double *rt  = malloc (sizeof (double) * nrows);
  double *g   = malloc (sizeof (double) * ncols);
  ok_pair *cand = malloc (sizeof (ok_pair) * ncols);
  double rnorm;
  int i, j, k;

  ws.nrows   = nrows;
  ws.ncols   = ncols;
  ws.umethod = umethod;
  ws.A       = A;
  ws.f     = f;
  ws.M     = malloc (sizeof (double) * (size_t) ncols * ncols);
  ws.b     = malloc (sizeof (double) * ncols);
  ws.S     = malloc (sizeof (double) * (size_t) ncols * ncols);
  ws.bs    = malloc (sizeof (double) * ncols);
  ws.P     = malloc (sizeof (int) * ncols);

  /* M = A^T A (dsyrk, ncm_matrix.c:1547-1575), b = A^T f (dgemv); the QR
   * method works on A alone (ncm_nnls.c:795-797) */
  if (umethod != OK_NNLS_QR)
  {
    #pragma omp parallel for schedule(dynamic, 8) private(j, k)
    for (i = 0; i < ncols; i++)
    {
      for (j = i; j < ncols; j++)
      {
        double s = 0.0;

        for (k = 0; k < nrows; k++)
          s = s + A[(size_t) k * ncols + i] * A[(size_t) k * ncols + j];

        ws.M[(size_t) i * ncols + j] = s;
        ws.M[(size_t) j * ncols + i] = s;
      }
    }

    for (j = 0; j < ncols; j++)
    {
      double s = 0.0;

      for (k = 0; k < nrows; k++)
        s = s + A[(size_t) k * ncols + j] * f[k];

      ws.b[j] = s;
    }
  }

  memset (P, 1, ncols);

  ok_nnls_feasible (&ws, P, x, ncols);
  rnorm = ok_nnls_residuals (&ws, x, r);
  ok_nnls_mgrad (&ws, r, g);

  while (1)
  {
    double add_frac = 1.0;
    int finish      = 0;
    double lrnorm;
    unsigned added;

    while (1)
    {
      /* ncm_iset.c:993-1078: candidates j not in P with g_j > rnorm*reltol,
       * add the largest k*add_frac of them (at least one) */
      const double gmin = rnorm * reltol;
      unsigned kc       = 0;

      memcpy (Ptry, P, ncols);

      for (j = 0; j < ncols; j++)
        if (!Ptry[j] && (g[j] > gmin))
        {
          cand[kc].d2 = -g[j]; /* ascending on -g == descending on g */
          cand[kc].j  = j;
          kc++;
        }

      {
        const double kf  = kc * add_frac;
        const double mx  = (kf > 1) ? kf : 1;
        const double adf = (kc < mx) ? kc : mx;

        added = (unsigned) adf;
      }

      if (kc == 0)
        added = 0;

      add_frac *= 0.5;

      if (added == 0)
      {
        finish = 1;
        break;
      }

      qsort (cand, kc, sizeof (ok_pair), ok_pair_cmp);

      for (j = 0; j < (int) added; j++)
        Ptry[cand[j].j] = 1;

      ok_nnls_feasible (&ws, Ptry, xt, (int) added);
      lrnorm = ok_nnls_residuals (&ws, xt, rt);

      if (rnorm - lrnorm > rnorm * reltol)
      {
        memcpy (P, Ptry, ncols);
        memcpy (x, xt, sizeof (double) * ncols);
        memcpy (r, rt, sizeof (double) * nrows);
        rnorm = lrnorm;
        break;
      }

      if (added == 1)
      {
        finish = 1;
        break;
      }
    }

    if (finish)
      break;

    ok_nnls_mgrad (&ws, r, g);
  }

  free (ws.M);
  free (ws.b);
  free (ws.S);
  free (ws.bs);
  free (ws.P);
  free (P);
  free (Ptry);
  free (xt);
  free (r);
  free (rt);
  free (g);
  free (cand);

  return rnorm;
}

/* ======================================================================= */
/* NcmStatsDist / KDE / VKDE                                               */
/* ======================================================================= */

struct ok_sd
{
  int kind, kernel, d;
  double nu;
  double over_smooth, shrink, local_frac, split_frac;
  int use_rot_href, cov_type, nearPD_maxiter, use_threads, cv_type, nnls_umethod;
  uint64_t cv_draws;
  double *cov_fixed, *cov_fixed_decomp;

  double *X;            /* sample array (add_obs copies), row-major */
  int n_samples, n_alloc;

  int prepared, n_obs, n_kernels, k;
  double href, kernel_lnnorm, rnorm, min_m2lnp, max_m2lnp;
  double *weights, *wcum;
  int wcum_ready;
  double *cov, *cov_decomp;  /* global d x d */
  double *invU;              /* n_obs x d whitened sample */
  double *U_arr;             /* VKDE: n_kernels x d x d */
  double *lnnorms;           /* VKDE: n_kernels */
  int *knn;                  /* VKDE: n_kernels x k */
};

ok_sd *
ok_new (int kind, int kernel, double nu, int d)
{
  ok_sd *sd = calloc (1, sizeof (ok_sd));

  sd->kind           = kind;
  sd->kernel         = kernel;
  sd->nu             = nu;
  sd->d              = d;
  sd->over_smooth    = 1.0;   /* ncm_stats_dist.c:389-396 */
  sd->shrink         = 0.01;  /* ncm_stats_dist.c:421-427 */
  sd->local_frac     = 0.05;  /* ncm_stats_dist_vkde.c:276-282 */
  sd->split_frac     = 0.5;   /* ncm_stats_dist.c:430-434 */
  sd->cv_type        = OK_CV_NONE;
  sd->use_rot_href   = 0;
  sd->cov_type       = OK_COV_SAMPLE;
  sd->nearPD_maxiter = 200;   /* ncm_stats_dist_kde.c:290-296 */
  sd->use_threads    = 0;
  sd->nnls_umethod   = OK_NNLS_NORMAL; /* ncm_stats_dist.c:861 */

  return sd;
}

static void
ok_clear_prepared (ok_sd *sd)
{
  free (sd->weights);
  free (sd->wcum);
  free (sd->cov);
  free (sd->cov_decomp);
  free (sd->invU);
  free (sd->U_arr);
  free (sd->lnnorms);
  free (sd->knn);
  sd->weights = sd->wcum = sd->cov = sd->cov_decomp = sd->invU = sd->U_arr = sd->lnnorms = NULL;
  sd->knn      = NULL;
  sd->prepared = 0;
}

void
ok_free (ok_sd *sd)
{
  if (sd == NULL)
    return;

  ok_clear_prepared (sd);
  free (sd->X);
  free (sd->cov_fixed);
  free (sd->cov_fixed_decomp);
  free (sd);
}

void ok_set_over_smooth (ok_sd *sd, double v) { sd->over_smooth = v; }
void ok_set_shrink (ok_sd *sd, double v) { sd->shrink = v; }
void ok_set_local_frac (ok_sd *sd, double v) { sd->local_frac = v; }
void ok_set_use_rot_href (ok_sd *sd, int v) { sd->use_rot_href = v; }
void ok_set_nearPD_maxiter (ok_sd *sd, int v) { sd->nearPD_maxiter = v; }
void ok_set_use_threads (ok_sd *sd, int v) { sd->use_threads = v; }
void ok_set_nnls_umethod (ok_sd *sd, int v) { sd->nnls_umethod = v; }

/* OpenMP team size of the batched calls (the CPU-baseline timings run at one
 * thread and at all cores) */
void
ok_set_num_threads (int n)
{
#ifdef _OPENMP
  omp_set_num_threads (n);
#else
  (void) n;
#endif
}

int
ok_get_max_threads (void)
{
#ifdef _OPENMP
  return omp_get_max_threads ();
#else
  return 1;
#endif
}

/* ncm_stats_dist_kde.c:731-745 and :770-791 */
static int
ok_decompose_fixed (ok_sd *sd)
{
  const int d = sd->d;

  if (sd->cov_fixed == NULL)
    return 0;

  if (sd->cov_fixed_decomp == NULL)
    sd->cov_fixed_decomp = malloc (sizeof (double) * d * d);

  memcpy (sd->cov_fixed_decomp, sd->cov_fixed, sizeof (double) * d * d);

  if (ok_cholesky_upper (d, sd->cov_fixed_decomp, d) != 0)
  {
    ok_set_error ("ncm_stats_dist_kde_set_cov_fixed: matrix cov_fixed is not positive definite.");
    return -1;
  }

  return 0;
}

int
ok_set_cov_type (ok_sd *sd, int cov_type)
{
  sd->cov_type = cov_type;

  if (cov_type == OK_COV_FIXED)
    return ok_decompose_fixed (sd);

  return 0;
}

int
ok_set_cov_fixed (ok_sd *sd, const double *cov)
{
  const int d = sd->d;

  if (sd->cov_fixed == NULL)
    sd->cov_fixed = malloc (sizeof (double) * d * d);

  memcpy (sd->cov_fixed, cov, sizeof (double) * d * d);

  if (sd->cov_type == OK_COV_FIXED)
    return ok_decompose_fixed (sd);

  return 0;
}

/*
 * Qn scale estimator of Rousseeuw & Croux (J. Am. Stat. Assoc. 88 (1993)
 * 1273-1283; Croux & Rousseeuw, Comput. Stat. 1 (1992) 411-428), the
 * estimator behind gsl_stats_Qn_from_sorted_data (GSL 2.8) as called in
 * ncm_stats_vec.c:1849-1855:
 *   Qn = c_n * (2.21914 * d_(k)),  d_(k) the k-th smallest of x_j - x_i
 *   (i < j) over the sorted data, k = h (h - 1) / 2, h = n / 2 + 1,
 *   c_n = .399 .994 .512 .844 .611 .857 .669 .872 for n = 2..9, else
 *   n / (n + 1.4) (n odd) or n / (n + 3.8) (n even).
 * Brute force: every pairwise difference, then the k-th order statistic.
 */
static int
ok_dcmp (const void *a, const void *b)
{
  const double x = *(const double *) a, y = *(const double *) b;

  return (x > y) - (x < y);
}

double
ok_qn (int n, const double *x)
{
  static const double cn_small[10] = {0.0, 0.0, 0.399, 0.994, 0.512, 0.844, 0.611, 0.857, 0.669, 0.872};
  double *v  = malloc (sizeof (double) * n);
  const long np = (long) n * (n - 1) / 2;
  double *dd = malloc (sizeof (double) * (np > 0 ? np : 1));
  const long h = n / 2 + 1;
  const long k = h * (h - 1) / 2;
  double cn, kth;
  long i, j, q = 0;

  memcpy (v, x, sizeof (double) * n);
  qsort (v, n, sizeof (double), ok_dcmp);

  for (i = 0; i < n; i++)
    for (j = i + 1; j < n; j++)
      dd[q++] = v[j] - v[i];

  qsort (dd, np, sizeof (double), ok_dcmp);
  kth = dd[k - 1];

  if (n <= 9)
    cn = cn_small[n];
  else
    cn = (n % 2) ? (double) n / (n + 1.4) : (double) n / (n + 3.8);

  free (v);
  free (dd);

  return cn * (2.21914 * kth);
}

/* ncm_stats_vec.c:1834-1875: diagonal covariance of squared Qn scales */
int
ok_cov_robust_diag (int d, int nrows, const double *X, const int *rows, double *cov)
{
  double *col = malloc (sizeof (double) * nrows);
  int i, a;

  if (nrows < 4)
  {
    ok_set_error_few (nrows);
    free (col);
    return -1;
  }

  memset (cov, 0, sizeof (double) * d * d);

  for (i = 0; i < d; i++)
  {
    double s;

    for (a = 0; a < nrows; a++)
      col[a] = X[(size_t) (rows ? rows[a] : a) * d + i];

    s              = ok_qn (nrows, col);
    cov[i * d + i] = s * s;
  }

  free (col);

  return 0;
}

/*
 * ncm_stats_vec.c:1889-2064: Orthogonalized Gnanadesikan-Kettenring.
 *   y = x / sigma_x (times the reciprocal, ncm_matrix_mul_col);
 *   U_ij = (Qn(y_i + y_j)^2 - Qn(y_i - y_j)^2) / 4, U_ii = 1;
 *   U = E^T diag E (rows of E the eigenvectors; ok_jacobi_eigen for dsyevr);
 *   z = E y^T, sigma_z = Qn of the rows of z;
 *   V_ij = sigma_z_i E_ij sigma_x_j;  cov = V^T V (both triangles written).
 * cov is invariant to the sign and order of the eigenvectors.
 */
int
ok_cov_robust_ogk (int d, int nrows, const double *X, const int *rows, double *cov)
{
  double *y    = malloc (sizeof (double) * (size_t) nrows * d);
  double *col  = malloc (sizeof (double) * nrows);
  double *sx   = malloc (sizeof (double) * d);
  double *sz   = malloc (sizeof (double) * d);
  double *U    = calloc ((size_t) d * d, sizeof (double));
  double *E    = malloc (sizeof (double) * d * d);
  double *eva  = malloc (sizeof (double) * d);
  double *V    = malloc (sizeof (double) * d * d);
  int i, j, a, l;

  if (nrows < 4)
  {
    ok_set_error_few (nrows);
    free (y); free (col); free (sx); free (sz); free (U); free (E); free (eva); free (V);
    return -1;
  }

  for (i = 0; i < d; i++)
  {
    double inv;

    for (a = 0; a < nrows; a++)
      col[a] = X[(size_t) (rows ? rows[a] : a) * d + i];

    sx[i] = ok_qn (nrows, col);
    inv   = 1.0 / sx[i];

    for (a = 0; a < nrows; a++)
      y[(size_t) a * d + i] = col[a] * inv;
  }

  for (i = 0; i < d; i++)
    for (j = 0; j < d; j++)
      U[i * d + j] = (i == j) ? 1.0 : 0.0;

  for (i = 0; i < d; i++)
    for (j = i + 1; j < d; j++)
    {
      double sp, sm;

      for (a = 0; a < nrows; a++)
        col[a] = y[(size_t) a * d + i] + y[(size_t) a * d + j];

      sp = ok_qn (nrows, col);

      for (a = 0; a < nrows; a++)
        col[a] = y[(size_t) a * d + i] - y[(size_t) a * d + j];

      sm           = ok_qn (nrows, col);
      U[i * d + j] = 0.25 * (sp * sp - sm * sm);
    }

  ok_jacobi_eigen (d, U, eva, E);

  for (i = 0; i < d; i++)
  {
    for (a = 0; a < nrows; a++)
    {
      double t = 0.0;

      for (l = 0; l < d; l++)
        t = t + E[i * d + l] * y[(size_t) a * d + l];

      col[a] = t;
    }

    sz[i] = ok_qn (nrows, col);
  }

  for (i = 0; i < d; i++)
    for (j = 0; j < d; j++)
      V[i * d + j] = sz[i] * E[i * d + j] * sx[j];

  for (j = 0; j < d; j++)
    for (l = j; l < d; l++)
    {
      double t = 0.0;

      for (i = 0; i < d; i++)
        t = t + V[i * d + j] * V[i * d + l];

      cov[j * d + l] = t;
      cov[l * d + j] = t;
    }

  free (y); free (col); free (sx); free (sz); free (U); free (E); free (eva); free (V);

  return 0;
}

/* covariance of the rows (all, or the listed ones) for cov_type: the
 * NcmStatsVec one-pass estimate, or one of the robust estimators */
static int
ok_cov_of_type (int cov_type, int d, int nrows, const double *X, const int *rows, double *cov)
{
  if (cov_type == OK_COV_ROBUST_DIAG)
    return ok_cov_robust_diag (d, nrows, X, rows, cov);

  if (cov_type == OK_COV_ROBUST)
    return ok_cov_robust_ogk (d, nrows, X, rows, cov);

  ok_cov_welford (d, nrows, X, rows, cov);

  return 0;
}

/* ncm_stats_dist.c:1096-1101 */
void
ok_reset (ok_sd *sd)
{
  sd->n_samples = 0;
}

/* ncm_stats_dist.c:1680-1686: the vector is duplicated */
void
ok_add_obs (ok_sd *sd, const double *x)
{
  if (sd->n_samples == sd->n_alloc)
  {
    sd->n_alloc = (sd->n_alloc == 0) ? 64 : 2 * sd->n_alloc;
    sd->X       = realloc (sd->X, sizeof (double) * (size_t) sd->n_alloc * sd->d);
  }

  memcpy (sd->X + (size_t) sd->n_samples * sd->d, x, sizeof (double) * sd->d);
  sd->n_samples++;
}

/* ncm_stats_dist.c:476-482, ncm_stats_dist_vkde.c:316-335 */
static double
ok_get_href_internal (ok_sd *sd)
{
  if (sd->kind == OK_KDE)
    return sd->over_smooth * ok_kernel_rot_bandwidth (sd->kernel, sd->nu, sd->d, sd->n_kernels);

  if (sd->use_rot_href)
    return sd->over_smooth * ok_kernel_rot_bandwidth (sd->kernel, sd->nu, sd->d, sd->n_kernels) / sd->local_frac;

  return sd->over_smooth;
}

/* ncm_stats_dist_kde.c:378-490 */
static int
ok_kde_prepare_kernel (ok_sd *sd)
{
  const int d = sd->d;
  int i, l;

  sd->cov        = malloc (sizeof (double) * d * d);
  sd->cov_decomp = malloc (sizeof (double) * d * d);

  if (sd->cov_type != OK_COV_FIXED)
  {
    if (ok_cov_of_type (sd->cov_type, d, sd->n_kernels, sd->X, NULL, sd->cov) != 0)
      return -1;

    ok_cholesky_decomp_fallback (d, sd->cov, sd->cov_decomp, sd->nearPD_maxiter);
  }
  else
  {
    if (sd->cov_fixed == NULL)
    {
      ok_set_error ("_ncm_stats_dist_kde_prepare_kernel: cov_type is FIXED but a fixed "
                    "covariance matrix was not provided, use ncm_stats_dist_kde_set_cov_fixed to set one.");
      return -1;
    }

    memcpy (sd->cov, sd->cov_fixed, sizeof (double) * d * d);
    memcpy (sd->cov_decomp, sd->cov_fixed_decomp, sizeof (double) * d * d);
  }

  sd->kernel_lnnorm = ok_kernel_lnnorm (sd->kernel, sd->nu, d, sd->cov_decomp);

  /* invUsample = X U^{-1} (gsl_blas_dtrsm Right/Upper/NoTrans): each row
   * solves y U = x by forward substitution */
  sd->invU = malloc (sizeof (double) * (size_t) sd->n_obs * d);

  for (i = 0; i < sd->n_obs; i++)
  {
    const double *x = sd->X + (size_t) i * d;
    double *y       = sd->invU + (size_t) i * d;

    for (l = 0; l < d; l++)
    {
      double t = x[l];
      int m;

      for (m = 0; m < l; m++)
        t = t - y[m] * sd->cov_decomp[m * d + l];

      y[l] = t / sd->cov_decomp[l * d + l];
    }
  }

  return 0;
}

/* ncm_stats_dist_vkde.c:362-496 */
static int
ok_vkde_build_cov_array (ok_sd *sd)
{
  const int d = sd->d;
  int i;

  sd->k       = (int) (size_t) ((sd->local_frac * sd->n_obs > 2) ? sd->local_frac * sd->n_obs : 2);
  sd->U_arr   = malloc (sizeof (double) * (size_t) sd->n_kernels * d * d);
  sd->lnnorms = malloc (sizeof (double) * sd->n_kernels);
  sd->knn     = malloc (sizeof (int) * (size_t) sd->n_kernels * sd->k);

  {
    int failed = 0;
    /* local covariances: the sample estimate for SAMPLE and FIXED
     * (ncm_stats_dist_vkde.c:460-473), the robust ones otherwise */
    const int ltype = (sd->cov_type == OK_COV_FIXED) ? OK_COV_SAMPLE : sd->cov_type;

    #pragma omp parallel for schedule(dynamic, 1) if (sd->use_threads)
    for (i = 0; i < sd->n_kernels; i++)
    {
      double *cov = malloc (sizeof (double) * d * d);
      int *nb     = sd->knn + (size_t) i * sd->k;
      double *U_i = sd->U_arr + (size_t) i * d * d;

      ok_knn (sd->n_obs, d, sd->invU, i, sd->k, nb, NULL);

      if (ok_cov_of_type (ltype, d, sd->k, sd->X, nb, cov) != 0)
      {
        #pragma omp atomic write
        failed = 1;
      }
      else
      {
        ok_cholesky_decomp_fallback (d, cov, U_i, sd->nearPD_maxiter);
        sd->lnnorms[i] = ok_kernel_lnnorm (sd->kernel, sd->nu, d, U_i);
      }

      free (cov);
    }

    if (failed)
    {
      ok_set_error_few (sd->k);
      return -1;
    }
  }

  return 0;
}

static int ok_cv_minimize (ok_sd *sd);
static int ok_cv_split_fit (ok_sd *sd, const double *m2lnp, double *IM, const double *f, const double *f1);

/* ncm_stats_dist.c:703-789 */
int
ok_prepare (ok_sd *sd)
{
  int i;

  ok_clear_prepared (sd);

  sd->n_obs     = sd->n_samples;
  sd->n_kernels = (sd->cv_type == OK_CV_SPLIT || sd->cv_type == OK_CV_SPLIT_NOFIT)
                    ? (int) ceil (sd->n_samples * sd->split_frac)
                    : sd->n_samples;

  if (sd->n_obs <= sd->d)
  {
    ok_set_error ("_ncm_stats_dist_prepare: the sample is too small.");
    return -1;
  }

  if (sd->kind == OK_VKDE)
  {
    /* ncm_stats_dist_vkde.c:498-515 */
    if (sd->local_frac * sd->n_obs < 2)
    {
      ok_set_error ("Too few observations.");
      return -1;
    }
  }

  if (ok_kde_prepare_kernel (sd) != 0)
    return -1;

  if (sd->kind == OK_VKDE)
    ok_vkde_build_cov_array (sd);

  sd->weights = malloc (sizeof (double) * sd->n_kernels);
  sd->wcum    = malloc (sizeof (double) * (sd->n_kernels + 1));
  sd->href    = ok_get_href_internal (sd);

  for (i = 0; i < sd->n_kernels; i++)
    sd->weights[i] = 1.0 / (1.0 * sd->n_kernels);

  sd->wcum_ready = 0;
  sd->prepared   = 1;
  sd->rnorm      = 0.0;

  /* ncm_stats_dist.c:769-788 */
  if (sd->cv_type == OK_CV_SPLIT_NOFIT || sd->cv_type == OK_CV_LOO)
    return ok_cv_minimize (sd);

  return 0;
}

/* ncm_stats_dist_kde.c:492-557 and ncm_stats_dist_vkde.c:517-606 */
int
ok_compute_IM (ok_sd *sd, double *IM)
{
  const int d      = sd->d;
  const int n_obs  = sd->n_obs;
  const int n_k    = sd->n_kernels;
  const double h   = sd->href;
  const double h2  = h * h;
  int i;

  if (sd->kind == OK_KDE)
  {
    for (i = 0; i < n_k; i++)
    {
      const double *yi = sd->invU + (size_t) i * d;
      int j;

      IM[(size_t) i * n_k + i] = 0.0;

      for (j = i + 1; j < n_k; j++)
      {
        const double *yj = sd->invU + (size_t) j * d;
        double chi2      = 0.0;
        int l;

        for (l = 0; l < d; l++)
        {
          const double t = yi[l] - yj[l];

          chi2 = chi2 + t * t;
        }

        chi2 = chi2 / h2;

        IM[(size_t) i * n_k + j] = chi2;
        IM[(size_t) j * n_k + i] = chi2;
      }
    }

    /* held-out rows of the split cross-validations (kde.c:526-546) */
    for (i = n_k; i < n_obs; i++)
    {
      const double *yi = sd->invU + (size_t) i * d;
      int j;

      for (j = 0; j < n_k; j++)
      {
        const double *yj = sd->invU + (size_t) j * d;
        double chi2      = 0.0;
        int l;

        for (l = 0; l < d; l++)
        {
          const double t = yi[l] - yj[l];

          chi2 = chi2 + t * t;
        }

        IM[(size_t) i * n_k + j] = chi2 / h2;
      }
    }

    for (i = 0; i < n_obs * n_k; i++)
      IM[i] = ok_kernel_eval_unnorm (sd->kernel, sd->nu, d, IM[i]);

    {
      const double s = exp (-(sd->kernel_lnnorm + d * log (h)));

      for (i = 0; i < n_obs * n_k; i++)
        IM[i] = IM[i] * s;
    }

    return 0;
  }

  {
    const double one_href2 = 1.0 / h2;

    #pragma omp parallel for schedule(dynamic, 1) if (sd->use_threads)
    for (i = 0; i < n_k; i++)
    {
      const double *U  = sd->U_arr + (size_t) i * d * d;
      const double *ti = sd->X + (size_t) i * d;
      double *z        = malloc (sizeof (double) * d);
      int j, l, m;

      for (j = 0; j < n_obs; j++)
      {
        const double *tj = sd->X + (size_t) j * d;
        double chi2      = 0.0;

        for (l = 0; l < d; l++)
          z[l] = tj[l] + (-1.0) * ti[l];

        /* dtrsm Right/Upper/NoTrans: z U = delta */
        for (l = 0; l < d; l++)
        {
          double t = z[l];

          for (m = 0; m < l; m++)
            t = t - z[m] * U[m * d + l];

          z[l] = t / U[l * d + l];
        }

        for (l = 0; l < d; l++)
          chi2 = chi2 + z[l] * z[l];

        IM[(size_t) j * n_k + i] = chi2 * one_href2;
      }

      free (z);
    }

    for (i = 0; i < n_obs * n_k; i++)
      IM[i] = ok_kernel_eval_unnorm (sd->kernel, sd->nu, d, IM[i]);

    {
      const double lnnorm_href = d * log (h);

      for (i = 0; i < n_k; i++)
      {
        const double norm_i = exp (sd->lnnorms[i] + lnnorm_href);
        const double s      = 1.0 / norm_i;
        int j;

        for (j = 0; j < n_obs; j++)
          IM[(size_t) j * n_k + i] = IM[(size_t) j * n_k + i] * s;
      }
    }
  }

  return 0;
}

static int
ok_sort_cmp_val (const void *a, const void *b)
{
  const ok_pair *pa = a, *pb = b;

  return ok_pair_cmp (pa, pb);
}

/* ncm_stats_dist.c:878-1094 */
int
ok_prepare_interp (ok_sd *sd, const double *m2lnp, int len)
{
  const double limit = -2.0 * 2.0 * OK_LOG_DBL_EPSILON;
  int i;

  if (ok_prepare (sd) != 0)
    return -1;

  if (len != sd->n_obs)
  {
    ok_set_error ("ncm_stats_dist_prepare_interp: m2lnp length differs from the sample size.");
    return -1;
  }

  sd->min_m2lnp = INFINITY;
  sd->max_m2lnp = -INFINITY;

  for (i = 0; i < sd->n_kernels; i++)
  {
    sd->min_m2lnp = fmin (sd->min_m2lnp, m2lnp[i]);
    sd->max_m2lnp = fmax (sd->max_m2lnp, m2lnp[i]);
  }

  if (sd->max_m2lnp - sd->min_m2lnp > limit)
  {
    /* every observation is sorted (ncm_stats_dist.c:911-914) */
    ok_pair *srt = malloc (sizeof (ok_pair) * sd->n_obs);
    int n_cut    = 0;

    for (i = 0; i < sd->n_obs; i++)
    {
      srt[i].d2 = m2lnp[i];
      srt[i].j  = i;
    }

    qsort (srt, sd->n_obs, sizeof (ok_pair), ok_sort_cmp_val);

    for (i = 0; i < sd->n_kernels; i++)
    {
      if (srt[i].d2 - sd->min_m2lnp > limit)
      {
        n_cut = i;
        break;
      }
    }

    if (n_cut < (int) (unsigned) (0.5 * sd->n_obs))
    {
      for (i = 0; i < sd->n_kernels; i++)
        sd->weights[i] = 0.1 / (sd->n_kernels - n_cut);

      /* held-out observations inside the window get no weight: the
       * reference indexes past its weight vector there (ncm_stats_dist.c:933-940) */
      for (i = 0; i < n_cut; i++)
        if (srt[i].j < sd->n_kernels)
          sd->weights[srt[i].j] = 0.9 / n_cut;

      sd->wcum_ready = 0;
      free (srt);
      return 0;
    }

    {
      double *m2lnp_cut = malloc (sizeof (double) * n_cut);
      double *Xcut      = malloc (sizeof (double) * (size_t) n_cut * sd->d);
      int j             = 0;
      int ret;

      for (i = 0; i < sd->n_obs; i++)
      {
        if (m2lnp[i] - sd->min_m2lnp <= limit)
        {
          m2lnp_cut[j] = m2lnp[i];
          memcpy (Xcut + (size_t) j * sd->d, sd->X + (size_t) i * sd->d, sizeof (double) * sd->d);
          j++;
        }
      }

      memcpy (sd->X, Xcut, sizeof (double) * (size_t) n_cut * sd->d);
      sd->n_samples = n_cut;

      ret = ok_prepare_interp (sd, m2lnp_cut, n_cut);

      free (m2lnp_cut);
      free (Xcut);
      free (srt);

      return ret;
    }
  }

  {
    const int n_obs = sd->n_obs;
    const int n_k   = sd->n_kernels;
    double *IM      = malloc (sizeof (double) * (size_t) n_obs * n_k);
    double *f       = malloc (sizeof (double) * n_obs);
    double *f1      = malloc (sizeof (double) * n_obs);

    for (i = 0; i < n_obs; i++)
    {
      f[i]  = exp (-0.5 * (m2lnp[i] - sd->min_m2lnp));
      f1[i] = 1.0;
    }

    if (sd->cv_type == OK_CV_SPLIT)
    {
      if (ok_cv_split_fit (sd, m2lnp, IM, f, f1) != 0)
      {
        free (IM);
        free (f);
        free (f1);
        return -1;
      }
    }
    else
    {
      ok_compute_IM (sd, IM);

      /* ncm_stats_dist.c:791-804: row i scaled by 1/f_i */
      for (i = 0; i < n_obs; i++)
      {
        const double s = 1.0 / f[i];
        int j;

        for (j = 0; j < n_k; j++)
          IM[(size_t) i * n_k + j] = IM[(size_t) i * n_k + j] * s;
      }

      sd->rnorm = ok_nnls_m (n_obs, n_k, IM, f1, sd->weights, sd->nnls_umethod);
    }

    free (IM);
    free (f);
    free (f1);
  }

  {
    double total = 0.0;

    for (i = 0; i < sd->n_kernels; i++)
      total += sd->weights[i];

    if (!(total > 0.0))
    {
      ok_set_error ("_ncm_stats_dist_prepare_interp: total weight is not positive.");
      return -1;
    }

    for (i = 0; i < sd->n_kernels; i++)
      sd->weights[i] = sd->weights[i] * ((1.0 - sd->shrink) / total);

    for (i = 0; i < sd->n_kernels; i++)
      sd->weights[i] = sd->weights[i] + sd->shrink / sd->n_kernels;
  }

  sd->wcum_ready = 0;

  return 0;
}

/* chi2 of every kernel at x (first loop of the eval_weights* functions) */
static void
ok_chi2_all (ok_sd *sd, const double *x, double *chi2)
{
  const int d     = sd->d;
  const double h  = sd->href;
  const double h2 = h * h;
  double *v       = malloc (sizeof (double) * d);
  int i, l, m;

  if (sd->kind == OK_KDE)
  {
    /* ncm_stats_dist_kde.c:648-667: v = U^{-T} x (dtrsv Upper/Trans) */
    for (l = 0; l < d; l++)
    {
      double t = x[l];

      for (m = 0; m < l; m++)
        t = t - sd->cov_decomp[m * d + l] * v[m];

      v[l] = t / sd->cov_decomp[l * d + l];
    }

    for (i = 0; i < sd->n_kernels; i++)
    {
      const double *yi = sd->invU + (size_t) i * d;
      double c         = 0.0;

      for (l = 0; l < d; l++)
      {
        const double t = yi[l] - v[l];

        c = c + t * t;
      }

      chi2[i] = c / h2;
    }
  }
  else
  {
    /* ncm_stats_dist_vkde.c:695-712 */
    const double one_href2 = 1.0 / h2;

    for (i = 0; i < sd->n_kernels; i++)
    {
      const double *U  = sd->U_arr + (size_t) i * d * d;
      const double *ti = sd->X + (size_t) i * d;
      double c         = 0.0;

      for (l = 0; l < d; l++)
        v[l] = x[l] + (-1.0) * ti[l];

      for (l = 0; l < d; l++)
      {
        double t = v[l];

        for (m = 0; m < l; m++)
          t = t - U[m * d + l] * v[m];

        v[l] = t / U[l * d + l];
      }

      for (l = 0; l < d; l++)
        c = c + v[l] * v[l];

      chi2[i] = c * one_href2;
    }
  }

  free (v);
}

/* ncm_stats_dist_kde.c:639-681, ncm_stats_dist_vkde.c:681-723 */
double
ok_eval_m2lnp (ok_sd *sd, const double *x)
{
  double *chi2 = malloc (sizeof (double) * sd->n_kernels);
  double gamma, lambda;

  ok_chi2_all (sd, x, chi2);

  if (sd->kind == OK_KDE)
    ok_kernel_sum1 (sd->kernel, sd->nu, sd->d, sd->n_kernels, chi2, sd->weights, sd->kernel_lnnorm, &gamma, &lambda);
  else
    ok_kernel_sum0 (sd->kernel, sd->nu, sd->d, sd->n_kernels, chi2, sd->weights, sd->lnnorms, &gamma, &lambda);

  free (chi2);

  return -2.0 * (gamma + log1p (lambda) - sd->d * log (sd->href));
}

/* ncm_stats_dist_kde.c:596-637, ncm_stats_dist_vkde.c:631-679 */
double
ok_eval (ok_sd *sd, const double *x)
{
  double *chi2 = malloc (sizeof (double) * sd->n_kernels);
  double res   = 0.0;
  int i;

  ok_chi2_all (sd, x, chi2);

  for (i = 0; i < sd->n_kernels; i++)
    chi2[i] = ok_kernel_eval_unnorm (sd->kernel, sd->nu, sd->d, chi2[i]);

  if (sd->kind == OK_KDE)
  {
    double s = 0.0;

    for (i = 0; i < sd->n_kernels; i++)
      s = s + chi2[i] * sd->weights[i];

    res = s * exp (-(sd->kernel_lnnorm + sd->d * log (sd->href)));
  }
  else
  {
    double s = 0.0;

    for (i = 0; i < sd->n_kernels; i++)
    {
      const double u_i = exp (sd->lnnorms[i]);

      s += sd->weights[i] * (chi2[i] / u_i);
    }

    res = s / pow (sd->href, sd->d);
  }

  free (chi2);

  return res;
}

void
ok_eval_m2lnp_batch (ok_sd *sd, const double *X, int m, double *out)
{
  int p;

  #pragma omp parallel for schedule(dynamic, 4)
  for (p = 0; p < m; p++)
    out[p] = ok_eval_m2lnp (sd, X + (size_t) p * sd->d);
}

void
ok_eval_batch (ok_sd *sd, const double *X, int m, double *out)
{
  int p;

  #pragma omp parallel for schedule(dynamic, 4)
  for (p = 0; p < m; p++)
    out[p] = ok_eval (sd, X + (size_t) p * sd->d);
}

/* ncm_stats_dist.c:1565-1606; p = u because gsl_ran_flat(r, 0, 1) = u */
int
ok_kernel_choose (ok_sd *sd, double p)
{
  int ilo = 0, ihi = sd->n_kernels;

  if (!sd->wcum_ready)
  {
    double cum = 0.0;
    int i;

    sd->wcum[0] = cum;

    for (i = 0; i < sd->n_kernels; i++)
    {
      cum           += sd->weights[i];
      sd->wcum[i + 1] = cum;
    }

    {
      const double s = 1.0 / cum;

      for (i = 0; i <= sd->n_kernels; i++)
        sd->wcum[i] = sd->wcum[i] * s;
    }

    sd->wcum_ready = 1;
  }

  while (ihi > ilo + 1)
  {
    const int mi = (ihi + ilo) / 2;

    if (sd->wcum[mi] > p)
      ihi = mi;
    else
      ilo = mi;
  }

  return ilo;
}

/* ncm_stats_dist.c:1618-1627 with ncm_stats_dist_kernel_gauss.c:335-355 and
 * ncm_stats_dist_kernel_st.c:388-414; z holds the d standard normals and
 * chi2 the chi-squared(nu) draw (ignored for the Gauss kernel). */
int
ok_sample (ok_sd *sd, double u, const double *z, double chi2, double *x)
{
  const int i = ok_kernel_choose (sd, u);

  ok_sample_at (sd, i, z, chi2, x);

  return i;
}

/* the kernel sample of ok_sample for a given kernel i */
void
ok_sample_at (ok_sd *sd, int i, const double *z, double chi2, double *x)
{
  const int d = sd->d;
  const double *U  = (sd->kind == OK_KDE) ? sd->cov_decomp : sd->U_arr + (size_t) i * d * d;
  const double *mu = sd->X + (size_t) i * d;
  int l, m;

  for (l = 0; l < d; l++)
    x[l] = z[l] * sd->href;

  /* dtrmv Upper/Trans: x = U^T x */
  for (l = d - 1; l >= 0; l--)
  {
    double t = U[l * d + l] * x[l];

    for (m = 0; m < l; m++)
      t = t + U[m * d + l] * x[m];

    x[l] = t;
  }

  if (sd->kernel == OK_ST)
  {
    const double chi_scale = sqrt (sd->nu / chi2);

    for (l = 0; l < d; l++)
      x[l] = x[l] * chi_scale;
  }

  for (l = 0; l < d; l++)
    x[l] = x[l] + mu[l];
}

void
ok_set_weights (ok_sd *sd, const double *w)
{
  memcpy (sd->weights, w, sizeof (double) * sd->n_kernels);
  sd->wcum_ready = 0;
}

/* Checker hook: a prepared VKDE (CV_NONE) from factors computed elsewhere,
 * so that the evaluation and sampling paths (ncm_stats_dist_vkde.c:631-723,
 * ncm_stats_dist.c:1565-1627) can be checked at sizes where the oracle's own
 * prepare would take too long.  X is n x d (kernel centres = observations),
 * U the n upper Cholesky factors (d x d row-major, cov_i = U_i^T U_i), w the
 * weights, href the bandwidth.  The lnnorms are the oracle's own
 * (ncm_stats_dist_kernel_*.c) of the given factors. */
void
ok_set_prepared_vkde (ok_sd *sd, int n, const double *X, const double *U, const double *w, double href)
{
  const int d     = sd->d;
  const size_t dd = (size_t) d * d;
  int i;

  ok_clear_prepared (sd);
  free (sd->X);
  sd->X         = malloc (sizeof (double) * (size_t) n * d);
  sd->n_alloc   = n;
  sd->n_samples = n;
  memcpy (sd->X, X, sizeof (double) * (size_t) n * d);

  sd->kind      = OK_VKDE;
  sd->n_obs     = n;
  sd->n_kernels = n;
  sd->k         = 0;
  sd->href      = href;
  sd->weights   = malloc (sizeof (double) * n);
  sd->wcum      = malloc (sizeof (double) * (n + 1));
  sd->U_arr     = malloc (sizeof (double) * (size_t) n * dd);
  sd->lnnorms   = malloc (sizeof (double) * n);
  memcpy (sd->weights, w, sizeof (double) * n);
  memcpy (sd->U_arr, U, sizeof (double) * (size_t) n * dd);

  #pragma omp parallel for schedule(static)
  for (i = 0; i < n; i++)
    sd->lnnorms[i] = ok_kernel_lnnorm (sd->kernel, sd->nu, d, U + (size_t) i * dd);

  sd->wcum_ready = 0;
  sd->prepared   = 1;
}

int ok_get_dim (ok_sd *sd) { return sd->d; }
int ok_get_sample_size (ok_sd *sd) { return sd->n_obs; }
int ok_get_n_kernels (ok_sd *sd) { return sd->n_kernels; }
int ok_get_n_samples (ok_sd *sd) { return sd->n_samples; }
int ok_get_k (ok_sd *sd) { return sd->k; }
double ok_get_href (ok_sd *sd) { return sd->href; }

/* ncm_stats_dist.c:1663-1669 returns the squared residual norm */
double ok_get_rnorm (ok_sd *sd) { return sd->rnorm * sd->rnorm; }

/* ncm_stats_dist_kde.c:579-587, ncm_stats_dist_vkde.c:619-629 */
double
ok_get_lnnorm (ok_sd *sd, int i)
{
  if (sd->kind == OK_KDE)
    return sd->kernel_lnnorm + sd->d * log (sd->href);

  return sd->lnnorms[i] + sd->d * log (sd->href);
}

void
ok_get_weights (ok_sd *sd, double *w)
{
  memcpy (w, sd->weights, sizeof (double) * sd->n_kernels);
}

void
ok_get_cov_decomp (ok_sd *sd, int i, double *U)
{
  const int d = sd->d;

  if (sd->kind == OK_KDE)
    memcpy (U, sd->cov_decomp, sizeof (double) * d * d);
  else
    memcpy (U, sd->U_arr + (size_t) i * d * d, sizeof (double) * d * d);
}

void
ok_get_full_cov (ok_sd *sd, double *cov)
{
  memcpy (cov, sd->cov, sizeof (double) * sd->d * sd->d);
}

void
ok_get_full_cov_decomp (ok_sd *sd, double *U)
{
  memcpy (U, sd->cov_decomp, sizeof (double) * sd->d * sd->d);
}

void
ok_get_sample (ok_sd *sd, double *X)
{
  memcpy (X, sd->X, sizeof (double) * (size_t) sd->n_samples * sd->d);
}

void
ok_get_invUsample (ok_sd *sd, double *Y)
{
  memcpy (Y, sd->invU, sizeof (double) * (size_t) sd->n_obs * sd->d);
}

void
ok_get_knn (ok_sd *sd, int *idx)
{
  if (sd->knn)
    memcpy (idx, sd->knn, sizeof (int) * (size_t) sd->n_kernels * sd->k);
}

/* ======================================================================== */
/* Cross-validated bandwidth (ncm_stats_dist.c:484-701, 769-788, 1018-1073)  */
/* ======================================================================== */

void ok_set_cv_type (ok_sd *sd, int cv) { sd->cv_type = cv; }
void ok_set_split_frac (ok_sd *sd, double v) { sd->split_frac = v; }
double ok_get_over_smooth (ok_sd *sd) { return sd->over_smooth; }

/* GSL 2.8 nmsimplex2 (multimin/simplex2.c), the minimiser of
 * ncm_stats_dist.c:175,664-701, in one dimension: P = 2 corners, corner moves
 * x' = a c + b x_hi with a = (1-t) P/(P-1), b = (P t - 1)/(P-1), the squared
 * size S2 = sum (x_i - c)^2 / P updated incrementally by each move and
 * recomputed after a contraction towards the best corner. */
static double
ok_nm_size (const double *x1, double c, double *S2)
{
  double ss = 0.0;
  int i;

  for (i = 0; i < 2; i++)
  {
    const double t = x1[i] - c;

    ss += t * t;
  }

  *S2 = ss / 2;

  return sqrt (ss / 2);
}

static void
ok_nm_update (double *x1, double *y1, double *c, double *S2, int i, double x, double val)
{
  const int P        = 2;
  const double delta = x - x1[i];
  const double xmc   = x1[i] - *c;

  *S2 += (2.0 / P) * (xmc * delta) + ((P - 1.0) / P) * ((delta * delta) / P);
  *c   = *c - x1[i] / P;
  *c   = *c + x / P;
  x1[i] = x;
  y1[i] = val;
}

int
ok_nm_minimize1 (ok_fn1 f, void *ctx, double *x, double step, double epsabs, int maxiter, int *iter_out,
                 double *fval_out)
{
  const int P = 2;
  double x1[2], y1[2], c, S2, size, val, val2, xc, xc2, fval;
  int iter = 0, status = 1;

  x1[0] = *x;
  y1[0] = f (ctx, x1[0]);
  x1[1] = *x + step;
  y1[1] = f (ctx, x1[1]);

  if (!isfinite (y1[0]) || !isfinite (y1[1]))
    return -1;

  c    = (0.0 + x1[0] + x1[1]) * (1.0 / P);
  size = ok_nm_size (x1, c, &S2);
  fval = y1[0];

  do
  {
    int hi = 0, lo = 0, s_hi = 1, failed = 0;
    double dhi = y1[0], dlo = y1[0];

    iter++;

    if (y1[1] < dlo)
    {
      dlo = y1[1];
      lo  = 1;
    }
    else if (y1[1] > dhi)
    {
      s_hi = hi;
      dhi  = y1[1];
      hi   = 1;
    }

    /* reflection (t = -1) */
    xc  = (1.0 - -1.0) * P / (P - 1.0) * c + (P * -1.0 - 1.0) / (P - 1.0) * x1[hi];
    val = f (ctx, xc);

    if (isfinite (val) && val < y1[lo])
    {
      /* expansion (t = -2) */
      xc2  = (1.0 - -2.0) * P / (P - 1.0) * c + (P * -2.0 - 1.0) / (P - 1.0) * x1[hi];
      val2 = f (ctx, xc2);

      if (isfinite (val2) && val2 < y1[lo])
        ok_nm_update (x1, y1, &c, &S2, hi, xc2, val2);
      else
        ok_nm_update (x1, y1, &c, &S2, hi, xc, val);
    }
    else if (!isfinite (val) || val > y1[s_hi])
    {
      if (isfinite (val) && val <= y1[hi])
        ok_nm_update (x1, y1, &c, &S2, hi, xc, val);

      /* one-dimensional contraction (t = 1/2) */
      xc2  = (1.0 - 0.5) * P / (P - 1.0) * c + (P * 0.5 - 1.0) / (P - 1.0) * x1[hi];
      val2 = f (ctx, xc2);

      if (isfinite (val2) && val2 <= y1[hi])
      {
        ok_nm_update (x1, y1, &c, &S2, hi, xc2, val2);
      }
      else
      {
        const int o = 1 - lo;

        x1[o] = 0.5 * (x1[o] + x1[lo]);
        y1[o] = f (ctx, x1[o]);

        if (!isfinite (y1[o]))
          failed = 1;

        c = (0.0 + x1[0] + x1[1]) * (1.0 / P);
        ok_nm_size (x1, c, &S2);
      }
    }
    else
    {
      ok_nm_update (x1, y1, &c, &S2, hi, xc, val);
    }

    if (failed)
    {
      status = -1;
      break;
    }

    lo     = (y1[1] < y1[0]) ? 1 : 0;
    *x     = x1[lo];
    fval   = y1[lo];
    size   = (S2 > 0.0) ? sqrt (S2) : ok_nm_size (x1, c, &S2);
    status = (size < epsabs) ? 0 : 1;
  } while (status == 1 && iter < maxiter);

  if (iter_out)
    *iter_out = iter;

  if (fval_out)
    *fval_out = fval;

  return status;
}

/* levmar 2.6 dlevmar_dif (external/levmar/lm_core.c:439-850, forward
 * differences of misc_core.c:137-171) for one parameter and a zero
 * measurement vector, as called by ncm_stats_dist.c:1063-1065.  Returns the
 * stop reason (info[6] of levmar). */
int
ok_lm_dif1 (ok_res_fn func, void *ctx, double *p_io, int n, int itmax, const double *opts, int *iter_out)
{
  const double tau = opts[0], eps1 = opts[1], eps2 = opts[2], eps2_sq = opts[2] * opts[2], eps3 = opts[3];
  const double delta = opts[4];
  const int K        = 10;
  double *e    = malloc (sizeof (double) * n), *hx = malloc (sizeof (double) * n), *jac = malloc (sizeof (double) * n);
  double *wrk  = malloc (sizeof (double) * n), *wrk2 = malloc (sizeof (double) * n);
  double p     = *p_io, mu = 0.0, jacTe_inf = 0.0, p_L2 = 0.0, Dp_L2 = DBL_MAX, jacTjac = 0.0, jacTe = 0.0;
  double diag  = 0.0, p_eL2 = 0.0;
  int nu = 20, updjac = 0, updp = 1, newjac = 0, stop = 0, k, i, l;

  if (func (ctx, p, hx) != 0)
    stop = 8;

  for (i = 0; i < n; i++)
  {
    e[i] = -hx[i];
    p_eL2 += e[i] * e[i];
  }

  if (!isfinite (p_eL2))
    stop = 7;

  for (k = 0; k < itmax && !stop; ++k)
  {
    if (p_eL2 <= eps3)
    {
      stop = 6;
      break;
    }

    if ((updp && nu > 16) || updjac == K)
    {
      double dstep = fabs (1e-4 * p), inv;

      if (dstep < delta)
        dstep = delta;

      if (func (ctx, p + dstep, wrk) != 0)
      {
        stop = 8;
        break;
      }

      inv = 1.0 / dstep;

      for (i = 0; i < n; i++)
        jac[i] = (wrk[i] - hx[i]) * inv;

      nu     = 2;
      updjac = 0;
      updp   = 0;
      newjac = 1;
    }

    if (newjac)
    {
      newjac  = 0;
      jacTjac = 0.0;
      jacTe   = 0.0;

      for (l = n; l-- > 0;)
      {
        jacTjac += jac[l] * jac[l];
        jacTe   += jac[l] * e[l];
      }

      jacTe_inf = fabs (jacTe);
      diag      = jacTjac;
      p_L2      = p * p;
    }

    if (jacTe_inf <= eps1)
    {
      Dp_L2 = 0.0;
      stop  = 1;
      break;
    }

    if (k == 0)
      mu = tau * ((diag > DBL_MIN) ? diag : DBL_MIN);

    jacTjac += mu;

    if (jacTjac != 0.0)
    {
      const double Dp  = jacTe / jacTjac;
      const double pDp = p + Dp;
      double pDp_eL2   = 0.0, dF, dL;

      Dp_L2 = Dp * Dp;

      if (Dp_L2 <= eps2_sq * p_L2)
      {
        stop = 2;
        break;
      }

      if (Dp_L2 >= (p_L2 + eps2) / (DBL_EPSILON * DBL_EPSILON))
      {
        stop = 4;
        break;
      }

      if (func (ctx, pDp, wrk) != 0)
      {
        stop = 8;
        break;
      }

      for (i = 0; i < n; i++)
      {
        wrk2[i] = -wrk[i];
        pDp_eL2 += wrk2[i] * wrk2[i];
      }

      if (!isfinite (pDp_eL2))
      {
        stop = 7;
        break;
      }

      dF = p_eL2 - pDp_eL2;

      if (updp || dF > 0.0)
      {
        for (i = 0; i < n; i++)
        {
          const double t = (wrk[i] - hx[i] - jac[i] * Dp) / Dp_L2;

          jac[i] += t * Dp;
        }

        ++updjac;
        newjac = 1;
      }

      dL = Dp * (mu * Dp + jacTe);

      if (dL > 0.0 && dF > 0.0)
      {
        double t = 2.0 * dF / dL - 1.0;

        t  = 1.0 - t * t * t;
        mu = mu * ((t >= 1.0 / 3.0) ? t : 1.0 / 3.0);
        nu = 2;
        p  = pDp;

        for (i = 0; i < n; i++)
        {
          e[i]  = wrk2[i];
          hx[i] = wrk[i];
        }

        p_eL2 = pDp_eL2;
        updp  = 1;
        continue;
      }
    }

    mu *= nu;

    if ((nu << 1) <= nu)
    {
      stop = 5;
      break;
    }

    nu      = nu << 1;
    jacTjac = diag;
  }

  if (k >= itmax && !stop)
    stop = 3;

  *p_io = p;

  if (iter_out)
    *iter_out = k;

  free (e);
  free (hx);
  free (jac);
  free (wrk);
  free (wrk2);

  return stop;
}

/* Weighted Welford / West update of NcmStatsVec for two variables and unit
 * weight (ncm_stats_vec.c:510-565): the covariance term pairs the updated
 * mean of the first variable with the old mean of the second. */
void
ok_pair_stats_update (double *st, double x0, double x1)
{
  /* st: W, W2, bias, mean0, mean1, var0, var1, cov01 */
  const double cw = st[0] + 1.0;
  const double x[2] = {x0, x1};
  int i;

  for (i = 0; i < 2; i++)
  {
    const double delta = x[i] - st[3 + i];
    const double R     = delta * 1.0 / cw;
    const double dvar  = st[0] * delta * R;

    st[3 + i] += R;
    st[5 + i] += dvar;

    if (i == 0)
      st[7] += 1.0 * (x[0] - st[3]) * (x[1] - st[4]);
  }

  st[0] = cw;
  st[1] += 1.0;
  st[2] = 1.0 / (st[0] - st[1] / st[0]);
}

/* The build's counter-based variates (Philox4x32-10 of Salmon, Moraes, Dror
 * and Shaw, SC'11): draw j of stream (key, ctr) uses counters (j, q, ctr_lo,
 * ctr_hi), q = 0, 1, ...: a 53-bit uniform, d Box-Muller normals (cosine
 * first, sine kept for the next), then nu_int squared normals summed. */
static void
ok_philox (uint32_t c[4], uint32_t k0, uint32_t k1)
{
  int r;

  for (r = 0; r < 10; r++)
  {
    const uint64_t p0 = (uint64_t) 0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t) 0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t) (p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n1 = (uint32_t) p1;
    const uint32_t n2 = (uint32_t) (p0 >> 32) ^ c[3] ^ k1;
    const uint32_t n3 = (uint32_t) p0;

    c[0] = n0;
    c[1] = n1;
    c[2] = n2;
    c[3] = n3;
    k0  += 0x9E3779B9u;
    k1  += 0xBB67AE85u;
  }
}

static double
ok_u53 (uint32_t a, uint32_t b)
{
  return ((double) (a >> 5) * 67108864.0 + (double) (b >> 6)) * (1.0 / 9007199254740992.0);
}

typedef struct
{
  uint32_t j, q, c0, c1, k0, k1;
  double spare;
  int has_spare;
} ok_stream;

static void
ok_stream_next (ok_stream *st, uint32_t r[4])
{
  r[0] = st->j;
  r[1] = st->q++;
  r[2] = st->c0;
  r[3] = st->c1;
  ok_philox (r, st->k0, st->k1);
}

static double
ok_stream_normal (ok_stream *st)
{
  uint32_t r[4];
  double a, b, s;

  if (st->has_spare)
  {
    st->has_spare = 0;
    return st->spare;
  }

  ok_stream_next (st, r);
  a = 1.0 - ok_u53 (r[0], r[1]);
  b = ok_u53 (r[2], r[3]);
  s = sqrt (-2.0 * log (a));

  st->spare     = s * sin (2.0 * M_PI * b);
  st->has_spare = 1;

  return s * cos (2.0 * M_PI * b);
}

void
ok_draw (uint64_t key, uint64_t ctr, uint32_t j, int d, int nu_int, double *u, double *z, double *chi2)
{
  ok_stream st = {j, 0u, (uint32_t) ctr, (uint32_t) (ctr >> 32), (uint32_t) key, (uint32_t) (key >> 32), 0.0, 0};
  uint32_t r[4];
  int i;

  ok_stream_next (&st, r);
  *u = ok_u53 (r[0], r[1]);

  for (i = 0; i < d; i++)
    z[i] = ok_stream_normal (&st);

  *chi2 = 0.0;

  for (i = 0; i < nu_int; i++)
  {
    const double g = ok_stream_normal (&st);

    *chi2 += g * g;
  }
}

/* normal number ctr of the build's host stream (key, tag 0x43565350) */
double
ok_host_normal (uint64_t key, uint64_t ctr)
{
  uint32_t c[4] = {(uint32_t) ctr, 0u, (uint32_t) (ctr >> 32), 0x43565350u};
  double a, b;

  ok_philox (c, (uint32_t) key, (uint32_t) (key >> 32));
  a = 1.0 - ok_u53 (c[0], c[1]);
  b = ok_u53 (c[2], c[3]);

  return sqrt (-2.0 * log (a)) * cos (2.0 * M_PI * b);
}

static void
ok_cv_set_lnos (ok_sd *sd, double lnos, double factor)
{
  sd->over_smooth = exp (lnos);
  sd->href        = factor * ok_get_href_internal (sd);
}

/* _ncm_stats_dist_m2lnp (ncm_stats_dist.c:484-510) */
static double
ok_cv_obj_m2lnp (void *ctx, double lnos)
{
  ok_sd *sd    = ctx;
  double m2lnp = 0.0;
  int i;

  ok_cv_set_lnos (sd, lnos, 1.0);

  for (i = sd->n_kernels; i < sd->n_obs; i++)
    m2lnp += ok_eval_m2lnp (sd, sd->X + (size_t) i * sd->d);

  return m2lnp;
}

/* _ncm_stats_dist_amise_kde_gauss (ncm_stats_dist.c:512-559) */
static double
ok_cv_obj_amise_kde_gauss (void *ctx, double lnos)
{
  ok_sd *sd           = ctx;
  const int n         = sd->n_kernels;
  const double npairs = (double) (unsigned) ((unsigned) n * ((unsigned) n - 1u));
  double *IM          = malloc (sizeof (double) * (size_t) n * n);
  double amise        = 0.0;
  int i, j;

  ok_cv_set_lnos (sd, lnos, sqrt (2.0));
  ok_compute_IM (sd, IM);

  for (i = 0; i < n; i++)
    for (j = 0; j < n; j++)
      amise += IM[(size_t) i * n + j] / ((double) n * (double) n);

  ok_cv_set_lnos (sd, lnos, 1.0);
  ok_compute_IM (sd, IM);

  for (i = 0; i < n; i++)
  {
    for (j = 0; j < i; j++)
      amise -= 2.0 * IM[(size_t) i * n + j] / npairs;

    for (j = i + 1; j < n; j++)
      amise -= 2.0 * IM[(size_t) i * n + j] / npairs;
  }

  free (IM);

  return amise;
}

/* pairs of one device batch of the Monte-Carlo integral; pair t uses draws
 * 2 (t mod B) and 2 (t mod B) + 1 of stream (key 0, t / B) */
#define OK_CV_MC_BATCH 4096

static int
ok_cmp_lp (const void *a, const void *b)
{
  return ok_pair_cmp (a, b);
}

/* _ncm_stats_dist_amise (ncm_stats_dist.c:563-660) with the build's
 * counter-based variates in place of the reference's mt19937 stream */
static double
ok_cv_obj_amise (void *ctx, double lnos)
{
  ok_sd *sd           = ctx;
  const int n         = sd->n_kernels;
  const int d         = sd->d;
  const double npairs = (double) (unsigned) ((unsigned) n * ((unsigned) n - 1u));
  const int nu_int    = (sd->kernel == OK_ST) ? (int) sd->nu : 0;
  double *IM          = malloc (sizeof (double) * (size_t) n * n);
  ok_pair *srt        = malloc (sizeof (ok_pair) * n);
  double *z1 = malloc (sizeof (double) * d), *z2 = malloc (sizeof (double) * d);
  double *x1 = malloc (sizeof (double) * d), *x2 = malloc (sizeof (double) * d);
  double st[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double amise = 0.0, mean = 0.0;
  uint64_t t;
  int i, j;

  ok_cv_set_lnos (sd, lnos, 1.0);
  ok_compute_IM (sd, IM);

  for (i = 0; i < n; i++)
  {
    double row_sum = 0.0;

    for (j = 0; j < i; j++)
      row_sum += IM[(size_t) i * n + j];

    for (j = i + 1; j < n; j++)
      row_sum += IM[(size_t) i * n + j];

    amise -= 2.0 * row_sum / npairs;

    srt[i].d2 = (row_sum + IM[(size_t) i * n + i]) / n;
    srt[i].j  = i;
  }

  qsort (srt, n, sizeof (ok_pair), ok_cmp_lp);

  for (t = 0;; t++)
  {
    const uint64_t b  = t / OK_CV_MC_BATCH;
    const uint32_t jj = (uint32_t) (t % OK_CV_MC_BATCH);
    double u1, u2, c1, c2, p1, p2;
    int ci;

    ok_draw (0, b, 2 * jj, d, nu_int, &u1, z1, &c1);
    ok_draw (0, b, 2 * jj + 1, d, nu_int, &u2, z2, &c2);
    ci = ok_kernel_choose (sd, u1);
    ok_sample_at (sd, srt[ci].j, z1, c1, x1);
    ok_sample_at (sd, srt[n - 1 - ci].j, z2, c2, x2);
    p1 = ok_eval (sd, x1);
    p2 = ok_eval (sd, x2);
    ok_pair_stats_update (st, p1, p2);

    if (t + 1 <= 100)
      continue;

    {
      const uint64_t it = t + 1 - 101;
      double var, msd;

      mean = 0.5 * (st[3] + st[4]);
      var  = 0.25 * (st[5] * st[2] + st[6] * st[2] + 2.0 * (st[7] * st[2]));
      msd  = sqrt (var / (it + 101.0)) / mean;

      if (msd < 1.0e-2 || it + 1 >= 100000000)
        break;
    }
  }

  amise += mean;

  free (IM);
  free (srt);
  free (z1);
  free (z2);
  free (x1);
  free (x2);

  return amise;
}

/* the objective of the CV type at ln(over_smooth) (sets over_smooth and
 * href as the reference's objectives do); for the pinning tests */
double
ok_cv_objective (ok_sd *sd, double lnos)
{
  if (sd->cv_type == OK_CV_SPLIT_NOFIT)
    return ok_cv_obj_m2lnp (sd, lnos);

  if (sd->kind == OK_KDE && sd->kernel == OK_GAUSS)
    return ok_cv_obj_amise_kde_gauss (sd, lnos);

  return ok_cv_obj_amise (sd, lnos);
}

/* _ncm_stats_dist_minimize_obj (ncm_stats_dist.c:664-701) */
static int
ok_cv_minimize (ok_sd *sd)
{
  double x = log (sd->over_smooth), fval;
  int iter;
  ok_fn1 obj;

  if (sd->cv_type == OK_CV_LOO && sd->kernel == OK_ST && (double) (int) sd->nu != sd->nu)
  {
    ok_set_error ("the leave-one-out Monte-Carlo integral needs an integer Student-t nu");
    return -1;
  }

  if (sd->cv_type == OK_CV_SPLIT_NOFIT)
    obj = ok_cv_obj_m2lnp;
  else if (sd->kind == OK_KDE && sd->kernel == OK_GAUSS)
    obj = ok_cv_obj_amise_kde_gauss;
  else
    obj = ok_cv_obj_amise;

  if (ok_nm_minimize1 (obj, sd, &x, 0.1, 1.0e-3, 1000, &iter, &fval) == -1 && iter == 0)
  {
    ok_set_error ("nm_minimize: non-finite function value at the start");
    return -1;
  }

  return 0;
}

typedef struct
{
  ok_sd *sd;
  const double *m2lnp, *f, *f1;
  double *IM;
  double rnorm;
} ok_split_ctx;

/* IM scaled by 1/f and its NNLS fit at ln(over_smooth) = lnos */
static void
ok_cv_fit_at (ok_split_ctx *c, double lnos)
{
  ok_sd *sd = c->sd;
  const int n_obs = sd->n_obs, n_k = sd->n_kernels;
  int i, j;

  ok_cv_set_lnos (sd, lnos, 1.0);
  ok_compute_IM (sd, c->IM);

  for (i = 0; i < n_obs; i++)
  {
    const double s = 1.0 / c->f[i];

    for (j = 0; j < n_k; j++)
      c->IM[(size_t) i * n_k + j] = c->IM[(size_t) i * n_k + j] * s;
  }

  c->rnorm = ok_nnls_m (n_obs, n_k, c->IM, c->f1, sd->weights, sd->nnls_umethod);
}

/* _ncm_stats_dist_prepare_interp_fit_nnls_f (ncm_stats_dist.c:826-855) */
static int
ok_cv_split_res (void *ctx, double p, double *hx)
{
  ok_split_ctx *c = ctx;
  ok_sd *sd       = c->sd;
  int i;

  ok_cv_fit_at (c, p);

  for (i = 0; i < sd->n_obs; i++)
  {
    const double m2lnpt = c->m2lnp[i] - sd->min_m2lnp;
    const double m2lnpi = ok_eval_m2lnp (sd, sd->X + (size_t) i * sd->d);

    hx[i] = expm1 (-0.5 * (m2lnpi - m2lnpt));
  }

  return 0;
}

/* CV_SPLIT branch of _ncm_stats_dist_prepare_interp (ncm_stats_dist.c:1018-1073);
 * the ten restarts draw from the build's host normal stream (key 0) */
static int
ok_cv_split_fit (ok_sd *sd, const double *m2lnp, double *IM, const double *f, const double *f1)
{
  ok_split_ctx c = {sd, m2lnp, f, f1, IM, 0.0};
  const double opts[5] = {1.0e-3, 1.0e-7, 1.0e-7, 1.0e-10, 1.0e-6};
  double ln_os = log (sd->over_smooth), rnorm0;
  int i, it;

  ok_cv_fit_at (&c, ln_os);
  rnorm0 = c.rnorm;

  for (i = 0; i < 10; i++)
  {
    const double ln_os_try = ln_os + 0.5 * ok_host_normal (0, sd->cv_draws++);

    ok_cv_fit_at (&c, ln_os_try);

    if (c.rnorm < rnorm0)
    {
      ln_os  = ln_os_try;
      rnorm0 = c.rnorm;
    }
  }

  ok_lm_dif1 (ok_cv_split_res, &c, &ln_os, sd->n_obs, 10000, opts, &it);
  ok_cv_fit_at (&c, ln_os);
  sd->rnorm = c.rnorm;

  return 0;
}

/* one Philox4x32-10 block (the known-answer tests of Random123 pin it) */
void
ok_philox4x32_10 (uint32_t *c, const uint32_t *k)
{
  ok_philox (c, k[0], k[1]);
}
