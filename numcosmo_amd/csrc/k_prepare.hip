// k_prepare.hip -- FP64 prepare kernels: covariance, batched Cholesky with
// nearPD/diagonal fallback, whitening, exact kNN, evaluation factors.
//
// These kernels keep the reference's operation order per element (no FMA
// contraction, IEEE division and square root), so the local covariances, the
// Cholesky factors and the neighbour sets equal the CPU restatement's bit for
// bit.  Parallelism comes from independent matrices, matrix entries and
// queries, never from reassociating a sum.
#include "nckde_internal.h"

#pragma clang fp contract(off)

namespace nck
{

// --------------------------------------------------------------------------
// NcmStatsVec covariance (ncm_stats_vec.c:510-565, :2376-2399).  One thread
// per (a <= b) entry follows the one-pass West/Welford recurrence over the
// rows in their insertion order; the per-entry recurrence needs only the
// running means of columns a and b, which the thread tracks itself.
//
// The division R = delta / W_t (W_t = t + 1, the weight after row t) sits on
// the recurrence's critical path.  The reciprocals y_t = RN(1 / W_t) are
// formed beside each staged chunk, off that path, and the quotient is
// rebuilt from them: q = RN(delta y_t), r = delta - q W_t (exact in an FMA),
// RN(q + r y_t), which is the correctly rounded delta / W_t for a faithful q
// and a correctly rounded y_t (Markstein's theorem); r = 0 keeps q, and with
// it the sign of a zero quotient.
// --------------------------------------------------------------------------
constexpr int WF_LDS = 4096;   // doubles of LDS per block: 4096 / d rows staged per step

__device__ inline double
div_by_recip (double a, double b, double y)
{
  const double q = a * y;
  const double r = fma (-q, b, a);

  return (r == 0.0) ? q : fma (r, y, q);
}

// Thread p of a matrix: the diagonal entries (p < d) take the first d
// threads, the off-diagonal pairs a < b start at the next multiple of 64, so
// that no wave runs both recurrences.
__device__ inline bool
welford_pair (int p, int d, int &a, int &b)
{
  const int dpad = (d + 63) & ~63;

  if (p < d)
  {
    a = b = p;
    return true;
  }

  int q = p - dpad;

  if (q < 0 || q >= d * (d - 1) / 2)
    return false;

  a = 0;

  while (q >= d - 1 - a)
  {
    q -= d - 1 - a;
    ++a;
  }

  b = a + 1 + q;

  return true;
}

__global__ void __launch_bounds__ (256)
k_welford_cov (const double *__restrict__ X, int d, int nrows, const int *__restrict__ rows,
               int rows_stride, int nmats, double *__restrict__ cov_out)
{
  extern __shared__ double xs[];   // (WF_LDS / d) rows x d, then their reciprocal weights
  const long mat   = blockIdx.y + (long) blockIdx.z * gridDim.y;
  const int p      = blockIdx.x * blockDim.x + threadIdx.x;

  if (mat >= nmats)
    return;

  int a = 0, b = 0;
  const bool active = welford_pair (p, d, a, b);

  const int *rw  = rows ? rows + mat * (size_t) rows_stride : nullptr;
  const double w = 1.0;
  double mean_a = 0.0, mean_b = 0.0, acc = 0.0, weight = 0.0, weight2 = 0.0;

  const int chunk = WF_LDS / d;

  for (int t0 = 0; t0 < nrows; t0 += chunk)
  {
    const int cnt = (nrows - t0 < chunk) ? nrows - t0 : chunk;

    double *rc = xs + chunk * d;

    for (int e = threadIdx.x; e < cnt * d; e += blockDim.x)
    {
      const int t    = e / d, col = e - t * d;
      const size_t r = rw ? (size_t) rw[t0 + t] : (size_t) (t0 + t);

      xs[e] = X[r * d + col];
    }

    for (int t = threadIdx.x; t < cnt; t += blockDim.x)
      rc[t] = 1.0 / (double) (t0 + t + 1);

    __syncthreads ();

    // one row of the recurrence
    auto step = [&] (double xa, double xb, double y) {
      const double curweight = weight + w;
      const double delta_a   = xa - mean_a;
      const double R_a       = div_by_recip (delta_a * w, curweight, y);

      if (a == b)
      {
        const double dvar = weight * delta_a * R_a;

        acc    = acc + dvar;
        mean_a = mean_a + R_a;
      }
      else
      {
        const double mean_a_new = mean_a + R_a;
        const double dC         = w * (xa - mean_a_new) * (xb - mean_b);
        const double delta_b    = xb - mean_b;
        const double R_b        = div_by_recip (delta_b * w, curweight, y);

        acc    = acc + dC;
        mean_a = mean_a_new;
        mean_b = mean_b + R_b;
      }

      weight  = curweight;
      weight2 = weight2 + w * w;
    };

    if (active)
    {
      // rows four at a time, their LDS reads issued before the recurrence
      int t = 0;

      for (; t + 4 <= cnt; t += 4)
      {
        double xa[4], xb[4], y[4];

#pragma unroll
        for (int u = 0; u < 4; ++u)
        {
          xa[u] = xs[(t + u) * d + a];
          xb[u] = xs[(t + u) * d + b];
          y[u]  = rc[t + u];
        }

#pragma unroll
        for (int u = 0; u < 4; ++u)
          step (xa[u], xb[u], y[u]);
      }

      for (; t < cnt; ++t)
        step (xs[t * d + a], xs[t * d + b], rc[t]);
    }

    __syncthreads ();
  }

  if (active)
  {
    const double bias_wt = 1.0 / (weight - weight2 / weight);
    double *C            = cov_out + mat * (size_t) d * d;
    const double v       = acc * bias_wt;

    C[a * d + b] = v;
    C[b * d + a] = v;
  }
}

hipError_t
launch_welford_cov (const double *X, int d, int nrows, const int *rows, int rows_stride, int nmats,
                    double *cov, hipStream_t s)
{
  const int nthr   = ((d + 63) & ~63) + d * (d - 1) / 2;
  const int bs     = 256;
  const int gy     = nmats < 32768 ? nmats : 32768;
  const int gz     = (nmats + gy - 1) / gy;
  dim3 grid ((nthr + bs - 1) / bs, gy, gz);

  hipLaunchKernelGGL (k_welford_cov, grid, dim3 (bs), sizeof (double) * ((WF_LDS / d) * d + WF_LDS / d), s, X, d,
                      nrows, rows, rows_stride, nmats, cov);

  return hipGetLastError ();
}

// --------------------------------------------------------------------------
// log-determinant with the frexp guard (ncm_matrix.c:1157-1186) and kernel
// log-normalisation (ncm_stats_dist_kernel_gauss.c:201-207,
// ncm_stats_dist_kernel_st.c:241-253).  For the Student-t kernel the host
// passes lg + nc = lgamma(nu/2) - lgamma((nu+d)/2) + (d/2)(ln pi + ln nu).
// --------------------------------------------------------------------------
__device__ double
chol_lndet (const double *U, int d, int lda)
{
  const double lb = 1.0e-200;
  const double ub = 1.0e+200;
  double detL     = 1.0;
  long exponent   = 0;

  for (int i = 0; i < d; i++)
  {
    const double Lii   = fabs (U[(size_t) i * lda + i]);
    const double ndetL = detL * Lii;

    if ((ndetL < lb) || (ndetL > ub))
    {
      int e = 0;

      detL      = frexp (ndetL, &e);
      exponent += e;
    }
    else
    {
      detL = ndetL;
    }
  }

  return 2.0 * (log (detL) + exponent * 0.69314718055994530942);
}

__device__ double
kernel_lnnorm (const double *U, int d, int lda, int kernel, double lnnorm_const)
{
  const double lndet = chol_lndet (U, d, lda);

  if (kernel == NCKDE_KERNEL_GAUSS)
    return 0.5 * ((double) (unsigned) d * LN2PI + lndet);

  return lnnorm_const + 0.5 * lndet;
}

// --------------------------------------------------------------------------
// Batched Cholesky, dpotrf on the row-major upper triangle (ncm_lapack.c:
// 215-237), one 64-thread block per matrix held in LDS.  Column j: the pivot
// by one lane, then the row j entries right of the pivot in parallel; each
// entry's dot product runs in the serial order.  The strict lower triangle is
// left as the input had it, as LAPACK does.
// --------------------------------------------------------------------------
__global__ void
k_chol_batched (const double *__restrict__ cov, double *__restrict__ Uout, int *__restrict__ info,
                double *__restrict__ lnnorm, int d, int nmats, int kernel, double lnnorm_const)
{
  extern __shared__ double S[];
  __shared__ int s_info;
  const long mat = blockIdx.x + (long) blockIdx.y * gridDim.x;
  const int tid  = threadIdx.x;

  if (mat >= nmats)
    return;

  const double *A = cov + mat * (size_t) d * d;

  for (int i = tid; i < d * d; i += blockDim.x)
    S[i] = A[i];

  if (tid == 0)
    s_info = 0;

  __syncthreads ();

  for (int j = 0; j < d; ++j)
  {
    if (tid == 0)
    {
      double s = S[j * d + j];

      for (int k = 0; k < j; ++k)
      {
        const double ukj = S[k * d + j];

        s = s - ukj * ukj;
      }

      if (!(s > 0.0))
      {
        S[j * d + j] = s;
        s_info       = j + 1;
      }
      else
      {
        S[j * d + j] = sqrt (s);
      }
    }

    __syncthreads ();

    if (s_info != 0)
      break;

    {
      const double ujj = S[j * d + j];

      for (int i = j + 1 + tid; i < d; i += blockDim.x)
      {
        double t = S[j * d + i];

        for (int k = 0; k < j; ++k)
          t = t - S[k * d + j] * S[k * d + i];

        S[j * d + i] = t / ujj;
      }
    }

    __syncthreads ();
  }

  __syncthreads ();

  double *U = Uout + mat * (size_t) d * d;

  for (int i = tid; i < d * d; i += blockDim.x)
    U[i] = S[i];

  if (tid == 0)
  {
    info[mat] = s_info;

    if (s_info == 0)
      lnnorm[mat] = kernel_lnnorm (S, d, d, kernel, lnnorm_const);
  }
}

hipError_t
launch_chol_batched (const double *cov, double *U, int *info, double *lnnorm, int d, int nmats, int kernel,
                     double lnnorm_const, hipStream_t s)
{
  const int gx = nmats < 32768 ? nmats : 32768;
  const int gy = (nmats + gx - 1) / gx;

  hipLaunchKernelGGL (k_chol_batched, dim3 (gx, gy), dim3 (64), sizeof (double) * d * d, s, cov, U, info,
                      lnnorm, d, nmats, kernel, lnnorm_const);

  return hipGetLastError ();
}

// --------------------------------------------------------------------------
// nearPD (ncm_matrix.c:1247-1344) and the diagonal fallback
// (ncm_stats_dist_vkde.c:337-360) for the matrices whose Cholesky failed:
// cyclic Jacobi for the symmetric eigenproblem, clamp of the negative
// eigenvalues to min_pos * DBL_EPSILON, re-assembly of the upper triangle,
// Dykstra correction, diagonal restored, until the Cholesky succeeds or
// maxiter + 2 projections have run.
//
// One wave per failing matrix.  The algorithm is serial over rotations,
// projections and Cholesky columns, but each step's element updates are
// independent: a rotation (p, q) changes row k's entries (k, p), (k, q) in
// its column pass and column k's entries (p, k), (q, k) in its row pass, and
// a Cholesky column j computes every entry (j, i > j) from columns < j.  Lane
// k does element k of each pass with the serial code's own expression, so
// every value is the one the serial algorithm computes (the sums of the
// re-assembly and of the Cholesky keep their serial order per element; the
// zero test of the off-diagonal sum only asks whether some square is
// non-zero).  The Jacobi matrices and the Cholesky work sit in LDS (D = 64:
// 96 KiB); the projection's cm, D_S and eigenvectors in global scratch.
// --------------------------------------------------------------------------

// Cholesky (upper, row-major, lda = d) of the LDS matrix A by one wave;
// returns 0 or the 1-based failing column, whose pivot is left in A
__device__ int
wave_cholesky_upper (int d, double *A)
{
  const int l = threadIdx.x;

  for (int j = 0; j < d; j++)
  {
    double s = A[j * d + j];

    for (int k = 0; k < j; k++)
    {
      const double ukj = A[k * d + j];

      s = s - ukj * ukj;
    }

    __syncthreads ();

    if (!(s > 0.0))
    {
      if (l == 0)
        A[j * d + j] = s;

      __syncthreads ();

      return j + 1;
    }

    s = sqrt (s);

    if (l == 0)
      A[j * d + j] = s;

    if (l > j && l < d)
    {
      double t = A[j * d + l];

      for (int k = 0; k < j; k++)
        t = t - A[k * d + j] * A[k * d + l];

      A[j * d + l] = t / s;
    }

    __syncthreads ();
  }

  return 0;
}

// eigenvalues (ascending, eva) and eigenvectors (rows of eve, global) of the
// symmetric matrix whose upper triangle is cm (global); a, v, ord in LDS
__device__ void
wave_jacobi_eigen (int d, const double *cm, double *eva, double *eve, double *a, double *v, int *ord)
{
  const int l = threadIdx.x;
  const int dd = d * d;

  for (int e = l; e < dd; e += 64)
  {
    const int i = e / d, j = e - (e / d) * d;

    a[e] = (j >= i) ? cm[i * d + j] : cm[j * d + i];
    v[e] = (i == j) ? 1.0 : 0.0;
  }

  __syncthreads ();

  for (int sweep = 0; sweep < 100; sweep++)
  {
    int nz = 0;

    for (int e = l; e < dd; e += 64)
    {
      const int i = e / d, j = e - (e / d) * d;

      if (j > i && a[e] * a[e] != 0.0)
        nz = 1;
    }

    if (!__syncthreads_or (nz))
      break;

    for (int p = 0; p < d; p++)
    {
      for (int q = p + 1; q < d; q++)
      {
        const double apq = a[p * d + q];
        const double app = a[p * d + p];
        const double aqq = a[q * d + q];

        __syncthreads ();

        if (fabs (apq) <= 2.220446049250313e-16 * 1.0e-3 * sqrt (fabs (app * aqq)) || apq == 0.0)
        {
          if (l == 0)
          {
            a[p * d + q] = 0.0;
            a[q * d + p] = 0.0;
          }

          __syncthreads ();
          continue;
        }

        const double theta = (aqq - app) / (2.0 * apq);
        const double t     = (theta >= 0.0 ? 1.0 : -1.0) / (fabs (theta) + sqrt (theta * theta + 1.0));
        const double c     = 1.0 / sqrt (t * t + 1.0);
        const double s     = t * c;

        if (l < d)
        {
          const double akp = a[l * d + p];
          const double akq = a[l * d + q];

          a[l * d + p] = c * akp - s * akq;
          a[l * d + q] = s * akp + c * akq;
        }

        __syncthreads ();

        if (l < d)
        {
          const double apk = a[p * d + l];
          const double aqk = a[q * d + l];

          a[p * d + l] = c * apk - s * aqk;
          a[q * d + l] = s * apk + c * aqk;

          const double vkp = v[l * d + p];
          const double vkq = v[l * d + q];

          v[l * d + p] = c * vkp - s * vkq;
          v[l * d + q] = s * vkp + c * vkq;
        }

        __syncthreads ();
      }
    }
  }

  if (l == 0)
  {
    for (int i = 0; i < d; i++)
      ord[i] = i;

    for (int i = 1; i < d; i++)
    {
      const int oi = ord[i];
      int j        = i - 1;

      while (j >= 0 && a[ord[j] * d + ord[j]] > a[oi * d + oi])
      {
        ord[j + 1] = ord[j];
        j--;
      }

      ord[j + 1] = oi;
    }
  }

  __syncthreads ();

  if (l < d)
    eva[l] = a[ord[l] * d + ord[l]];

  for (int e = l; e < dd; e += 64)
  {
    const int i = e / d, j = e - (e / d) * d;

    eve[e] = v[j * d + ord[i]];
  }

  __syncthreads ();
}

template <int D>
__global__ void __launch_bounds__ (64)
k_nearpd_fix (const int *__restrict__ which, const double *__restrict__ cov, double *__restrict__ Uout,
              double *__restrict__ lnnorm, double *__restrict__ scratch, int d, int maxiter, int kernel,
              double lnnorm_const, int *__restrict__ status)
{
  __shared__ double a[D * D], v[D * D], R[D * D];
  __shared__ double eva[D], diag[D];
  __shared__ int ord[D];
  const int f     = blockIdx.x;
  const int l     = threadIdx.x;
  const int mat   = which[f];
  const int dd    = d * d;
  double *cm      = scratch + (size_t) f * 3 * dd;
  double *D_S     = cm + dd;
  double *eve     = cm + 2 * dd;
  const double *C = cov + (size_t) mat * dd;
  double *U       = Uout + (size_t) mat * dd;
  int ret         = 0;
  int iter        = 0;

  for (int e = l; e < dd; e += 64)
  {
    cm[e]  = C[e];
    D_S[e] = 0.0;
  }

  if (l < d)
    diag[l] = C[l * d + l];

  __syncthreads ();

  while (true)
  {
    for (int e = l; e < dd; e += 64)
    {
      cm[e] = cm[e] - D_S[e];
      R[e]  = cm[e];
    }

    __syncthreads ();

    wave_jacobi_eigen (d, cm, eva, eve, a, v, ord);

    double min_pos_ev = INFINITY;

    for (int i = 0; i < d; i++)
      if (eva[i] > 0.0)
        min_pos_ev = fmin (eva[i], min_pos_ev);

    // eigenvector rows scaled by the square roots of the clamped eigenvalues
    for (int e = l; e < dd; e += 64)
    {
      const int i     = e / d;
      const double sc = (eva[i] < 0.0) ? sqrt (min_pos_ev * 2.220446049250313e-16) : sqrt (eva[i]);

      eve[e] = sc * eve[e];
    }

    __syncthreads ();

    // upper triangle of E^T E, each entry summed over k in order
    for (int e = l; e < dd; e += 64)
    {
      const int i = e / d, j = e - (e / d) * d;

      if (j >= i)
      {
        double s = 0.0;

        for (int k = 0; k < d; k++)
          s = s + eve[k * d + i] * eve[k * d + j];

        cm[e] = s;
      }
    }

    __syncthreads ();

    for (int e = l; e < dd; e += 64)
    {
      const int i = e / d, j = e - (e / d) * d;

      D_S[e] = cm[e] - R[e];

      if (i == j)
        cm[e] = diag[i];

      R[e] = cm[e];
    }

    __syncthreads ();

    if ((ret = wave_cholesky_upper (d, R)) == 0)
      break;

    if (iter > maxiter)
      break;

    iter++;
  }

  if (ret != 0)
  {
    // the diagonal fallback
    for (int e = l; e < dd; e += 64)
    {
      const int i = e / d, j = e - (e / d) * d;

      R[e] = (i == j) ? C[e] : 0.0;
    }

    __syncthreads ();

    if (wave_cholesky_upper (d, R) != 0 && l == 0)
      status[0] = 1;
  }

  for (int e = l; e < dd; e += 64)
    U[e] = R[e];

  if (l == 0)
    lnnorm[mat] = kernel_lnnorm (R, d, d, kernel, lnnorm_const);
}

hipError_t
launch_nearpd_fix (const int *which, int nfix, const double *cov, double *U, double *lnnorm, double *scratch,
                   int d, int maxiter, int kernel, double lnnorm_const, int *status, hipStream_t s)
{
  if (nfix <= 0)
    return hipSuccess;

#define NEARPD_T(DM)                                                                                                \
  hipLaunchKernelGGL (k_nearpd_fix<DM>, dim3 (nfix), dim3 (64), 0, s, which, cov, U, lnnorm, scratch, d, maxiter,  \
                      kernel, lnnorm_const, status)

  if (d <= 8)
    NEARPD_T (8);
  else if (d <= 16)
    NEARPD_T (16);
  else if (d <= 32)
    NEARPD_T (32);
  else if (d <= 64)
    NEARPD_T (64);
  else
    return hipErrorInvalidValue;

#undef NEARPD_T

  return hipGetLastError ();
}

// --------------------------------------------------------------------------
// invUsample = X U^{-1} (gsl_blas_dtrsm Right/Upper/NoTrans,
// ncm_stats_dist_kde.c:477-480): one lane per row, forward substitution.
// --------------------------------------------------------------------------
__global__ void
k_whiten (const double *__restrict__ X, const double *__restrict__ Ug, double *__restrict__ Y, int n, int d)
{
  const long j = blockIdx.x * (long) blockDim.x + threadIdx.x;

  if (j >= n)
    return;

  const double *x = X + j * d;
  double *y       = Y + j * d;

  for (int l = 0; l < d; l++)
  {
    double t = x[l];

    for (int m = 0; m < l; m++)
      t = t - y[m] * Ug[m * d + l];

    y[l] = t / Ug[l * d + l];
  }
}

hipError_t
launch_whiten (const double *X, const double *Ug, double *Y, int n, int d, hipStream_t s)
{
  hipLaunchKernelGGL (k_whiten, dim3 ((n + 255) / 256), dim3 (256), 0, s, X, Ug, Y, n, d);

  return hipGetLastError ();
}

// --------------------------------------------------------------------------
// Exact k nearest neighbours in whitened space (kdtree.c:27-43 distance,
// :192-320 search; rb_knn_list.c:31-40 order).  One 256-thread block per
// query:
//   1. squared distances to every point, summed over coordinates in order,
//      stored as their IEEE bit patterns (monotone for non-negative doubles);
//      the zero keys (the query itself, duplicates) are counted apart;
//   2. radix select (passes of 8 bits) of the k-th smallest key among the
//      positive ones, starting after the leading bytes that all of them share
//      (those of the smallest and the largest);
//   3. collect keys below it, then the smallest-index points equal to it;
//   4. bitonic sort of the k (key, index) pairs; write the indices.
// The keys sit in LDS when a query's keys fit (small samples, e.g. an APES
// half-ensemble), else in HBM.  The histogram is kept per wave (4 x 256
// bins), and the lanes of a wave that share a digit add to it with one atomic.
// --------------------------------------------------------------------------
constexpr int KNN_THREADS = 256;
constexpr int KNN_KMAX    = 4096;

size_t
knn_max_k ()
{
  return KNN_KMAX;
}

// Dynamic LDS: keys[n] (when they fit), then the sort buffers skey[npow],
// sidx[npow].
constexpr int KNN_LDS_BYTES = 56 * 1024;

static int
knn_npow (int k)
{
  int p = 1;

  while (p < k)
    p <<= 1;

  return p;
}

static size_t
knn_lds_bytes (int n, int k)
{
  return (size_t) n * sizeof (uint64_t) + (size_t) knn_npow (k) * (sizeof (uint64_t) + sizeof (int));
}

bool
knn_keys_in_lds (int n, int k)
{
  return knn_lds_bytes (n, k) <= (size_t) KNN_LDS_BYTES;
}

// Histogram increment with the lanes of a wave that share a digit combined
// into one LDS atomic (the leading digits of squared distances of similar
// size are the same for most keys, so plain atomics would serialise).
__device__ inline void
hist_add_wave (unsigned *hist, bool active, unsigned digit)
{
  uint64_t pending = __ballot (active);

  while (pending)
  {
    const int leader   = __ffsll ((long long) pending) - 1;
    const unsigned dl  = __shfl ((int) digit, leader, 64);
    const uint64_t grp = __ballot (active && digit == dl) & pending;

    if ((int) (threadIdx.x & 63) == leader)
      atomicAdd (&hist[dl], (unsigned) __popcll (grp));

    pending &= ~grp;
  }
}

template <bool LDS_KEYS>
__global__ void __launch_bounds__ (KNN_THREADS)
k_knn (const double *__restrict__ Yt, int n, int d, int q0, int nq, int k, int npow, uint64_t *__restrict__ keys_all,
       int *__restrict__ out, int *__restrict__ status)
{
  extern __shared__ uint64_t knn_smem[];
  uint64_t *keys = LDS_KEYS ? knn_smem : keys_all + (size_t) blockIdx.x * n;
  uint64_t *skey = LDS_KEYS ? knn_smem + n : knn_smem;
  int *sidx      = (int *) (skey + npow);
  __shared__ unsigned hist[KNN_THREADS / 64][256];
  __shared__ uint64_t s_prefix, s_wmin[KNN_THREADS / 64], s_wmax[KNN_THREADS / 64];
  __shared__ unsigned s_wzero[KNN_THREADS / 64];
  __shared__ int s_kk, s_nlt, s_eq_taken, s_pass0, s_cnt, s_zero;
  __shared__ int s_scan[KNN_THREADS];
  __shared__ unsigned s_wsum[KNN_THREADS / 64];

  const int qb   = blockIdx.x;
  const int tid  = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;

  if (qb >= nq)
    return;

  const int q = q0 + qb;

  // 1. distances (coordinates summed in order); the extremes of the positive
  //    keys and the number of zero keys (the query itself and its
  //    duplicates).  The query's coordinates sit in LDS and each point's are
  //    loaded 16 at a time, all in flight together (a load-and-wait per
  //    coordinate made this phase latency-bound).
  __shared__ double qs[64];

  if (tid < d)
    qs[tid] = Yt[(size_t) tid * n + q];

  __syncthreads ();

  uint64_t kmin = ~(uint64_t) 0, kmax = 0;
  unsigned nzero = 0;

  for (int j = tid; j < n; j += KNN_THREADS)
  {
    double dist = 0.0;

    for (int l0 = 0; l0 < d; l0 += 16)
    {
      double yv[16];

#pragma unroll
      for (int u = 0; u < 16; ++u)
        yv[u] = (l0 + u < d) ? Yt[(size_t) (l0 + u) * n + j] : 0.0;

#pragma unroll
      for (int u = 0; u < 16; ++u)
        if (l0 + u < d)
        {
          const double t = yv[u] - qs[l0 + u];

          dist = dist + t * t;
        }
    }

    const uint64_t key = (uint64_t) __double_as_longlong (dist);

    keys[j] = key;

    if (key == 0)
      nzero++;
    else
    {
      kmin = (key < kmin) ? key : kmin;
      kmax = (key > kmax) ? key : kmax;
    }
  }

  for (int o = 32; o > 0; o >>= 1)
  {
    const uint64_t a = __shfl_xor (kmin, o, 64);
    const uint64_t b = __shfl_xor (kmax, o, 64);

    kmin = (a < kmin) ? a : kmin;
    kmax = (b > kmax) ? b : kmax;
    nzero += __shfl_xor (nzero, o, 64);
  }

  if (lane == 0)
  {
    s_wmin[wave]  = kmin;
    s_wmax[wave]  = kmax;
    s_wzero[wave] = nzero;
  }

  __syncthreads ();

  if (tid == 0)
  {
    uint64_t mn = s_wmin[0], mx = s_wmax[0];
    unsigned nz = s_wzero[0];

    for (int w = 1; w < KNN_THREADS / 64; ++w)
    {
      mn = (s_wmin[w] < mn) ? s_wmin[w] : mn;
      mx = (s_wmax[w] > mx) ? s_wmax[w] : mx;
      nz += s_wzero[w];
    }

    if ((int) nz >= k || mx == 0)
    {
      // the k smallest keys are zeros: T = 0, the first k zero keys by index
      s_pass0  = 8;
      s_prefix = 0;
      s_kk     = k;
      s_cnt    = (int) nz;
      s_zero   = 1;
    }
    else
    {
      // rank k - nz among the positive keys, which share their leading bytes
      // with the smallest and largest of them
      const uint64_t x = mn ^ mx;
      const int lead   = (x == 0) ? 8 : __clzll ((long long) x) / 8;

      s_pass0  = lead;
      s_prefix = (lead == 0) ? 0 : (lead == 8 ? mn : (mn & (~(uint64_t) 0 << (64 - 8 * lead))));
      s_kk     = k - (int) nz;
      s_cnt    = n - (int) nz;   // every positive key shares the prefix
      s_zero   = 0;
    }
  }

  __syncthreads ();

  // 2. radix select over the remaining bytes, positive keys only
  const int pass0 = s_pass0;
  uint64_t mask   = (pass0 == 0) ? 0 : (pass0 == 8 ? ~(uint64_t) 0 : (~(uint64_t) 0 << (64 - 8 * pass0)));

  for (int pass = pass0; pass < 8; ++pass)
  {
    const int shift = 56 - 8 * pass;

    for (int b = tid; b < 256 * (KNN_THREADS / 64); b += KNN_THREADS)
      (&hist[0][0])[b] = 0;

    __syncthreads ();

    const uint64_t prefix = s_prefix;

    // every lane of the wave takes part in each round (the combining needs
    // whole-wave ballots)
    for (int j0 = wave * 64; j0 < n; j0 += KNN_THREADS)
    {
      const int j         = j0 + lane;
      const uint64_t key  = (j < n) ? keys[j] : 0;
      const bool active   = (j < n) && key != 0 && (key & mask) == prefix;

      hist_add_wave (hist[wave], active, (unsigned) ((key >> shift) & 255u));
    }

    __syncthreads ();

    {
      unsigned h = 0;

      for (int w = 0; w < KNN_THREADS / 64; ++w)
        h += hist[w][tid];

      unsigned v = h;

#pragma unroll
      for (int o = 1; o < 64; o <<= 1)
      {
        const unsigned t = __shfl_up (v, o, 64);

        if (lane >= o)
          v += t;
      }

      if (lane == 63)
        s_wsum[wave] = v;

      const int kk = s_kk;

      __syncthreads ();

      unsigned base = 0;

      for (int w = 0; w < wave; ++w)
        base += s_wsum[w];

      const unsigned incl = base + v;
      const unsigned excl = incl - h;

      if (excl < (unsigned) kk && incl >= (unsigned) kk)
      {
        s_kk     = kk - (int) excl;
        s_cnt    = (int) h;
        s_prefix = prefix | ((uint64_t) tid << shift);
      }
    }

    mask |= (uint64_t) 255 << shift;
    __syncthreads ();

    // the bucket holding rank k is needed whole: the remaining bytes cannot
    // change which keys are taken
    if (s_cnt == s_kk)
      break;
  }

  // the k nearest are the keys below the bucket (key & mask) < T (zero keys
  // too, on the positive path) and need_eq keys of the bucket itself, all of
  // it unless the bucket is a single key value with more copies than needed
  // (ties, taken by smallest index)
  const uint64_t T  = s_prefix;
  const int need_eq = s_kk;
  const bool zpath  = s_zero != 0;
  const bool whole  = (s_cnt == need_eq);

  // 3. collect
  if (tid == 0)
  {
    s_nlt      = 0;
    s_eq_taken = 0;
  }

  __syncthreads ();

  for (int j = tid; j < n; j += KNN_THREADS)
  {
    const uint64_t key = keys[j];
    const bool below   = !zpath && (key == 0 || (key & mask) < T);
    const bool inb     = zpath ? (key == 0) : (key != 0 && (key & mask) == T);

    if (below || (whole && inb))
    {
      const int slot = atomicAdd (&s_nlt, 1);

      if (slot < npow)
      {
        skey[slot] = key;
        sidx[slot] = j;
      }
    }
  }

  __syncthreads ();

  const int nlt = s_nlt;

  for (int base = 0; base < n && !whole; base += KNN_THREADS)
  {
    if (s_eq_taken >= need_eq)
      break;

    const int j    = base + tid;
    const int flag = (j < n && (zpath ? keys[j] == 0 : (keys[j] != 0 && keys[j] == T))) ? 1 : 0;

    s_scan[tid] = flag;
    __syncthreads ();

    for (int off = 1; off < KNN_THREADS; off <<= 1)
    {
      const int v = (tid >= off) ? s_scan[tid - off] : 0;

      __syncthreads ();
      s_scan[tid] += v;
      __syncthreads ();
    }

    const int taken = s_eq_taken;
    const int rank  = s_scan[tid] - flag;

    if (flag && taken + rank < need_eq)
    {
      const int slot = nlt + taken + rank;

      if (slot < npow)
      {
        skey[slot] = T;
        sidx[slot] = j;
      }
    }

    __syncthreads ();

    if (tid == KNN_THREADS - 1)
      s_eq_taken = taken + s_scan[KNN_THREADS - 1];

    __syncthreads ();
  }

  if ((whole ? nlt : nlt + need_eq) != k && tid == 0)
    status[0] = 2;

  int *o = out + (size_t) q * k;

  // 4. (key, idx) order of the k taken.  Up to 512 of them: each one's rank
  // is the number of pairs below it (the indices are distinct, so the ranks
  // are a permutation), one barrier instead of a bitonic network's dozens.
  if (k <= 512)
  {
    __syncthreads ();

    for (int i = tid; i < k; i += KNN_THREADS)
    {
      const uint64_t ki = skey[i];
      const int ii      = sidx[i];
      int r             = 0;

#pragma unroll 8
      for (int jj = 0; jj < k; ++jj)
      {
        const uint64_t kj = skey[jj];

        r += (kj < ki || (kj == ki && sidx[jj] < ii)) ? 1 : 0;
      }

      o[r] = ii;
    }

    return;
  }

  // bitonic sort of (key, idx) over npow slots
  for (int i = k + tid; i < npow; i += KNN_THREADS)
  {
    skey[i] = ~(uint64_t) 0;
    sidx[i] = 0x7fffffff;
  }

  __syncthreads ();

  for (int size = 2; size <= npow; size <<= 1)
  {
    for (int stride = size >> 1; stride > 0; stride >>= 1)
    {
      for (int i = tid; i < npow; i += KNN_THREADS)
      {
        const int pj = i ^ stride;

        if (pj > i)
        {
          const bool up     = ((i & size) == 0);
          const uint64_t ka = skey[i], kb = skey[pj];
          const int ia      = sidx[i], ib = sidx[pj];
          const bool a_gt_b = (ka > kb) || (ka == kb && ia > ib);

          if (a_gt_b == up)
          {
            skey[i]  = kb;
            skey[pj] = ka;
            sidx[i]  = ib;
            sidx[pj] = ia;
          }
        }
      }

      __syncthreads ();
    }
  }

  for (int i = tid; i < k; i += KNN_THREADS)
    o[i] = sidx[i];
}

hipError_t
launch_knn (const double *Yt, int n, int d, int q0, int nq, int k, uint64_t *keys, int *out, int *status,
            hipStream_t s)
{
  const size_t sort_bytes = (size_t) knn_npow (k) * (sizeof (uint64_t) + sizeof (int));

  if (knn_keys_in_lds (n, k))
    hipLaunchKernelGGL (k_knn<true>, dim3 (nq), dim3 (KNN_THREADS), knn_lds_bytes (n, k), s, Yt, n, d, q0, nq, k,
                        knn_npow (k), keys, out, status);
  else
    hipLaunchKernelGGL (k_knn<false>, dim3 (nq), dim3 (KNN_THREADS), sort_bytes, s, Yt, n, d, q0, nq, k,
                        knn_npow (k), keys, out, status);

  return hipGetLastError ();
}

// --------------------------------------------------------------------------
// Evaluation factors.  VKDE: W_i = U_i^{-T} (lower triangular), packed by
// rows and zero-padded to D, and the padded centres.  KDE: padded whitened
// centres.  Each column of W_i solves U_i^T w = e_c by forward substitution.
// --------------------------------------------------------------------------
// One thread per (kernel, column c): column c of W_i by forward substitution,
// held in registers (the columns are independent; each entry's products are
// subtracted in the serial order), then stored with the zero padding; thread
// c also writes the padded centre coordinate c.
template <int D>
__global__ void __launch_bounds__ (256)
k_vkde_factors (const double *__restrict__ Uall, const double *__restrict__ X, int nk, int d,
                double *__restrict__ Wp, double *__restrict__ Th)
{
  const long gid = blockIdx.x * (long) blockDim.x + threadIdx.x;
  const long i   = gid / D;
  const int c    = (int) (gid - i * D);

  if (i >= nk)
    return;

  const double *U = Uall + i * (size_t) d * d;
  double *W       = Wp + i * (size_t) (D * (D + 1) / 2);
  double w[D];

#pragma unroll
  for (int l = 0; l < D; ++l)
  {
    if (c < d && l >= c && l < d)
    {
      double t = (l == c) ? 1.0 : 0.0;

#pragma unroll
      for (int m = 0; m < l; ++m)
        if (m >= c)
          t = t - U[m * d + l] * w[m];

      w[l] = t / U[l * d + l];
    }
    else
    {
      w[l] = 0.0;
    }
  }

#pragma unroll
  for (int l = 0; l < D; ++l)
    if (l >= c)
      W[l * (l + 1) / 2 + c] = w[l];

  Th[i * D + c] = (c < d) ? X[i * d + c] : 0.0;
}

hipError_t
launch_vkde_factors (const double *U, const double *X, int nk, int d, int D, double *Wp, double *Th, hipStream_t s)
{
  const long nt = (long) nk * D;
  const dim3 grid ((unsigned) ((nt + 255) / 256));

  switch (D)
  {
    case 4: hipLaunchKernelGGL (k_vkde_factors<4>, grid, dim3 (256), 0, s, U, X, nk, d, Wp, Th); break;
    case 8: hipLaunchKernelGGL (k_vkde_factors<8>, grid, dim3 (256), 0, s, U, X, nk, d, Wp, Th); break;
    case 12: hipLaunchKernelGGL (k_vkde_factors<12>, grid, dim3 (256), 0, s, U, X, nk, d, Wp, Th); break;
    case 16: hipLaunchKernelGGL (k_vkde_factors<16>, grid, dim3 (256), 0, s, U, X, nk, d, Wp, Th); break;
    case 24: hipLaunchKernelGGL (k_vkde_factors<24>, grid, dim3 (256), 0, s, U, X, nk, d, Wp, Th); break;
    case 32: hipLaunchKernelGGL (k_vkde_factors<32>, grid, dim3 (256), 0, s, U, X, nk, d, Wp, Th); break;
    case 48: hipLaunchKernelGGL (k_vkde_factors<48>, grid, dim3 (256), 0, s, U, X, nk, d, Wp, Th); break;
    case 64: hipLaunchKernelGGL (k_vkde_factors<64>, grid, dim3 (256), 0, s, U, X, nk, d, Wp, Th); break;
    default: return hipErrorInvalidValue;
  }

  return hipGetLastError ();
}

__global__ void
k_kde_factors (const double *__restrict__ Y, int nk, int d, int D, double *__restrict__ Th)
{
  const long i = blockIdx.x * (long) blockDim.x + threadIdx.x;

  if (i >= nk)
    return;

  for (int l = 0; l < D; ++l)
    Th[i * D + l] = (l < d) ? Y[i * d + l] : 0.0;
}

hipError_t
launch_kde_factors (const double *Y, int nk, int d, int D, double *Th, hipStream_t s)
{
  hipLaunchKernelGGL (k_kde_factors, dim3 ((nk + 255) / 256), dim3 (256), 0, s, Y, nk, d, D, Th);

  return hipGetLastError ();
}

// b_i = ln w_i - lnnorm_i (the per-kernel constant of sum0/sum1)
__global__ void
k_bvec (const double *__restrict__ w, const double *__restrict__ lnnorm, int nk, double *__restrict__ b)
{
  const long i = blockIdx.x * (long) blockDim.x + threadIdx.x;

  if (i >= nk)
    return;

  b[i] = log (w[i]) - (lnnorm ? lnnorm[i] : 0.0);
}

hipError_t
launch_bvec (const double *w, const double *lnnorm, int nk, double *b, hipStream_t s)
{
  hipLaunchKernelGGL (k_bvec, dim3 ((nk + 255) / 256), dim3 (256), 0, s, w, lnnorm, nk, b);

  return hipGetLastError ();
}

} // namespace nck
