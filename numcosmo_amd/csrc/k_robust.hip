// k_robust.hip -- robust covariances of row sets: the Qn scale estimator,
// the diagonal robust covariance and the Orthogonalized
// Gnanadesikan-Kettenring (OGK) covariance.
//
// Reference: ncm_stats_vec_compute_cov_robust_diag / _ogk
// (ncm_stats_vec.c:1834-2064), called for the global covariance by
// ncm_stats_dist_kde.c:423-441 and for the local ones by
// ncm_stats_dist_vkde.c:467-472.  Qn is the estimator of Rousseeuw & Croux
// (J. Am. Stat. Assoc. 88 (1993) 1273) as computed by
// gsl_stats_Qn_from_sorted_data:
//   Qn = c_n * (2.21914 * d_(k)),  d_(k) the k-th smallest x_j - x_i, i < j,
//   over the sorted data, k = h (h - 1) / 2, h = n / 2 + 1, with the
//   finite-sample factors c_n of Croux & Rousseeuw (1992).
//
// One Qn problem = one set of n values.  The values are sorted (bitonic) and
// d_(k) is found by bisection over the bit patterns of non-negative doubles
// (ordered like the values): for a threshold t, the number of pairs with
// fl(x_j - x_i) <= t is a sum over i of binary searches in the sorted data,
// and the smallest t reaching k is d_(k) itself.  Sets of up to 2048 values
// are handled by one wave with the values in LDS; larger sets (the global
// covariance of a big sample) by a 1024-thread block over a global-memory
// copy.
#include "nckde_internal.h"

#pragma clang fp contract(off)

namespace nck
{

namespace
{

constexpr int QN_WAVE_MAX = 2048;

__device__ inline double
qn_factor (int n)
{
  if (n <= 9)
  {
    switch (n)
    {
      case 2: return 0.399;
      case 3: return 0.994;
      case 4: return 0.512;
      case 5: return 0.844;
      case 6: return 0.611;
      case 7: return 0.857;
      case 8: return 0.669;
      default: return 0.872;
    }
  }

  return (n % 2) ? (double) n / (n + 1.4) : (double) n / (n + 3.8);
}

// number of j in (i, n) with v[j] - v[i] <= t (v sorted ascending)
__device__ inline long long
pairs_below (const double *v, int n, int i, double t)
{
  int lo = i + 1, hi = n;

  while (lo < hi)
  {
    const int mid = (lo + hi) >> 1;

    if (v[mid] - v[i] <= t)
      lo = mid + 1;
    else
      hi = mid;
  }

  return lo - (i + 1);
}

// Work of one problem: which values to gather.
//   mode 0: x_col                          (column scales sigma_x)
//   mode 1: y_i + y_j, mode 2: y_i - y_j   (pair scales), y = x * (1 / sigma_x)
//   mode 3: z_i = sum_l E_il y_l           (rotated scales sigma_z)
struct QnArgs
{
  const double *X;
  int d, nrows;
  const int *rows;  // nmats x rows_stride indices, or NULL: rows 0..nrows-1
  int rows_stride;
  int nmats;
  int mode;
  const double *inv_sx;  // nmats x d
  const double *E;       // nmats x d x d
  double *out;           // one value per problem
  int P;                 // padded (power-of-two) set size
  double *scratch;       // block path: nprob x P
};

__device__ inline void
problem_of (const QnArgs &a, int p, int &m, int &i, int &j, int &sgn)
{
  const int d = a.d;

  if (a.mode == 1)
  {
    // p = m * npairs + q, q enumerating i < j row by row
    const int npairs = d * (d - 1) / 2;

    m = p / npairs;
    int q = p % npairs;

    i = 0;

    while (q >= d - 1 - i)
    {
      q -= d - 1 - i;
      ++i;
    }

    j   = i + 1 + q;
    sgn = 1;
  }
  else
  {
    m   = p / d;
    i   = p % d;
    j   = 0;
    sgn = 0;
  }
}

__device__ inline double
gather_value (const QnArgs &a, int m, int i, int j, int sgn, int r)
{
  const int d      = a.d;
  const int row    = a.rows ? a.rows[(size_t) m * a.rows_stride + r] : r;
  const double *x  = a.X + (size_t) row * d;

  if (a.mode == 0)
    return x[i];

  const double *inv = a.inv_sx + (size_t) m * d;

  if (a.mode == 1)
  {
    const double yi = x[i] * inv[i];
    const double yj = x[j] * inv[j];

    return sgn > 0 ? yi + yj : yi - yj;
  }

  const double *Ei = a.E + ((size_t) m * d + i) * d;
  double t         = 0.0;

  for (int l = 0; l < d; ++l)
    t = t + Ei[l] * (x[l] * inv[l]);

  return t;
}

// ---- one wave per problem, values in LDS ---------------------------------

__device__ inline void
wave_sort (double *v, int P, int lane)
{
  for (int size = 2; size <= P; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1)
    {
      for (int e = lane; e < P / 2; e += 64)
      {
        const int lo  = (e / stride) * 2 * stride + (e % stride);
        const int hi  = lo + stride;
        const bool up = ((lo & size) == 0);
        const double a = v[lo], b = v[hi];

        if ((a > b) == up)
        {
          v[lo] = b;
          v[hi] = a;
        }
      }

      __syncthreads ();
    }
}

__device__ inline double
wave_kth_diff (const double *v, int n, long long K, int lane)
{
  unsigned long long lo = 0, hi = (unsigned long long) __double_as_longlong (v[n - 1] - v[0]);

  while (lo < hi)
  {
    const unsigned long long mid = lo + ((hi - lo) >> 1);
    const double t               = __longlong_as_double ((long long) mid);
    long long c                  = 0;

    for (int i = lane; i < n - 1; i += 64)
      c += pairs_below (v, n, i, t);

    for (int o = 32; o > 0; o >>= 1)
      c += __shfl_xor (c, o, 64);

    if (c >= K)
      hi = mid;
    else
      lo = mid + 1;
  }

  return __longlong_as_double ((long long) lo);
}

__global__ void __launch_bounds__ (64)
k_qn_wave (QnArgs a, double *out2)
{
  extern __shared__ double v[];
  const int lane = threadIdx.x;
  const int p    = blockIdx.x;
  const int n    = a.nrows;
  int m, i, j, sgn;

  problem_of (a, p, m, i, j, sgn);

  const long long h = n / 2 + 1;
  const long long K = h * (h - 1) / 2;
  double res[2];

  for (int pass = 0; pass < (a.mode == 1 ? 2 : 1); ++pass)
  {
    const int s = (pass == 0) ? 1 : -1;

    for (int r = lane; r < a.P; r += 64)
      v[r] = (r < n) ? gather_value (a, m, i, j, s, r) : INFINITY;

    __syncthreads ();
    wave_sort (v, a.P, lane);

    res[pass] = qn_factor (n) * (2.21914 * wave_kth_diff (v, n, K, lane));
    __syncthreads ();
  }

  if (lane == 0)
  {
    if (a.mode == 1)
    {
      const int d = a.d;

      // U_ij = (Qn(y_i + y_j)^2 - Qn(y_i - y_j)^2) / 4
      a.out[((size_t) m * d + i) * d + j] = 0.25 * (res[0] * res[0] - res[1] * res[1]);
    }
    else
    {
      a.out[p] = res[0];
    }
  }
}

// ---- one block per problem, values in global memory -------------------------

__global__ void __launch_bounds__ (1024)
k_qn_block (QnArgs a)
{
  __shared__ long long red[16];
  const int tid = threadIdx.x;
  const int p   = blockIdx.x;
  const int n   = a.nrows;
  double *v     = a.scratch + (size_t) p * a.P;
  int m, i, j, sgn;

  problem_of (a, p, m, i, j, sgn);

  const long long h = n / 2 + 1;
  const long long K = h * (h - 1) / 2;
  double res[2];

  for (int pass = 0; pass < (a.mode == 1 ? 2 : 1); ++pass)
  {
    const int s = (pass == 0) ? 1 : -1;

    for (int r = tid; r < a.P; r += 1024)
      v[r] = (r < n) ? gather_value (a, m, i, j, s, r) : INFINITY;

    __syncthreads ();

    for (int size = 2; size <= a.P; size <<= 1)
      for (int stride = size >> 1; stride > 0; stride >>= 1)
      {
        for (int e = tid; e < a.P / 2; e += 1024)
        {
          const int lo  = (e / stride) * 2 * stride + (e % stride);
          const int hi  = lo + stride;
          const bool up = ((lo & size) == 0);
          const double x = v[lo], y = v[hi];

          if ((x > y) == up)
          {
            v[lo] = y;
            v[hi] = x;
          }
        }

        __syncthreads ();
      }

    unsigned long long lo = 0, hi = (unsigned long long) __double_as_longlong (v[n - 1] - v[0]);

    while (lo < hi)
    {
      const unsigned long long mid = lo + ((hi - lo) >> 1);
      const double t               = __longlong_as_double ((long long) mid);
      long long c                  = 0;

      for (int q = tid; q < n - 1; q += 1024)
        c += pairs_below (v, n, q, t);

      for (int o = 32; o > 0; o >>= 1)
        c += __shfl_xor (c, o, 64);

      if ((tid & 63) == 0)
        red[tid >> 6] = c;

      __syncthreads ();

      long long tot = 0;

      for (int w = 0; w < 16; ++w)
        tot += red[w];

      __syncthreads ();

      if (tot >= K)
        hi = mid;
      else
        lo = mid + 1;
    }

    res[pass] = qn_factor (n) * (2.21914 * __longlong_as_double ((long long) lo));
    __syncthreads ();
  }

  if (tid == 0)
  {
    if (a.mode == 1)
    {
      const int d = a.d;

      a.out[((size_t) m * d + i) * d + j] = 0.25 * (res[0] * res[0] - res[1] * res[1]);
    }
    else
    {
      a.out[p] = res[0];
    }
  }
}

// ---- per-matrix steps ------------------------------------------------------

// inv_sx = 1 / sx; U = identity (its upper off-diagonal is filled by the
// pair problems)
__global__ void
k_ogk_init (const double *__restrict__ sx, int d, int nmats, double *__restrict__ inv_sx, double *__restrict__ U)
{
  const int m = blockIdx.x * blockDim.x + threadIdx.x;

  if (m >= nmats)
    return;

  for (int i = 0; i < d; ++i)
  {
    inv_sx[(size_t) m * d + i] = 1.0 / sx[(size_t) m * d + i];

    for (int j = 0; j < d; ++j)
      U[((size_t) m * d + i) * d + j] = (i == j) ? 1.0 : 0.0;
  }
}

// Cyclic Jacobi of the symmetric matrix whose upper triangle is U (the same
// sweep and rotation sequence as the oracle's stand-in for dsyevr); E gets
// the eigenvectors as rows, eigenvalues ascending.  One thread per matrix,
// scratch a, v: nmats x d x d each.
__global__ void
k_jacobi (const double *__restrict__ U, int d, int nmats, double *__restrict__ wa, double *__restrict__ wv,
          double *__restrict__ E)
{
  const int m = blockIdx.x * blockDim.x + threadIdx.x;

  if (m >= nmats)
    return;

  const double *A = U + (size_t) m * d * d;
  double *a       = wa + (size_t) m * d * d;
  double *v       = wv + (size_t) m * d * d;

  for (int i = 0; i < d; i++)
    for (int j = 0; j < d; j++)
    {
      a[i * d + j] = (j >= i) ? A[i * d + j] : A[j * d + i];
      v[i * d + j] = (i == j) ? 1.0 : 0.0;
    }

  for (int sweep = 0; sweep < 100; sweep++)
  {
    double off = 0.0;

    for (int p = 0; p < d; p++)
      for (int q = p + 1; q < d; q++)
        off += a[p * d + q] * a[p * d + q];

    if (off == 0.0)
      break;

    for (int p = 0; p < d; p++)
      for (int q = p + 1; q < d; q++)
      {
        const double apq = a[p * d + q];
        const double app = a[p * d + p];
        const double aqq = a[q * d + q];

        if (fabs (apq) <= 2.220446049250313e-16 * 1.0e-3 * sqrt (fabs (app * aqq)) || apq == 0.0)
        {
          a[p * d + q] = 0.0;
          a[q * d + p] = 0.0;
          continue;
        }

        const double theta = (aqq - app) / (2.0 * apq);
        const double t     = (theta >= 0.0 ? 1.0 : -1.0) / (fabs (theta) + sqrt (theta * theta + 1.0));
        const double c     = 1.0 / sqrt (t * t + 1.0);
        const double s     = t * c;

        for (int k = 0; k < d; k++)
        {
          const double akp = a[k * d + p];
          const double akq = a[k * d + q];

          a[k * d + p] = c * akp - s * akq;
          a[k * d + q] = s * akp + c * akq;
        }

        for (int k = 0; k < d; k++)
        {
          const double apk = a[p * d + k];
          const double aqk = a[q * d + k];

          a[p * d + k] = c * apk - s * aqk;
          a[q * d + k] = s * apk + c * aqk;
        }

        for (int k = 0; k < d; k++)
        {
          const double vkp = v[k * d + p];
          const double vkq = v[k * d + q];

          v[k * d + p] = c * vkp - s * vkq;
          v[k * d + q] = s * vkp + c * vkq;
        }
      }
  }

  // ascending eigenvalues (insertion sort on the diagonal), eigenvector i =
  // column ord[i] of v, stored as row i of E
  double *Em = E + (size_t) m * d * d;
  int ord[64];

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

  for (int i = 0; i < d; i++)
    for (int k = 0; k < d; k++)
      Em[i * d + k] = v[k * d + ord[i]];
}

// cov = V^T V with V_ij = sz_i E_ij sx_j (OGK), or diag(sx^2) (robust diag)
__global__ void
k_robust_assemble (int ogk, const double *__restrict__ sx, const double *__restrict__ sz,
                   const double *__restrict__ E, int d, int nmats, double *__restrict__ cov)
{
  const size_t e = (size_t) blockIdx.x * blockDim.x + threadIdx.x;

  if (e >= (size_t) nmats * d * d)
    return;

  const int m = (int) (e / (d * d));
  const int j = (int) (e % (d * d)) / d;
  const int l = (int) (e % d);
  double t    = 0.0;

  if (!ogk)
  {
    t = (j == l) ? sx[(size_t) m * d + j] * sx[(size_t) m * d + j] : 0.0;
  }
  else
  {
    const double *Em = E + (size_t) m * d * d;
    const double *sxm = sx + (size_t) m * d;
    const double *szm = sz + (size_t) m * d;
    const int jj      = (j <= l) ? j : l;
    const int ll      = (j <= l) ? l : j;

    for (int i = 0; i < d; i++)
    {
      const double vj = szm[i] * Em[i * d + jj] * sxm[jj];
      const double vl = szm[i] * Em[i * d + ll] * sxm[ll];

      t = t + vj * vl;
    }
  }

  cov[e] = t;
}

int
pow2_at_least (int n)
{
  int p = 1;

  while (p < n)
    p <<= 1;

  return p;
}

hipError_t
launch_qn (QnArgs a, int nprob, hipStream_t s)
{
  if (nprob <= 0)
    return hipSuccess;

  if (a.P <= QN_WAVE_MAX)
    hipLaunchKernelGGL (k_qn_wave, dim3 (nprob), dim3 (64), sizeof (double) * a.P, s, a, nullptr);
  else
    hipLaunchKernelGGL (k_qn_block, dim3 (nprob), dim3 (1024), 0, s, a);

  return hipGetLastError ();
}

} // namespace

size_t
robust_work_doubles (int type, int d, int nrows, int nmats)
{
  const size_t P  = (size_t) pow2_at_least (nrows);
  const size_t dd = (size_t) d * d;
  // sx, inv_sx, sz: 3 d; U, E, two Jacobi scratch: 4 d^2 per matrix
  size_t w = (size_t) nmats * (3 * (size_t) d + 4 * dd);

  if (P > (size_t) QN_WAVE_MAX)
  {
    const size_t nprob = (size_t) nmats * ((type == NCKDE_COV_ROBUST) ? (size_t) d * (d - 1) / 2 + d : d);

    w += nprob * P;
  }

  return w;
}

hipError_t
launch_robust_cov (int type, const double *X, int d, int nrows, const int *rows, int rows_stride, int nmats,
                   double *cov, double *work, hipStream_t s)
{
  const size_t dd = (size_t) d * d;
  double *sx      = work;
  double *inv_sx  = sx + (size_t) nmats * d;
  double *sz      = inv_sx + (size_t) nmats * d;
  double *U       = sz + (size_t) nmats * d;
  double *E       = U + (size_t) nmats * dd;
  double *wa      = E + (size_t) nmats * dd;
  double *wv      = wa + (size_t) nmats * dd;
  double *scr     = wv + (size_t) nmats * dd;
  hipError_t e;

  QnArgs a;

  a.X           = X;
  a.d           = d;
  a.nrows       = nrows;
  a.rows        = rows;
  a.rows_stride = rows_stride;
  a.nmats       = nmats;
  a.inv_sx      = inv_sx;
  a.E           = E;
  a.P           = pow2_at_least (nrows);
  a.scratch     = scr;

  // column scales
  a.mode = 0;
  a.out  = sx;

  if ((e = launch_qn (a, nmats * d, s)) != hipSuccess)
    return e;

  if (type == NCKDE_COV_ROBUST)
  {
    hipLaunchKernelGGL (k_ogk_init, dim3 ((nmats + 63) / 64), dim3 (64), 0, s, sx, d, nmats, inv_sx, U);

    // pair scales into the upper triangle of U
    a.mode = 1;
    a.out  = U;

    if ((e = launch_qn (a, nmats * (d * (d - 1) / 2), s)) != hipSuccess)
      return e;

    hipLaunchKernelGGL (k_jacobi, dim3 ((nmats + 63) / 64), dim3 (64), 0, s, U, d, nmats, wa, wv, E);

    // scales of the rotated coordinates
    a.mode = 3;
    a.out  = sz;

    if ((e = launch_qn (a, nmats * d, s)) != hipSuccess)
      return e;
  }

  const size_t tot = (size_t) nmats * dd;

  hipLaunchKernelGGL (k_robust_assemble, dim3 ((unsigned) ((tot + 255) / 256)), dim3 (256), 0, s,
                      (type == NCKDE_COV_ROBUST) ? 1 : 0, sx, sz, E, d, nmats, cov);

  return hipGetLastError ();
}

} // namespace nck
