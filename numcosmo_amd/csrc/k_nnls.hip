// k_nnls.hip -- the fallback rungs and the refinement of the NNLS
// unconstrained solve (ncm_nnls.c:574-640, 655-706).
//
// * Symmetric solve (the reference's dsysv rung, ncm_nnls.c:574-605):
//   Gaussian elimination with partial pivoting (first maximum) on the
//   symmetrised normal matrix, one column step per pair of launches over the
//   whole chip: a one-block pivot/swap kernel and an elimination kernel with
//   one block per (row, 256-column chunk) that also records |S[i][k+1]| for
//   the next pivot search.  The products and differences are rounded as
//   written (no contraction), as in the reference's LAPACK build.
// * QR rung (dgels 'N', ncm_nnls.c:608-640): Householder QR of A_P for
//   nrows >= |P|, the minimum-norm solution through the QR of A_P^T for
//   nrows < |P|.  The matrix is held as contiguous columns W[c][0..L), one
//   launch pair per reflector (norm and reflector in one block, then one
//   block per remaining column).  A column whose entries below the current
//   row are exactly zero (an exact dependency, e.g. duplicated sample points)
//   gets zero weight and the next reflector starts on the same row.
// * Refinement helpers: A_P^T r over the active list, x_P += delta, and the
//   extreme pivots of a Cholesky factor (the refinement trigger).
#include "nckde_internal.h"

namespace nck
{

// deterministic block sum (blockDim.x == 256)
__device__ inline double
block_sum256 (double v, double *red)
{
  const int tid = threadIdx.x;

  for (int o = 32; o > 0; o >>= 1)
    v += __shfl_xor (v, o, 64);

  __syncthreads ();

  if ((tid & 63) == 0)
    red[tid >> 6] = v;

  __syncthreads ();

  return (red[0] + red[1]) + (red[2] + red[3]);
}

// ---------------------------------------------------------------- LU ------

// lower triangle from the upper one; colv[i] = |S[i][0]|
__global__ void
k_lu_init (double *__restrict__ S, int np, double *__restrict__ colv)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;

  if (c >= np)
    return;

  if (c < r)
    S[(size_t) r * np + c] = S[(size_t) c * np + r];

  if (r == 0)
    colv[c] = fabs (S[c]);
}

// pivot of column k: first maximum of colv[k..np); info = k + 1 when it is
// zero; otherwise rows k and piv are swapped (columns k..np) with b
__global__ void __launch_bounds__ (1024)
k_lu_pivot (double *__restrict__ S, int np, int k, double *__restrict__ b, const double *__restrict__ colv,
            int *__restrict__ info)
{
  __shared__ double rv[1024];
  __shared__ int ri[1024];
  const int tid = threadIdx.x;

  if (*info != 0)
    return;

  double best = -1.0;
  int bi      = np;

  for (int i = k + tid; i < np; i += 1024)
  {
    const double v = colv[i];

    if (v > best)
    {
      best = v;
      bi   = i;
    }
  }

  rv[tid] = best;
  ri[tid] = bi;
  __syncthreads ();

  for (int o = 512; o > 0; o >>= 1)
  {
    if (tid < o)
    {
      const double v2 = rv[tid + o];
      const int i2    = ri[tid + o];

      if (v2 > rv[tid] || (v2 == rv[tid] && i2 < ri[tid]))
      {
        rv[tid] = v2;
        ri[tid] = i2;
      }
    }

    __syncthreads ();
  }

  const double amax = rv[0];
  const int piv     = ri[0];

  if (amax == 0.0)
  {
    if (tid == 0)
      *info = k + 1;

    return;
  }

  if (piv != k)
  {
    for (int j = k + tid; j < np; j += 1024)
    {
      const double t           = S[(size_t) k * np + j];
      S[(size_t) k * np + j]   = S[(size_t) piv * np + j];
      S[(size_t) piv * np + j] = t;
    }

    if (tid == 0)
    {
      const double t = b[k];

      b[k]   = b[piv];
      b[piv] = t;
    }
  }
}

// rows i > k: S[i][j] -= (S[i][k] / S[k][k]) S[k][j] for j > k, b likewise;
// colv[i] = |S[i][k+1]| after the update
__global__ void __launch_bounds__ (256)
k_lu_elim (double *__restrict__ S, int np, int k, double *__restrict__ b, double *__restrict__ colv,
           const int *__restrict__ info)
{
  if (*info != 0)
    return;

  const int i = k + 1 + blockIdx.y;
  const int j = k + 1 + blockIdx.x * 256 + threadIdx.x;

  const double l = __ddiv_rn (S[(size_t) i * np + k], S[(size_t) k * np + k]);

  if (j < np)
  {
    const double v = __dsub_rn (S[(size_t) i * np + j], __dmul_rn (l, S[(size_t) k * np + j]));

    S[(size_t) i * np + j] = v;

    if (j == k + 1)
      colv[i] = fabs (v);
  }

  if (blockIdx.x == 0 && threadIdx.x == 0)
    b[i] = __dsub_rn (b[i], __dmul_rn (l, b[k]));
}

hipError_t
launch_lu_factor (double *S, int np, double *b, double *colv, int *info, hipStream_t s)
{
  hipLaunchKernelGGL (k_lu_init, dim3 ((np + 255) / 256, np), dim3 (256), 0, s, S, np, colv);

  for (int k = 0; k < np; ++k)
  {
    hipLaunchKernelGGL (k_lu_pivot, dim3 (1), dim3 (1024), 0, s, S, np, k, b, colv, info);

    const int rows = np - k - 1;

    if (rows > 0)
      hipLaunchKernelGGL (k_lu_elim, dim3 ((rows + 255) / 256, rows), dim3 (256), 0, s, S, np, k, b, colv, info);
  }

  return hipGetLastError ();
}

// --------------------------------------------------------------- QR -------

// W[j][i] = A[i][P[j]] = AT[P[j]][i] (columns of A_P, length nobs);
// rh = f (ones when f is NULL)
__global__ void
k_qr_gather_cols (const double *__restrict__ AT, int nobs, const int *__restrict__ P, int np,
                  const double *__restrict__ f, double *__restrict__ W, double *__restrict__ rh)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;

  if (i >= nobs)
    return;

  if (j < np)
    W[(size_t) j * nobs + i] = AT[(size_t) P[j] * nobs + i];
  else
    rh[i] = f ? f[i] : 1.0;
}

// W[i][j] = A[i][P[j]] (rows of A_P, length np: the columns of A_P^T)
__global__ void
k_qr_gather_rows (const double *__restrict__ AT, int nobs, const int *__restrict__ P, int np,
                  double *__restrict__ W)
{
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;

  if (j < np)
    W[(size_t) i * np + j] = AT[(size_t) P[j] * nobs + i];
}

// Reflector of column k (length L) from row r = *rowp (QR) or r = k (LQ):
// alpha = -sign(w_r) |w[r..L)|, w_r -= alpha, vtv = |w[r..L)|^2 after.
// prow[k] = r, or -1 when the column is zero below r (no reflector).
__global__ void __launch_bounds__ (256)
k_qr_house (double *__restrict__ W, int L, int k, int lq, int *__restrict__ rowp, int *__restrict__ prow,
            double *__restrict__ alpha, double *__restrict__ vtv)
{
  __shared__ double red[4];
  const int r = lq ? k : *rowp;
  double *w   = W + (size_t) k * L;

  if (r >= L)
  {
    if (threadIdx.x == 0)
      prow[k] = -1;

    return;
  }

  double s = 0.0;

  for (int i = r + threadIdx.x; i < L; i += 256)
    s = fma (w[i], w[i], s);

  const double nrm2 = block_sum256 (s, red);

  if (nrm2 == 0.0)
  {
    if (threadIdx.x == 0)
    {
      prow[k]  = -1;
      alpha[k] = 0.0;
      vtv[k]   = 0.0;
    }

    return;
  }

  const double nrm = sqrt (nrm2);
  const double w0  = w[r];
  const double a   = (w0 >= 0.0) ? -nrm : nrm;
  const double v0  = w0 - a;

  // |v|^2 = nrm2 - w0^2 + v0^2, summed as the first pass with v0 in place
  double t = 0.0;

  for (int i = r + threadIdx.x; i < L; i += 256)
  {
    const double vi = (i == r) ? v0 : w[i];

    t = fma (vi, vi, t);
  }

  const double vv = block_sum256 (t, red);

  if (threadIdx.x == 0)
  {
    w[r]     = v0;
    prow[k]  = r;
    alpha[k] = a;
    vtv[k]   = vv;

    if (!lq)
      *rowp = r + 1;
  }
}

// apply reflector k to the columns j = k+1 .. ncols-1 (one block each); the
// column index ncols stands for the right-hand side rh when rh != NULL
__global__ void __launch_bounds__ (256)
k_qr_apply (double *__restrict__ W, int L, int k, int ncols, double *__restrict__ rh, const int *__restrict__ prow,
            const double *__restrict__ vtv)
{
  __shared__ double red[4];
  const int r = prow[k];

  if (r < 0)
    return;

  const int j     = k + 1 + blockIdx.x;
  const double *v = W + (size_t) k * L;
  double *c       = (j < ncols) ? W + (size_t) j * L : rh;
  double s        = 0.0;

  for (int i = r + threadIdx.x; i < L; i += 256)
    s = fma (v[i], c[i], s);

  const double sc = 2.0 * block_sum256 (s, red) / vtv[k];

  for (int i = r + threadIdx.x; i < L; i += 256)
    c[i] = fma (-sc, v[i], c[i]);
}

// R x = rh[0..rank) by columns, last first: x_k = rh[row_k] / alpha_k, then
// rh[i] -= R[i][k] x_k for the pivot rows above (R[i][k] = W[k][i]); a
// column without a reflector gets x_k = 0.  One block.
__global__ void __launch_bounds__ (1024)
k_qr_back (const double *__restrict__ W, int L, int nc, double *__restrict__ rh, const int *__restrict__ prow,
           const double *__restrict__ alpha, double *__restrict__ x)
{
  __shared__ double xk;

  for (int k = nc - 1; k >= 0; --k)
  {
    const int row = prow[k];

    if (row < 0)
    {
      if (threadIdx.x == 0)
        x[k] = 0.0;

      continue;
    }

    if (threadIdx.x == 0)
    {
      xk   = rh[row] / alpha[k];
      x[k] = xk;
    }

    __syncthreads ();

    const double *col = W + (size_t) k * L;
    const double xv   = xk;

    for (int i = threadIdx.x; i < row; i += 1024)
      rh[i] = fma (-col[i], xv, rh[i]);

    __syncthreads ();
  }
}

// minimum norm (nrows = m < np): R^T y = f (forward, R[i][k] = W[k][i] for
// i < k, R[k][k] = alpha_k; y_k = 0 without a reflector), then
// x = H_0 ... H_{m-1} [y; 0].  One block; f NULL means ones.
__global__ void __launch_bounds__ (1024)
k_lq_solve (const double *__restrict__ W, int m, int np, const double *__restrict__ f, const int *__restrict__ prow,
            const double *__restrict__ alpha, const double *__restrict__ vtv, double *__restrict__ x)
{
  __shared__ double red[16];
  const int tid = threadIdx.x;

  for (int j = tid; j < np; j += 1024)
    x[j] = 0.0;

  __syncthreads ();

  for (int k = 0; k < m; ++k)
  {
    const double *col = W + (size_t) k * np;
    double s          = 0.0;

    for (int i = tid; i < k; i += 1024)
      s = fma (col[i], x[i], s);

    for (int o = 32; o > 0; o >>= 1)
      s += __shfl_xor (s, o, 64);

    if ((tid & 63) == 0)
      red[tid >> 6] = s;

    __syncthreads ();

    if (tid == 0)
    {
      double t = 0.0;

      for (int q = 0; q < 16; ++q)
        t += red[q];

      const double fk = f ? f[k] : 1.0;

      x[k] = (prow[k] >= 0) ? (fk - t) / alpha[k] : 0.0;
    }

    __syncthreads ();
  }

  for (int k = m - 1; k >= 0; --k)
  {
    if (prow[k] < 0)
      continue;

    const double *v = W + (size_t) k * np;
    double s        = 0.0;

    for (int j = k + tid; j < np; j += 1024)
      s = fma (v[j], x[j], s);

    for (int o = 32; o > 0; o >>= 1)
      s += __shfl_xor (s, o, 64);

    if ((tid & 63) == 0)
      red[tid >> 6] = s;

    __syncthreads ();

    double t = 0.0;

    for (int q = 0; q < 16; ++q)
      t += red[q];

    const double sc = 2.0 * t / vtv[k];

    __syncthreads ();

    for (int j = k + tid; j < np; j += 1024)
      x[j] = fma (-sc, v[j], x[j]);

    __syncthreads ();
  }
}

size_t
qr_work_doubles (int nobs, int np)
{
  // W, rh, alpha, vtv
  return (size_t) nobs * np + (size_t) nobs + 2 * (size_t) np + 64;
}

hipError_t
launch_qr_lstsq (const double *AT, int nobs, const int *P, int np, const double *f, double *work, int *iwork,
                 double *x, hipStream_t s)
{
  double *W     = work;
  double *rh    = W + (size_t) nobs * np;
  double *alpha = rh + nobs;
  double *vtv   = alpha + np;
  int *rowp     = iwork;
  int *prow     = iwork + 1;

  const hipError_t e = hipMemsetAsync (rowp, 0, sizeof (int), s);

  if (e != hipSuccess)
    return e;

  if (nobs >= np)
  {
    hipLaunchKernelGGL (k_qr_gather_cols, dim3 ((nobs + 255) / 256, np + 1), dim3 (256), 0, s, AT, nobs, P, np, f, W,
                        rh);

    for (int k = 0; k < np; ++k)
    {
      hipLaunchKernelGGL (k_qr_house, dim3 (1), dim3 (256), 0, s, W, nobs, k, 0, rowp, prow, alpha, vtv);
      hipLaunchKernelGGL (k_qr_apply, dim3 (np - k), dim3 (256), 0, s, W, nobs, k, np, rh, prow, vtv);
    }

    hipLaunchKernelGGL (k_qr_back, dim3 (1), dim3 (1024), 0, s, W, nobs, np, rh, prow, alpha, x);
  }
  else
  {
    hipLaunchKernelGGL (k_qr_gather_rows, dim3 ((np + 255) / 256, nobs), dim3 (256), 0, s, AT, nobs, P, np, W);

    for (int k = 0; k < nobs; ++k)
    {
      hipLaunchKernelGGL (k_qr_house, dim3 (1), dim3 (256), 0, s, W, np, k, 1, rowp, prow, alpha, vtv);

      if (nobs - k - 1 > 0)
        hipLaunchKernelGGL (k_qr_apply, dim3 (nobs - k - 1), dim3 (256), 0, s, W, np, k, nobs, nullptr, prow, vtv);
    }

    hipLaunchKernelGGL (k_lq_solve, dim3 (1), dim3 (1024), 0, s, W, nobs, np, f, prow, alpha, vtv, x);
  }

  return hipGetLastError ();
}

// ------------------------------------------------------- refinement ------

// out[a] = sum_j AT[P[a]][j] v[j], one wave per active column
__global__ void
k_row_dot_idx (const double *__restrict__ AT, const int *__restrict__ P, int np, int K, const double *__restrict__ v,
               double *__restrict__ out)
{
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / 64;
  const int lane = threadIdx.x % 64;

  if (wave >= np)
    return;

  const double *row = AT + (size_t) P[wave] * K;
  double s          = 0.0;

  for (int j = lane; j < K; j += 64)
    s = fma (row[j], v[j], s);

  for (int o = 32; o > 0; o >>= 1)
    s += __shfl_xor (s, o, 64);

  if (lane == 0)
    out[wave] = s;
}

hipError_t
launch_row_dot_idx (const double *AT, const int *P, int np, int K, const double *v, double *out, hipStream_t s)
{
  hipLaunchKernelGGL (k_row_dot_idx, dim3 ((np + 3) / 4), dim3 (256), 0, s, AT, P, np, K, v, out);

  return hipGetLastError ();
}

// x += dx over np entries; stat[0] = max |dx|, stat[1] = max |x| (after)
__global__ void __launch_bounds__ (1024)
k_add_update (double *__restrict__ x, const double *__restrict__ dx, int np, double *__restrict__ stat)
{
  __shared__ double r0[1024], r1[1024];
  const int tid = threadIdx.x;
  double a = 0.0, b = 0.0;

  for (int i = tid; i < np; i += 1024)
  {
    const double v = x[i] + dx[i];

    x[i] = v;
    a    = fmax (a, fabs (dx[i]));
    b    = fmax (b, fabs (v));
  }

  r0[tid] = a;
  r1[tid] = b;
  __syncthreads ();

  for (int o = 512; o > 0; o >>= 1)
  {
    if (tid < o)
    {
      r0[tid] = fmax (r0[tid], r0[tid + o]);
      r1[tid] = fmax (r1[tid], r1[tid + o]);
    }

    __syncthreads ();
  }

  if (tid == 0)
  {
    stat[0] = r0[0];
    stat[1] = r1[0];
  }
}

hipError_t
launch_add_update (double *x, const double *dx, int np, double *stat, hipStream_t s)
{
  hipLaunchKernelGGL (k_add_update, dim3 (1), dim3 (1024), 0, s, x, dx, np, stat);

  return hipGetLastError ();
}

// largest and smallest diagonal entry of the factor U (row-major np x np)
__global__ void __launch_bounds__ (1024)
k_diag_range (const double *__restrict__ S, int np, double *__restrict__ out)
{
  __shared__ double r0[1024], r1[1024];
  const int tid = threadIdx.x;
  double mx = 0.0, mn = INFINITY;

  for (int i = tid; i < np; i += 1024)
  {
    const double v = S[(size_t) i * np + i];

    mx = fmax (mx, v);
    mn = fmin (mn, v);
  }

  r0[tid] = mx;
  r1[tid] = mn;
  __syncthreads ();

  for (int o = 512; o > 0; o >>= 1)
  {
    if (tid < o)
    {
      r0[tid] = fmax (r0[tid], r0[tid + o]);
      r1[tid] = fmin (r1[tid], r1[tid + o]);
    }

    __syncthreads ();
  }

  if (tid == 0)
  {
    out[0] = r0[0];
    out[1] = r1[0];
  }
}

hipError_t
launch_diag_range (const double *S, int np, double *out, hipStream_t s)
{
  hipLaunchKernelGGL (k_diag_range, dim3 (1), dim3 (1024), 0, s, S, np, out);

  return hipGetLastError ();
}

} // namespace nck
