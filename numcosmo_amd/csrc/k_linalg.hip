// k_linalg.hip -- interpolation matrix and the FP64 linear algebra of the
// active-set NNLS weight fit.
//
// The interpolation matrix is stored column-major, AT[i][j] = A[j][i]
// (kernel i, observation j), so that building it, scaling its rows and both
// matrix-vector products of the NNLS loop read and write unit-stride rows.
//
// Reference: _ncm_stats_dist_compute_IM_full (ncm_stats_dist.c:791-804),
// compute_IM (ncm_stats_dist_kde.c:492-557, ncm_stats_dist_vkde.c:517-606),
// ncm_nnls_solve and its normal-equation solvers (ncm_nnls.c:464-871).
#include "nckde_internal.h"

namespace nck
{

// --------------------------------------------------------------------------
// A[j][i] = K(chi2_ij) / (u_i h^d) / f_j,  f_j = exp(-(m2lnp_j - min)/2)
// One thread per observation j, a block per (256 observations, kernel
// chunk).  The per-kernel normalisation 1 / exp(lnnorm_i + d ln h) is formed
// once per block for 256 kernels at a time in LDS, and the Student-t kernel
// is exp(kappa log1p(chi2 / nu)) (equal to the reference's pow form to a few
// ulp).  Threads past the last observation take part in the LDS staging only.
// --------------------------------------------------------------------------
constexpr int IM_KB = 256;

// The kernel factors come through scalar loads only when the compiler knows
// that the stores into AT cannot reach them: hence the separate __restrict__
// arguments (with the pointers inside IMArgs the loop loaded every factor
// entry with vector loads and waited on each, 3x slower).
template <int D, int KERNEL, int KIND>
__global__ void __launch_bounds__ (256)
k_build_IM (IMArgs a, const double *__restrict__ Wp, const double *__restrict__ Th, double *__restrict__ AT)
{
  __shared__ double kscale[IM_KB];
  const int j     = blockIdx.x * 256 + threadIdx.x;
  const int chunk = (a.n_k + gridDim.y - 1) / gridDim.y;
  const int i0    = blockIdx.y * chunk;
  const int i1    = min (a.n_k, i0 + chunk);
  const bool live = (j < a.n_obs);
  const int jj    = live ? j : (a.n_obs - 1);
  const int d     = a.d;
  // at D >= 32 the point is read again for every kernel (from L1/L2): held
  // in registers beside the differences it left one wave per SIMD at D = 64
  // and two at D = 32 (the build took 2.0x and 1.5x as long)
  constexpr bool XREG = (D < 32);
  const double *xrow  = ((KIND == NCKDE_VKDE) ? a.X : a.Y) + (size_t) jj * d;
  double xj[XREG ? D : 1];

  if (XREG)
  {
#pragma unroll
    for (int k = 0; k < D; ++k)
      xj[k] = (k < d) ? xrow[k] : 0.0;
  }

  double rowscale = 1.0;

  if (a.m2lnp)
    rowscale = 1.0 / exp (-0.5 * (a.m2lnp[jj] - a.min_m2lnp));

  const double kappa  = -0.5 * (a.nu + (double) (unsigned) d);
  const double inv_nu = 1.0 / a.nu;

  for (int ib = i0; ib < i1; ib += IM_KB)
  {
    const int ie = min (i1, ib + IM_KB);

    if (KIND == NCKDE_VKDE)
    {
      __syncthreads ();

      if (ib + (int) threadIdx.x < ie)
        kscale[threadIdx.x] = 1.0 / exp (a.lnnorm[ib + threadIdx.x] + a.lnnorm_href);

      __syncthreads ();
    }

    if (!live)
      continue;

    for (int i = ib; i < ie; ++i)
    {
      const double *th = Th + (size_t) i * D;
      double chi2      = 0.0;
      double scale;

      if (KIND == NCKDE_VKDE)
      {
        const double *W = Wp + (size_t) i * (D * (D + 1) / 2);
        double delta[D];

#pragma unroll
        for (int k = 0; k < D; ++k)
          delta[k] = (XREG ? xj[k] : ((k < d) ? xrow[k] : 0.0)) - th[k];

#pragma unroll
        for (int r = 0; r < D; ++r)
        {
          double z = 0.0;

#pragma unroll
          for (int c = 0; c <= r; ++c)
            z = fma (W[r * (r + 1) / 2 + c], delta[c], z);

          chi2 = fma (z, z, chi2);
        }

        chi2  = chi2 * a.inv_h2;
        scale = kscale[i - ib];
      }
      else
      {
#pragma unroll
        for (int k = 0; k < D; ++k)
        {
          const double t = th[k] - (XREG ? xj[k] : ((k < d) ? xrow[k] : 0.0));

          chi2 = fma (t, t, chi2);
        }

        chi2  = chi2 * a.inv_h2;
        scale = a.kde_scale;
      }

      double K;

      if (KERNEL == NCKDE_KERNEL_GAUSS)
        K = exp (-0.5 * chi2);
      else
        K = exp (kappa * log1p (chi2 * inv_nu));

      AT[(size_t) i * a.n_obs + j] = (K * scale) * rowscale;
    }
  }
}

// Grid: about 2048 blocks, rounded to whole rounds of the blocks the device
// holds at once, so that no nearly empty last round doubles a block's time
// (n_obs = 3200: 2054 blocks were 2 rounds and 6 blocks).
template <int D, int KERNEL, int KIND>
static hipError_t
im_t (const IMArgs &a, hipStream_t s)
{
  static int resident = 0;

  if (resident == 0)
  {
    int dev = 0, ncu = 0, nb = 0;

    (void) hipGetDevice (&dev);
    (void) hipDeviceGetAttribute (&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    (void) hipOccupancyMaxActiveBlocksPerMultiprocessor (&nb, k_build_IM<D, KERNEL, KIND>, 256, 0);
    resident = (nb > 0 ? nb : 1) * (ncu > 0 ? ncu : 1);
  }

  const int bx     = (a.n_obs + 255) / 256;
  const int total0 = bx * ((2048 + bx - 1) / bx);
  const int rounds = (total0 + resident / 2) / resident;
  int gy           = (rounds > 0 ? rounds : 1) * resident / bx;

  if (gy > a.n_k)
    gy = a.n_k;

  if (gy < 1)
    gy = 1;

  hipLaunchKernelGGL ((k_build_IM<D, KERNEL, KIND>), dim3 (bx, gy), dim3 (256), 0, s, a, a.Wp, a.Th, a.AT);

  return hipGetLastError ();
}

template <int D>
static hipError_t
im_d (const IMArgs &a, hipStream_t s)
{
  if (a.kind == NCKDE_VKDE)
    return (a.kernel == NCKDE_KERNEL_GAUSS) ? im_t<D, NCKDE_KERNEL_GAUSS, NCKDE_VKDE> (a, s)
                                            : im_t<D, NCKDE_KERNEL_ST, NCKDE_VKDE> (a, s);

  return (a.kernel == NCKDE_KERNEL_GAUSS) ? im_t<D, NCKDE_KERNEL_GAUSS, NCKDE_KDE> (a, s)
                                          : im_t<D, NCKDE_KERNEL_ST, NCKDE_KDE> (a, s);
}

hipError_t
launch_build_IM (const IMArgs &a, hipStream_t s)
{
  switch (a.D)
  {
    case 4: return im_d<4> (a, s);
    case 8: return im_d<8> (a, s);
    case 12: return im_d<12> (a, s);
    case 16: return im_d<16> (a, s);
    case 24: return im_d<24> (a, s);
    case 32: return im_d<32> (a, s);
    case 48: return im_d<48> (a, s);
    case 64: return im_d<64> (a, s);
    default: return hipErrorInvalidValue;
  }
}

// out[c][r] = in[r][c]
__global__ void
k_transpose (const double *__restrict__ in, int rows, int cols, double *__restrict__ out)
{
  __shared__ double tile[32][33];
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  const int tx = threadIdx.x, ty = threadIdx.y;

  for (int k = ty; k < 32; k += 8)
  {
    const int r = r0 + k, c = c0 + tx;

    tile[k][tx] = (r < rows && c < cols) ? in[(size_t) r * cols + c] : 0.0;
  }

  __syncthreads ();

  for (int k = ty; k < 32; k += 8)
  {
    const int c = c0 + k, r = r0 + tx;

    if (r < rows && c < cols)
      out[(size_t) c * rows + r] = tile[tx][k];
  }
}

hipError_t
launch_transpose (const double *in, int rows, int cols, double *out, hipStream_t s)
{
  dim3 grid ((cols + 31) / 32, (rows + 31) / 32);

  hipLaunchKernelGGL (k_transpose, grid, dim3 (32, 8), 0, s, in, rows, cols, out);

  return hipGetLastError ();
}

// --------------------------------------------------------------------------
// Symmetric rank-K products on the FP64 matrix cores (v_mfma_f64_16x16x4f64)
// over BM x BM tiles of the upper triangle.  256 threads = 4 waves in a
// 2 x 2 layout, each wave a BM/2 x BM/2 block of 16 x 16 MFMA tiles; K
// staged through LDS in slices of 16, the next slice's global loads in
// flight during the current slice's MFMAs.  BM = 128 serves large products
// (at least 1024 tiles for the Gram matrix, 16384 for the Cholesky update)
// and BM = 64 the smaller ones: 4x more blocks, and for the Cholesky update
// (K = 64) a 4x shorter MFMA chain per block, which suits the shrinking
// trailing matrix.
//   mode 0: M = AT AT^T (AT is n x K row-major): the Gram matrix A^T A of
//           the NNLS normal equations (cblas_dsyrk, ncm_matrix.c:1547-1575);
//           both triangles are written.
//   mode 1: trailing update of the blocked Cholesky, S[a][b] -= sum_r
//           S[kb+r][a] S[kb+r][b] over the panel rows kb..kb+K-1, for
//           off <= a <= b (upper triangle only).
// Tiles are numbered along the rows of the upper triangle; the grid is
// padded to a multiple of 8 and block b takes tile (b % 8) * (grid / 8) +
// b / 8, so that each XCD (blocks are dealt round-robin over the 8 XCDs)
// works on a contiguous band of tiles that shares operand rows in its L2.
// --------------------------------------------------------------------------
typedef double f64x4 __attribute__ ((ext_vector_type (4)));

constexpr int SY_BM = 128;   // tile of the Gram matrix (MODE 0)
constexpr int SY_BK = 16;

// BM: square tile of one block (128 or 64); every thread stages
// VPT = 16 BM / 256 values of each operand per K slice, 16 threads per K row
template <int MODE, int BM>
__device__ inline void
syrk_load (const double *__restrict__ P, int n, int K, int ld, int kb, int r0, int k0, double (&v)[BM / 16])
{
  constexpr int VPT = BM / 16;
  const int tid     = threadIdx.x;

  if (MODE == 0)
  {
    // 256 / BM threads per row, VPT consecutive K entries each
    const int a  = r0 + tid / (256 / BM);
    const int kq = k0 + (tid % (256 / BM)) * VPT;

#pragma unroll
    for (int q = 0; q < VPT; ++q)
      v[q] = (a < n && kq + q < K) ? P[(size_t) a * ld + kq + q] : 0.0;
  }
  else
  {
    const int kq = k0 + tid / 16;
    const int a  = r0 + (tid % 16) * VPT;

#pragma unroll
    for (int q = 0; q < VPT; ++q)
      v[q] = (a + q < n && kq < K) ? P[(size_t) (kb + kq) * ld + a + q] : 0.0;
  }
}

template <int MODE, int BM>
__device__ inline void
syrk_store (double (*T)[BM + 4], const double (&v)[BM / 16])
{
  constexpr int VPT = BM / 16;
  const int tid     = threadIdx.x;

  if (MODE == 0)
  {
    const int rr = tid / (256 / BM), kk = (tid % (256 / BM)) * VPT;

#pragma unroll
    for (int q = 0; q < VPT; ++q)
      T[kk + q][rr] = v[q];
  }
  else
  {
    const int kk = tid / 16, rr = (tid % 16) * VPT;

#pragma unroll
    for (int q = 0; q < VPT; ++q)
      T[kk][rr + q] = v[q];
  }
}

template <int MODE, int BM>
__global__ void __launch_bounds__ (256)
k_syrk_mfma (const double *__restrict__ P, int n, int K, int ld, int off, int kb, double *__restrict__ M,
             int ntiles)
{
  static_assert (BM == 128 || BM == 64, "128- or 64-wide tiles");
  constexpr int WM = BM / 2;    // tile of one wave
  constexpr int NI = WM / 16;   // 16 x 16 MFMA blocks per wave dimension

  __shared__ double As[SY_BK][BM + 4];
  __shared__ double Bs[SY_BK][BM + 4];

  const int per = gridDim.x / 8;
  int t         = (blockIdx.x % 8) * per + blockIdx.x / 8;

  if (t >= ntiles)
    return;

  const int nt = (n - off + BM - 1) / BM;
  int ti       = 0;

  while (t >= nt - ti)
  {
    t -= nt - ti;
    ++ti;
  }

  const int tj   = ti + t;
  const int r0   = off + ti * BM, c0 = off + tj * BM;
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
  const int wr   = (w >> 1) * WM, wc = (w & 1) * WM;
  const int li   = lane % 16, lk = lane / 16;

  f64x4 acc[NI][NI];

#pragma unroll
  for (int i = 0; i < NI; ++i)
#pragma unroll
    for (int j = 0; j < NI; ++j)
      acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};

  double va[BM / 16], vb[BM / 16];

  syrk_load<MODE, BM> (P, n, K, ld, kb, r0, 0, va);
  syrk_load<MODE, BM> (P, n, K, ld, kb, c0, 0, vb);

  for (int k0 = 0; k0 < K; k0 += SY_BK)
  {
    __syncthreads ();
    syrk_store<MODE, BM> (As, va);
    syrk_store<MODE, BM> (Bs, vb);
    __syncthreads ();

    if (k0 + SY_BK < K)
    {
      syrk_load<MODE, BM> (P, n, K, ld, kb, r0, k0 + SY_BK, va);
      syrk_load<MODE, BM> (P, n, K, ld, kb, c0, k0 + SY_BK, vb);
    }

#pragma unroll
    for (int ks = 0; ks < SY_BK / 4; ++ks)
    {
      const int kq = ks * 4 + lk;
      double fa[NI], fb[NI];

#pragma unroll
      for (int i = 0; i < NI; ++i)
      {
        fa[i] = As[kq][wr + i * 16 + li];
        fb[i] = Bs[kq][wc + i * 16 + li];
      }

#pragma unroll
      for (int i = 0; i < NI; ++i)
#pragma unroll
        for (int j = 0; j < NI; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64 (fa[i], fb[j], acc[i][j], 0, 0, 0);
    }
  }

  // D[i][j] of a 16 x 16 tile: register r of a lane holds i = lane / 16 + 4 r, j = lane % 16
  if (MODE == 0)
  {
#pragma unroll
    for (int i = 0; i < NI; ++i)
#pragma unroll
      for (int j = 0; j < NI; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r)
        {
          const int a = r0 + wr + i * 16 + lk + 4 * r;
          const int b = c0 + wc + j * 16 + li;

          if (a < n && b < n && a <= b)
          {
            M[(size_t) a * n + b] = acc[i][j][r];
            M[(size_t) b * n + a] = acc[i][j][r];
          }
        }
  }
  else
  {
    // read-modify-write of the tile, one 16-column block at a time with all
    // loads of a lane in flight (out-of-range entries read element 0)
#pragma unroll
    for (int j = 0; j < NI; ++j)
    {
      const int b = c0 + wc + j * 16 + li;
      double old[NI][4];

#pragma unroll
      for (int i = 0; i < NI; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r)
        {
          const int a  = r0 + wr + i * 16 + lk + 4 * r;
          const bool v = (a < n && b < n && a <= b);

          old[i][r] = M[v ? (size_t) a * ld + b : 0];
        }

#pragma unroll
      for (int i = 0; i < NI; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r)
        {
          const int a = r0 + wr + i * 16 + lk + 4 * r;

          if (a < n && b < n && a <= b)
            M[(size_t) a * ld + b] = old[i][r] - acc[i][j][r];
        }
    }
  }
}

// row_end > 0: only the tile rows below row_end (row_end - off a multiple
// of BM), every column from the tile row on
template <int MODE, int BM = SY_BM>
static void
launch_syrk_mfma (const double *P, int n, int K, int ld, int off, int kb, double *M, hipStream_t s, int row_end = 0)
{
  const int nt     = (n - off + BM - 1) / BM;
  const int nr     = (row_end > 0) ? (row_end - off + BM - 1) / BM : nt;
  const int ntiles = nr * nt - nr * (nr - 1) / 2;
  const int grid   = (ntiles + 7) / 8 * 8;

  if (ntiles > 0)
    hipLaunchKernelGGL ((k_syrk_mfma<MODE, BM>), dim3 (grid), dim3 (256), 0, s, P, n, K, ld, off, kb, M, ntiles);
}

hipError_t
launch_gram_nt (const double *AT, int n, int K, double *M, hipStream_t s)
{
  // 128-wide tiles reuse each staged operand twice as often; below about
  // four blocks per CU their tail leaves the chip half idle, and 64-wide
  // tiles (4x the blocks) balance better
  const int nt128 = (n + 127) / 128;

  if (nt128 * (nt128 + 1) / 2 >= 1024)
    launch_syrk_mfma<0, 128> (AT, n, K, K, 0, 0, M, s);
  else
    launch_syrk_mfma<0, 64> (AT, n, K, K, 0, 0, M, s);

  return hipGetLastError ();
}

// Row sums of IM for the leave-one-out criteria (ncm_stats_dist.c:512-600):
// thread j walks the kernels i = 0..nk-1 of observation j in ascending order,
// the order of the reference's row loops, skipping i == j; the loads go out
// in batches of 16 so that they overlap.
__global__ void
k_im_obs_sums (const double *__restrict__ AT, int nk, int nobs, double *__restrict__ off, double *__restrict__ diag)
{
  const int j = blockIdx.x * blockDim.x + threadIdx.x;

  if (j >= nobs)
    return;

  double s = 0.0;
  int i    = 0;

  for (; i + 16 <= nk; i += 16)
  {
    double v[16];

#pragma unroll
    for (int q = 0; q < 16; ++q)
      v[q] = AT[(size_t) (i + q) * nobs + j];

#pragma unroll
    for (int q = 0; q < 16; ++q)
      s += (i + q != j) ? v[q] : 0.0;
  }

  for (; i < nk; ++i)
    s += (i != j) ? AT[(size_t) i * nobs + j] : 0.0;

  off[j]  = s;
  diag[j] = (j < nk) ? AT[(size_t) j * nobs + j] : 0.0;
}

hipError_t
launch_im_obs_sums (const double *AT, int nk, int nobs, double *off, double *diag, hipStream_t s)
{
  if (nobs > 0)
    hipLaunchKernelGGL (k_im_obs_sums, dim3 ((nobs + 127) / 128), dim3 (128), 0, s, AT, nk, nobs, off, diag);

  return hipGetLastError ();
}

// out[a] = sum_j AT[a][j] v[j] (v == NULL: ones).  One wave per row.
__global__ void
k_row_dot (const double *__restrict__ AT, int n, int K, const double *__restrict__ v, double *__restrict__ out)
{
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / 64;
  const int lane = threadIdx.x % 64;

  if (wave >= n)
    return;

  const double *row = AT + (size_t) wave * K;
  double s          = 0.0;

  for (int j = lane; j < K; j += 64)
    s = fma (row[j], v ? v[j] : 1.0, s);

  for (int o = 32; o > 0; o >>= 1)
    s += __shfl_xor (s, o, 64);

  if (lane == 0)
    out[wave] = s;
}

hipError_t
launch_row_dot (const double *AT, int n, int K, const double *v, double *out, hipStream_t s)
{
  hipLaunchKernelGGL (k_row_dot, dim3 ((n + 3) / 4), dim3 (256), 0, s, AT, n, K, v, out);

  return hipGetLastError ();
}

// r = f - A x (f == NULL: ones) and |r|^2.  Pass 1: block (jb, c) sums the
// kernels of chunk c (res_chunk kernels, inactive ones skipped: x is uniform
// over the block) for 256 observations; pass 2 adds the chunks in a fixed
// order and forms per-block partial sums of r^2; a last pass adds those.
// Chunks: about 2048 blocks in all, at most 4096 chunks (n = 3200: 21
// kernels per chunk; 512-kernel chunks left 91 blocks whose threads each ran
// a 512-step serial chain, 188 us).
static int
res_chunk (int nk, int nobs)
{
  const int nb = (nobs + 255) / 256;
  int nch      = (2048 + nb - 1) / nb;

  nch = (nch < 1) ? 1 : ((nch > 4096) ? 4096 : nch);
  nch = (nch > nk) ? nk : nch;

  return (nk + nch - 1) / (nch > 0 ? nch : 1);
}

__global__ void __launch_bounds__ (256)
k_residual_part (const double *__restrict__ AT, int nk, int nobs, const double *__restrict__ x, int ch,
                 double *__restrict__ part)
{
  const int j  = blockIdx.x * 256 + threadIdx.x;
  const int a0 = blockIdx.y * ch;
  const int a1 = (a0 + ch < nk) ? a0 + ch : nk;
  double acc   = 0.0;

  if (j < nobs)
    for (int a = a0; a < a1; ++a)
    {
      const double xa = x[a];

      if (xa != 0.0)
        acc = fma (AT[(size_t) a * nobs + j], xa, acc);
    }

  if (j < nobs)
    part[(size_t) blockIdx.y * nobs + j] = acc;
}

__global__ void __launch_bounds__ (256)
k_residual (const double *__restrict__ part, int nch, int nobs, const double *__restrict__ f,
            double *__restrict__ r, double *__restrict__ partial)
{
  __shared__ double red[256];
  const int j = blockIdx.x * 256 + threadIdx.x;
  double rj   = 0.0;

  if (j < nobs)
  {
    double acc = 0.0;

    for (int c = 0; c < nch; ++c)
      acc += part[(size_t) c * nobs + j];

    rj   = (f ? f[j] : 1.0) - acc;
    r[j] = rj;
  }

  red[threadIdx.x] = rj * rj;
  __syncthreads ();

  for (int o = 128; o > 0; o >>= 1)
  {
    if (threadIdx.x < o)
      red[threadIdx.x] += red[threadIdx.x + o];

    __syncthreads ();
  }

  if (threadIdx.x == 0)
    partial[blockIdx.x] = red[0];
}

__global__ void
k_sum_partials (const double *__restrict__ partial, int np, double *__restrict__ out)
{
  if (threadIdx.x == 0 && blockIdx.x == 0)
  {
    double s = 0.0;

    for (int i = 0; i < np; ++i)
      s += partial[i];

    out[0] = s;
  }
}

size_t
residual_work_doubles (int nk, int nobs)
{
  const int ch = res_chunk (nk, nobs);

  return (size_t) ((nk + ch - 1) / ch) * nobs;
}

hipError_t
launch_residual (const double *AT, int nk, int nobs, const double *x, const double *f, double *r, double *partial,
                 double *rnorm2, double *work, hipStream_t s)
{
  const int nb  = (nobs + 255) / 256;
  const int ch  = res_chunk (nk, nobs);
  const int nch = (nk + ch - 1) / ch;

  hipLaunchKernelGGL (k_residual_part, dim3 (nb, nch), dim3 (256), 0, s, AT, nk, nobs, x, ch, work);
  hipLaunchKernelGGL (k_residual, dim3 (nb), dim3 (256), 0, s, work, nch, nobs, f, r, partial);
  hipLaunchKernelGGL (k_sum_partials, dim3 (1), dim3 (64), 0, s, partial, nb, rnorm2);

  return hipGetLastError ();
}

// S = M[P][P], bs = b[P]
__global__ void
k_gather_sym (const double *__restrict__ M, int n, const int *__restrict__ P, int np, double *__restrict__ S,
              const double *__restrict__ b, double *__restrict__ bs)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;

  if (r >= np || c >= np)
    return;

  S[(size_t) r * np + c] = M[(size_t) P[r] * n + P[c]];

  if (c == 0)
    bs[r] = b[P[r]];
}

hipError_t
launch_gather_sym (const double *M, int n, const int *P, int np, double *S, const double *b, double *bs,
                   hipStream_t s)
{
  hipLaunchKernelGGL (k_gather_sym, dim3 ((np + 255) / 256, np), dim3 (256), 0, s, M, n, P, np, S, b, bs);

  return hipGetLastError ();
}

// --------------------------------------------------------------------------
// Blocked right-looking Cholesky of the SPD system S = U^T U (upper, as
// dposv with the row-major 'U' of ncm_matrix.c:1198-1208), 64-wide panels:
//   diagonal block in LDS (dot-product form), row-panel triangular solve,
//   trailing rank-64 update of the upper triangle.
// info = first failing global pivot + 1 (LAPACK convention), else 0.
// --------------------------------------------------------------------------
constexpr int CH_NB = 64;

// Factor of the 64 x 64 diagonal block, right-looking with the block in
// registers: thread (tr, tc) of a 16 x 16 grid owns entries
// (tr + 16 a, tc + 16 b), a, b < 4.  Per column j the owner of (j, j)
// publishes the pivot, the owners of row j publish the scaled row, and every
// thread updates its own entries from two LDS vectors (two barriers per
// column).  Returns the failing pivot + 1 (LAPACK info), else 0; every
// thread of the block returns the same value.
__device__ inline int
factor_diag_block (const double *__restrict__ S, int np, int kb, int nb, double (&x)[4][4], double *row,
                   double &piv)
{
  const int tid = threadIdx.x;
  const int tr  = tid >> 4, tc = tid & 15;

#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
    {
      const int r = tr + 16 * a, c = tc + 16 * b;

      x[a][b] = (r < nb && c < nb) ? S[(size_t) (kb + r) * np + kb + c] : 0.0;
    }

  for (int j = 0; j < nb; ++j)
  {
    const int ja = j >> 4, jr = j & 15;

    if (tr == jr && tc == jr)
    {
#pragma unroll
      for (int a = 0; a < 4; ++a)
        if (a == ja)
          piv = x[a][a];
    }

    __syncthreads ();

    const double djj = piv;

    if (!(djj > 0.0))
      return kb + j + 1;

    const double ujj = sqrt (djj);

    if (tr == jr)
    {
#pragma unroll
      for (int a = 0; a < 4; ++a)
        if (a == ja)
        {
#pragma unroll
          for (int b = 0; b < 4; ++b)
          {
            const int c = tc + 16 * b;

            if (c >= j && c < nb)
            {
              const double u = (c == j) ? ujj : x[a][b] / ujj;

              x[a][b] = u;
              row[c]  = u;
            }
          }
        }
    }

    __syncthreads ();

    double ur[4], uc[4];

#pragma unroll
    for (int a = 0; a < 4; ++a)
    {
      const int r = tr + 16 * a, c = tc + 16 * a;

      ur[a] = (r > j && r < nb) ? row[r] : 0.0;
      uc[a] = (c > j && c < nb) ? row[c] : 0.0;
    }

#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b)
        x[a][b] = fma (-ur[a], uc[b], x[a][b]);
  }

  return 0;
}

// Panel step for a panel narrower than 64 (the last one when np is not a
// multiple of 64; full panels go through k_potrf_rowpanel): every block
// factors the diagonal block (redundantly, from
// the unfactored values, which stay in S until k_copy_diag), then solves
// U_kk^T x = S[kb:kb+nb][c] for its 64 columns c >= kb+nb in LDS (wave 0, one
// lane per column, four partial sums per row) and writes them back.  Block
// 0 stores the factor in diagbuf.  Staging loads of the panel columns are
// issued before the factorisation.
__global__ void __launch_bounds__ (256)
k_potrf_panel (double *__restrict__ S, int np, int kb, int nb, int *__restrict__ info, double *__restrict__ diagbuf)
{
  __shared__ double U[CH_NB][CH_NB + 1];
  __shared__ double Xs[CH_NB][64 + 1];
  __shared__ double row[CH_NB];
  __shared__ double piv;
  const int tid  = threadIdx.x;
  const int tr   = tid >> 4, tc = tid & 15;
  const int lane = tid & 63, w = tid >> 6;
  const int c0   = kb + nb + blockIdx.x * 64;
  const int c    = c0 + lane;

  if (*info != 0)
    return;

  double vx[16];

#pragma unroll
  for (int q = 0; q < 16; ++q)
  {
    const int r = w * 16 + q;

    vx[q] = (r < nb && c < np) ? S[(size_t) (kb + r) * np + c] : 0.0;
  }

  double x[4][4];
  const int fail = factor_diag_block (S, np, kb, nb, x, row, piv);

  if (fail)
  {
    if (blockIdx.x == 0 && tid == 0)
      *info = fail;

    return;
  }

#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
    {
      const int r = tr + 16 * a, cc = tc + 16 * b;

      if (r < nb && cc < nb)
      {
        U[r][cc] = x[a][b];

        if (blockIdx.x == 0 && cc >= r)
          diagbuf[r * CH_NB + cc] = x[a][b];
      }
    }

#pragma unroll
  for (int q = 0; q < 16; ++q)
    Xs[w * 16 + q][lane] = vx[q];

  __syncthreads ();

  if (w == 0 && c0 < np)
  {
    for (int r = 0; r < nb; ++r)
    {
      double t0 = Xs[r][lane], t1 = 0.0, t2 = 0.0, t3 = 0.0;
      int q     = 0;

      for (; q + 4 <= r; q += 4)
      {
        t0 = fma (-U[q][r], Xs[q][lane], t0);
        t1 = fma (-U[q + 1][r], Xs[q + 1][lane], t1);
        t2 = fma (-U[q + 2][r], Xs[q + 2][lane], t2);
        t3 = fma (-U[q + 3][r], Xs[q + 3][lane], t3);
      }

      for (; q < r; ++q)
        t0 = fma (-U[q][r], Xs[q][lane], t0);

      Xs[r][lane] = ((t0 + t1) + (t2 + t3)) / U[r][r];
    }
  }

  __syncthreads ();

  if (c < np)
#pragma unroll
    for (int q = 0; q < 16; ++q)
    {
      const int r = w * 16 + q;

      if (r < nb)
        S[(size_t) (kb + r) * np + c] = Xs[r][lane];
    }
}

// double broadcast from lane `src` (a compile-time constant after unrolling)
__device__ inline double
bcast_lane (double v, int src)
{
  const int lo = __builtin_amdgcn_readlane (__double2loint (v), src);
  const int hi = __builtin_amdgcn_readlane (__double2hiint (v), src);

  return __hiloint2double (hi, lo);
}

// A full panel (nb = 64) is one launch: the diagonal block's factor and the
// solve of the row panel right of it are one right-looking factorisation of
// the 64 x (64 + cols) row panel, each column held whole by one lane (its
// 64 entries in registers).  Block b has two waves: wave 0 holds the
// diagonal block (every block factors it, from the same values with the
// same arithmetic), wave 1 holds RP_EXTRA = 64 columns of the row panel.
// Step j: wave 0 takes the pivot from lane j by readlane, forms u_jj =
// sqrt(pivot) and r_j = 1 / u_jj, scales row j of the diagonal block by
// r_j and publishes it with r_j in LDS (double-buffered, one barrier per
// step); every lane of the row panel scales its own row-j entry by r_j; then
// each lane applies the rank-one update a_i -= u_ji * u_jc to its rows i >
// j, row j's entries read from LDS all at once.  Entries below the diagonal
// take part in the diagonal block's updates unmasked: they only ever
// combine with each other and are never stored.  A non-positive pivot is
// recorded (the first one), the steps run on, and nothing is stored.  Two
// waves per block measured fastest (2.0 ms for a 3200-column factor, against
// 2.3 ms with four waves and 2.8 ms for a one-wave diagonal factor followed
// by a separate panel solve): the step barrier is cheap between two waves.
// The factor goes to Ud (upper triangle, row-major 64 x 64) and the
// reciprocals of its diagonal to rd.
constexpr int RP_EXTRA = 64;

__global__ void __launch_bounds__ (64 + RP_EXTRA)
k_potrf_rowpanel (double *__restrict__ S, int np, int kb, int *__restrict__ info, double *__restrict__ Ud,
                  double *__restrict__ rd)
{
  __shared__ double R[2][CH_NB];
  __shared__ double Rr[2];
  __shared__ int s_fail;
  const int tid  = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  // wave 0: diagonal block column `lane`; wave 1: row-panel columns
  const int c    = (w == 0) ? kb + lane : kb + CH_NB + blockIdx.x * RP_EXTRA + (w - 1) * 64 + lane;
  const bool act = c < np;
  const int cc   = act ? c : kb;

  if (*info != 0)
    return;

  double a[CH_NB];
  double rdl = 0.0;   // wave 0, lane j: 1 / u_jj
  int fail   = 0;

#pragma unroll
  for (int r = 0; r < CH_NB; ++r)
    a[r] = (w > 0 || r <= lane) ? S[(size_t) (kb + r) * np + cc] : 0.0;

#pragma unroll
  for (int j = 0; j < CH_NB; ++j)
  {
    const int bsel = j & 1;

    if (w == 0)
    {
      const double dn = bcast_lane (a[j], j);

      fail = (fail == 0 && !(dn > 0.0)) ? j + 1 : fail;

      const double un = sqrt (dn);
      const double rn = 1.0 / un;

      a[j]          = (lane == j) ? un : a[j] * rn;
      R[bsel][lane] = a[j];
      rdl           = (lane == j) ? rn : rdl;

      if (lane == j)
        Rr[bsel] = rn;
    }

    __syncthreads ();

    const double rn = Rr[bsel];

    if (w > 0)
      a[j] = a[j] * rn;

    double rr[CH_NB];

#pragma unroll
    for (int r = j + 1; r < CH_NB; ++r)
      rr[r] = R[bsel][r];

    const double uj = a[j];

#pragma unroll
    for (int r = j + 1; r < CH_NB; ++r)
      a[r] = fma (-rr[r], uj, a[r]);
  }

  if (tid == 0)
    s_fail = fail;

  __syncthreads ();

  if (s_fail != 0)
  {
    if (blockIdx.x == 0 && tid == 0)
      *info = kb + s_fail;

    return;
  }

  if (w == 0)
  {
    if (blockIdx.x == 0)
    {
#pragma unroll
      for (int r = 0; r < CH_NB; ++r)
        if (r <= lane)
          Ud[r * CH_NB + lane] = a[r];

      rd[lane] = rdl;
    }
  }
  else if (act)
  {
#pragma unroll
    for (int r = 0; r < CH_NB; ++r)
      S[(size_t) (kb + r) * np + c] = a[r];
  }
}

// the factored diagonal blocks (upper triangles) into S
__global__ void __launch_bounds__ (256)
k_copy_diag (double *__restrict__ S, int np, const double *__restrict__ diagbuf)
{
  const int kb = blockIdx.x * CH_NB;
  const int nb = (np - kb < CH_NB) ? np - kb : CH_NB;
  const double *D = diagbuf + (size_t) blockIdx.x * CH_NB * CH_NB;

  for (int e = threadIdx.x; e < CH_NB * CH_NB; e += 256)
  {
    const int r = e >> 6, c = e & 63;

    if (r < nb && c < nb && c >= r)
      S[(size_t) (kb + r) * np + kb + c] = D[e];
  }
}

// lower triangle <- transpose of the upper factor (for the backward solve)
__global__ void
k_mirror_lower (double *__restrict__ S, int np)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;

  if (r < np && c < r)
    S[(size_t) r * np + c] = S[(size_t) c * np + r];
}

constexpr int SP_NB  = 4 * CH_NB;   // super-panel: four panels
constexpr int SP_MIN = 4096;        // columns left above which super-panels are used

// S[off:][off:] -= (rank-K product of the rows kb .. kb+K), upper triangle:
// 128-wide tiles while they number at least 16384 (trailing size above
// about 23000: more reuse per staged operand), 64-wide ones below (measured
// faster at n = 20000 and below)
static void
trailing_update (double *S, int np, int K, int off, int kb, hipStream_t s)
{
  const int nt128 = (np - off + 127) / 128;

  if (nt128 * (nt128 + 1) / 2 >= 16384)
    launch_syrk_mfma<1, 128> (S, np, K, np, off, kb, S, s);
  else
    launch_syrk_mfma<1, 64> (S, np, K, np, off, kb, S, s);
}

// Right-looking blocked Cholesky S = U^T U in 64-row panels: per panel, the
// row-panel factorisation (diagonal factor and row-panel solve in one
// launch), then the FP64-MFMA trailing update of the upper triangle; in a
// super-panel, each panel's update is applied to the super-panel's own rows
// only, and the rest of the trailing matrix gets one rank-256 update.  Factored diagonal blocks wait in
// diagbuf until the end; *info is the 1-based failing pivot, 0 on success.
hipError_t
launch_potrf (double *S, int np, int *info, double *diagbuf, hipStream_t s)
{
  const int npan = (np + CH_NB - 1) / CH_NB;
  double *rdall  = diagbuf + (size_t) npan * CH_NB * CH_NB;

  for (int kb = 0; kb < np;)
  {
    // while more than SP_MIN columns remain: super-panels of four full
    // panels, whose trailing update is one rank-256 product instead of four
    // rank-64 ones (a 4x shorter stream over the trailing matrix: 12000
    // columns factor in 18.7 ms against 25.3 ms); below that the per-panel
    // steps are as fast or faster
    if (np - kb > SP_MIN)
    {
      const int ke = kb + SP_NB;

      for (int kp = kb; kp < ke; kp += CH_NB)
      {
        const int rest = np - kp - CH_NB;
        const int gr   = (rest + RP_EXTRA - 1) / RP_EXTRA;

        hipLaunchKernelGGL (k_potrf_rowpanel, dim3 (gr), dim3 (64 + RP_EXTRA), 0, s, S, np, kp, info,
                            diagbuf + (size_t) (kp / CH_NB) * CH_NB * CH_NB, rdall + kp);

        // the rank-64 update of the super-panel's own rows below this panel
        if (kp + CH_NB < ke)
          launch_syrk_mfma<1, 64> (S, np, CH_NB, np, kp + CH_NB, kp, S, s, ke);
      }

      trailing_update (S, np, SP_NB, ke, kb, s);
      kb = ke;
      continue;
    }

    const int nb   = (np - kb < CH_NB) ? np - kb : CH_NB;
    const int rest = np - kb - nb;
    double *db     = diagbuf + (size_t) (kb / CH_NB) * CH_NB * CH_NB;

    if (nb == CH_NB)
    {
      const int gr = (rest + RP_EXTRA - 1) / RP_EXTRA;

      hipLaunchKernelGGL (k_potrf_rowpanel, dim3 (gr > 0 ? gr : 1), dim3 (64 + RP_EXTRA), 0, s, S, np, kb, info, db,
                          rdall + kb);
    }
    else
    {
      hipLaunchKernelGGL (k_potrf_panel, dim3 ((rest > 0) ? (rest + 63) / 64 : 1), dim3 (256), 0, s, S, np, kb, nb,
                          info, db);
    }

    if (rest > 0)
      trailing_update (S, np, nb, kb + nb, kb, s);

    kb += nb;
  }

  hipLaunchKernelGGL (k_copy_diag, dim3 ((np + CH_NB - 1) / CH_NB), dim3 (256), 0, s, S, np, diagbuf);
  hipLaunchKernelGGL (k_mirror_lower, dim3 ((np + 255) / 256, np), dim3 (256), 0, s, S, np);

  return hipGetLastError ();
}

size_t
potrf_diag_doubles (int np)
{
  // the factored diagonal blocks, then the reciprocals of their diagonals
  return (size_t) ((np + CH_NB - 1) / CH_NB) * CH_NB * CH_NB + (size_t) np + CH_NB;
}

// U^T y = b then U x = y with the factor of launch_potrf (upper triangle =
// U, lower triangle = U^T), in 64-row panels: one launch per panel and
// direction.  With T[i][j] = S[(p0+j) np + p0+i], the panel's diagonal
// block is L = U^T below its diagonal (forward) and U above it (backward),
// and the coupling of a row i outside the panel to panel column c is
// S[(p0+c) np + i] in both directions, a unit-stride read over i.  Every
// block stages T in LDS and solves the panel with one wave (64 serial
// steps, operands shuffled from the pivot lane); block 0 writes the panel's
// solution to `out`, and every block updates its 256 rows of the right-hand
// side outside the panel.  The panel rows of `rhs` are only read in a
// launch, so the blocks never race.
constexpr int TS_NB = 64;

template <int FWD>
__global__ void __launch_bounds__ (256)
k_trsv_panel (const double *__restrict__ S, int np, int p0, int pb, double *__restrict__ rhs,
              double *__restrict__ out)
{
  __shared__ double T[TS_NB][TS_NB + 1];
  __shared__ double y[TS_NB];
  __shared__ double rdiag[TS_NB];
  const int tid = threadIdx.x;

  {
    double v[16];

#pragma unroll
    for (int q = 0; q < 16; ++q)
    {
      const int e = tid + q * 256, j = e >> 6, i = e & 63;

      v[q] = (i < pb && j < pb) ? S[(size_t) (p0 + j) * np + p0 + i] : 0.0;
    }

#pragma unroll
    for (int q = 0; q < 16; ++q)
    {
      const int e = tid + q * 256;

      T[e & 63][e >> 6] = v[q];
    }
  }

  __syncthreads ();

  if (tid < 64)
  {
    const int i = tid;
    double bi   = (i < pb) ? rhs[p0 + i] : 0.0;

    // reciprocals of the diagonal, then 64 serial steps with the pivot
    // taken by readlane (j is wave-uniform) and a multiplication in place
    // of the division
    rdiag[i] = (i < pb) ? 1.0 / T[i][i] : 0.0;
    __builtin_amdgcn_fence (__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier ();
    __builtin_amdgcn_fence (__ATOMIC_ACQUIRE, "wavefront");

    if (pb == TS_NB)
    {
      // full panel: the lane's row of T in registers, zero outside the
      // triangle being solved, so each step is a readlane of lane j's
      // scaled value and one fma or select (the same arithmetic as below)
      const double rdl = rdiag[i];
      double t[TS_NB];

#pragma unroll
      for (int j = 0; j < TS_NB; ++j)
        t[j] = (FWD ? (j < i) : (j > i)) ? T[i][j] : 0.0;

#pragma unroll
      for (int step = 0; step < TS_NB; ++step)
      {
        const int j     = FWD ? step : TS_NB - 1 - step;
        const double yj = bcast_lane (bi * rdl, j);

        bi = (i == j) ? yj : fma (-t[j], yj, bi);
      }
    }
    else
    {
      for (int step = 0; step < pb; ++step)
      {
        const int j     = FWD ? step : pb - 1 - step;
        const double bj = bcast_lane (bi, j);
        const double yj = bj * rdiag[j];

        if (i == j)
          bi = yj;
        else if (FWD ? (i > j) : (i < j))
          bi = fma (-T[i][j], yj, bi);
      }
    }

    y[i] = bi;

    if (blockIdx.x == 0 && i < pb)
      out[p0 + i] = bi;
  }

  __syncthreads ();

  const int row = (FWD ? p0 + pb : 0) + blockIdx.x * 256 + tid;
  const int hi  = FWD ? np : p0;

  if (row < hi)
  {
    double t[4] = {rhs[row], 0.0, 0.0, 0.0};

    for (int c0 = 0; c0 < pb; c0 += 16)
    {
      double v[16];

#pragma unroll
      for (int q = 0; q < 16; ++q)
        v[q] = (c0 + q < pb) ? S[(size_t) (p0 + c0 + q) * np + row] : 0.0;

#pragma unroll
      for (int q = 0; q < 16; ++q)
        t[q & 3] = fma (-v[q], (c0 + q < pb) ? y[c0 + q] : 0.0, t[q & 3]);
    }

    rhs[row] = (t[0] + t[1]) + (t[2] + t[3]);
  }
}

// solution of (U^T U) x = b returned in b; tmp holds np doubles
hipError_t
launch_trsv_pair (const double *S, int np, double *b, double *tmp, hipStream_t s)
{
  for (int p0 = 0; p0 < np; p0 += TS_NB)
  {
    const int pb   = (np - p0 < TS_NB) ? np - p0 : TS_NB;
    const int rest = np - p0 - pb;
    const int grid = (rest + 255) / 256 > 0 ? (rest + 255) / 256 : 1;

    hipLaunchKernelGGL (k_trsv_panel<1>, dim3 (grid), dim3 (256), 0, s, S, np, p0, pb, b, tmp);
  }

  const int last = ((np - 1) / TS_NB) * TS_NB;

  for (int p0 = last; p0 >= 0; p0 -= TS_NB)
  {
    const int pb   = (np - p0 < TS_NB) ? np - p0 : TS_NB;
    const int grid = (p0 + 255) / 256 > 0 ? (p0 + 255) / 256 : 1;

    hipLaunchKernelGGL (k_trsv_panel<0>, dim3 (grid), dim3 (256), 0, s, S, np, p0, pb, tmp, b);
  }

  return hipGetLastError ();
}

// U x = y alone, U in the upper triangle and U^T mirrored into the lower one
// (the factor of launch_lu_factor after launch_mirror_lower); y is consumed
hipError_t
launch_trsv_backward (const double *S, int np, double *y, double *x, hipStream_t s)
{
  const int last = ((np - 1) / TS_NB) * TS_NB;

  for (int p0 = last; p0 >= 0; p0 -= TS_NB)
  {
    const int pb   = (np - p0 < TS_NB) ? np - p0 : TS_NB;
    const int grid = (p0 + 255) / 256 > 0 ? (p0 + 255) / 256 : 1;

    hipLaunchKernelGGL (k_trsv_panel<0>, dim3 (grid), dim3 (256), 0, s, S, np, p0, pb, y, x);
  }

  return hipGetLastError ();
}

// The same two sweeps as one launch each, chained by flags: one 64-lane
// block per panel, block k taking the k-th panel in sweep order.  A block
// applies the contributions of the panels before it in sweep order, each
// as soon as that panel's flag carries this call's epoch, with the order
// and arithmetic of k_trsv_panel (the solution is the same to the bit),
// then solves its diagonal block and publishes its part of the solution and
// its flag.  The factor entries of a contribution are loaded before its
// flag is awaited.  The host launches the chain only when every block of
// the grid can be resident at once, so each wait ends; a wait longer than
// spin_limit polls still ends the kernel, with *status = epoch (no later
// block of either sweep does anything then), and the host solves again with
// the panel launches.
template <int FWD>
__global__ void __launch_bounds__ (64)
k_trsv_chain (const double *__restrict__ S, int np, const double *__restrict__ rhs, double *__restrict__ out,
              unsigned *__restrict__ flags, unsigned epoch, unsigned spin_limit, unsigned *__restrict__ status)
{
  __shared__ double T[TS_NB][TS_NB + 1];
  __shared__ double rdiag[TS_NB];
  __shared__ double ys[TS_NB];
  const int lane = threadIdx.x;
  const int npan = (np + TS_NB - 1) / TS_NB;
  const int k    = blockIdx.x;
  const int pan  = FWD ? k : npan - 1 - k;
  const int p0   = pan * TS_NB;
  const int pb   = (np - p0 < TS_NB) ? np - p0 : TS_NB;

  if (__hip_atomic_load (status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch)
    return;

  for (int e = lane; e < TS_NB * TS_NB; e += 64)
  {
    const int j = e >> 6, i = e & 63;

    T[i][j] = (i < pb && j < pb) ? S[(size_t) (p0 + j) * np + p0 + i] : 0.0;
  }

  const int row = p0 + lane;
  double r      = (lane < pb) ? rhs[row] : 0.0;

  for (int kk = 0; kk < k; ++kk)
  {
    const int q  = FWD ? kk : npan - 1 - kk;
    const int q0 = q * TS_NB;
    const int qb = (np - q0 < TS_NB) ? np - q0 : TS_NB;
    double v[TS_NB];

#pragma unroll
    for (int c = 0; c < TS_NB; ++c)
      v[c] = (c < qb && lane < pb) ? S[(size_t) (q0 + c) * np + row] : 0.0;

    unsigned polls = 0;
    bool dead      = false;

    while (__hip_atomic_load (&flags[q], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != epoch)
    {
      if (++polls > spin_limit || __hip_atomic_load (status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch)
      {
        dead = true;
        break;
      }

      __builtin_amdgcn_s_sleep (1);
    }

    if (dead)
    {
      if (lane == 0)
        __hip_atomic_store (status, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);

      return;
    }

    ys[lane] = (lane < qb) ? out[q0 + lane] : 0.0;
    __builtin_amdgcn_fence (__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier ();
    __builtin_amdgcn_fence (__ATOMIC_ACQUIRE, "wavefront");

    double t[4] = {r, 0.0, 0.0, 0.0};

#pragma unroll
    for (int c0 = 0; c0 < TS_NB; c0 += 16)
#pragma unroll
      for (int qq = 0; qq < 16; ++qq)
        if (c0 < qb)
          t[qq & 3] = fma (-v[c0 + qq], (c0 + qq < qb) ? ys[c0 + qq] : 0.0, t[qq & 3]);

    r = (t[0] + t[1]) + (t[2] + t[3]);
    __builtin_amdgcn_wave_barrier ();
  }

  __syncthreads ();

  const int i = lane;
  double bi   = r;

  rdiag[i] = (i < pb) ? 1.0 / T[i][i] : 0.0;
  __builtin_amdgcn_fence (__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier ();
  __builtin_amdgcn_fence (__ATOMIC_ACQUIRE, "wavefront");

  if (pb == TS_NB)
  {
    const double rdl = rdiag[i];
    double t[TS_NB];

#pragma unroll
    for (int j = 0; j < TS_NB; ++j)
      t[j] = (FWD ? (j < i) : (j > i)) ? T[i][j] : 0.0;

#pragma unroll
    for (int step = 0; step < TS_NB; ++step)
    {
      const int j     = FWD ? step : TS_NB - 1 - step;
      const double yj = bcast_lane (bi * rdl, j);

      bi = (i == j) ? yj : fma (-t[j], yj, bi);
    }
  }
  else
  {
    for (int step = 0; step < pb; ++step)
    {
      const int j     = FWD ? step : pb - 1 - step;
      const double bj = bcast_lane (bi, j);
      const double yj = bj * rdiag[j];

      if (i == j)
        bi = yj;
      else if (FWD ? (i > j) : (i < j))
        bi = fma (-T[i][j], yj, bi);
    }
  }

  if (i < pb)
    out[p0 + i] = bi;

  __hip_atomic_store (&flags[pan], epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// panels the chained sweeps take at most: the blocks of k_trsv_chain the
// device holds at once
int
trsv_chain_max_panels ()
{
  static int cap = -1;

  if (cap < 0)
  {
    int dev = 0, ncu = 0, nb = 0;

    (void) hipGetDevice (&dev);
    (void) hipDeviceGetAttribute (&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    (void) hipOccupancyMaxActiveBlocksPerMultiprocessor (&nb, k_trsv_chain<1>, 64, 0);

    int nb0 = 0;

    (void) hipOccupancyMaxActiveBlocksPerMultiprocessor (&nb0, k_trsv_chain<0>, 64, 0);
    nb  = (nb0 < nb) ? nb0 : nb;
    cap = (nb > 0 && ncu > 0) ? nb * ncu : 0;
  }

  return cap;
}

// flags: status word, then 2 npan panel flags (forward, backward)
hipError_t
launch_trsv_chain (const double *S, int np, double *b, double *tmp, unsigned *flags, unsigned epoch, hipStream_t s)
{
  const int npan = (np + TS_NB - 1) / TS_NB;

  hipLaunchKernelGGL (k_trsv_chain<1>, dim3 (npan), dim3 (64), 0, s, S, np, b, tmp, flags + 1, epoch, 1u << 22,
                      flags);
  hipLaunchKernelGGL (k_trsv_chain<0>, dim3 (npan), dim3 (64), 0, s, S, np, tmp, b, flags + 1 + npan, epoch,
                      1u << 22, flags);

  return hipGetLastError ();
}

hipError_t
launch_mirror_lower (double *S, int np, hipStream_t s)
{
  hipLaunchKernelGGL (k_mirror_lower, dim3 ((np + 255) / 256, np), dim3 (256), 0, s, S, np);

  return hipGetLastError ();
}

// --------------------------------------------------------------------------
// Fallback when the normal matrix is not numerically SPD (the reference's
// dsysv path, ncm_nnls.c:574-605): Gaussian elimination with partial
// pivoting (first maximum) on the symmetrised matrix, one 1024-thread block.
// --------------------------------------------------------------------------
__global__ void __launch_bounds__ (1024)
k_lu_solve (double *__restrict__ S, int np, double *__restrict__ b, int *__restrict__ info)
{
  __shared__ double rv[1024];
  __shared__ int ri[1024];
  const int tid = threadIdx.x;

  for (size_t e = tid; e < (size_t) np * np; e += 1024)
  {
    const int i = (int) (e / np), j = (int) (e % np);

    if (j < i)
      S[e] = S[(size_t) j * np + i];
  }

  __syncthreads ();

  for (int k = 0; k < np; ++k)
  {
    double best = -1.0;
    int bi      = np;

    for (int i = k + tid; i < np; i += 1024)
    {
      const double v = fabs (S[(size_t) i * np + k]);

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

    __syncthreads ();

    if (amax == 0.0)
    {
      if (tid == 0)
        *info = k + 1;

      return;
    }

    if (piv != k)
    {
      for (int j = tid; j < np; j += 1024)
      {
        const double t          = S[(size_t) k * np + j];
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

    __syncthreads ();

    const int rows = np - k - 1;

    for (size_t e = tid; e < (size_t) rows * (rows + 1); e += 1024)
    {
      const int i    = k + 1 + (int) (e / (rows + 1));
      const int jj   = (int) (e % (rows + 1));
      const double l = S[(size_t) i * np + k] / S[(size_t) k * np + k];

      if (jj < rows)
      {
        const int j = k + 1 + jj;

        S[(size_t) i * np + j] = S[(size_t) i * np + j] - l * S[(size_t) k * np + j];
      }
      else
      {
        b[i] = b[i] - l * b[k];
      }
    }

    __syncthreads ();
  }

  for (int i = np - 1; i >= 0; --i)
  {
    if (tid == 0)
    {
      double t = b[i];

      for (int j = i + 1; j < np; ++j)
        t = t - S[(size_t) i * np + j] * b[j];

      b[i] = t / S[(size_t) i * np + i];
    }

    __syncthreads ();
  }

  if (tid == 0)
    *info = 0;
}

hipError_t
launch_lu_solve (double *S, int np, double *b, int *info, hipStream_t s)
{
  hipLaunchKernelGGL (k_lu_solve, dim3 (1), dim3 (1024), 0, s, S, np, b, info);

  return hipGetLastError ();
}

__global__ void
k_scatter (const double *__restrict__ sol, const int *__restrict__ P, int np, int n, double *__restrict__ x)
{
  const int a = blockIdx.x * blockDim.x + threadIdx.x;

  if (a < n)
    x[a] = 0.0;
}

__global__ void
k_scatter2 (const double *__restrict__ sol, const int *__restrict__ P, int np, double *__restrict__ x)
{
  const int a = blockIdx.x * blockDim.x + threadIdx.x;

  if (a < np)
    x[P[a]] = sol[a];
}

hipError_t
launch_scatter (const double *sol, const int *P, int np, int n, double *x, hipStream_t s)
{
  hipLaunchKernelGGL (k_scatter, dim3 ((n + 255) / 256), dim3 (256), 0, s, sol, P, np, n, x);
  hipLaunchKernelGGL (k_scatter2, dim3 ((np + 255) / 256), dim3 (256), 0, s, sol, P, np, x);

  return hipGetLastError ();
}

// w <- w (1 - s) / sum(w) + s / n  (ncm_stats_dist.c:1087-1093), serial sum
__global__ void
k_shrink (double *__restrict__ w, int n, double shrink, int *__restrict__ status)
{
  __shared__ double total;
  __shared__ double wl[4096];
  double t = 0.0;

  // serial sum in index order (thread 0) over chunks staged in LDS
  for (int i0 = 0; i0 < n; i0 += 4096)
  {
    const int cnt = (n - i0 < 4096) ? n - i0 : 4096;

    for (int i = threadIdx.x; i < cnt; i += blockDim.x)
      wl[i] = w[i0 + i];

    __syncthreads ();

    if (threadIdx.x == 0)
      for (int i = 0; i < cnt; ++i)
        t = __dadd_rn (t, wl[i]);

    __syncthreads ();
  }

  if (threadIdx.x == 0)
  {
    total = t;

    if (!(t > 0.0))
      status[0] = 3;
  }

  __syncthreads ();

  const double sc = (1.0 - shrink) / total;
  const double ad = shrink / n;

  for (int i = threadIdx.x; i < n; i += blockDim.x)
    w[i] = __dadd_rn (__dmul_rn (w[i], sc), ad);
}

hipError_t
launch_shrink (double *w, int n, double shrink, int *status, hipStream_t s)
{
  hipLaunchKernelGGL (k_shrink, dim3 (1), dim3 (1024), 0, s, w, n, shrink, status);

  return hipGetLastError ();
}

} // namespace nck
