// k_sample.hip -- mixture sampling: kernel choice and draw transform.
//
// The cumulative weights follow ncm_stats_dist_kernel_choose
// (ncm_stats_dist.c:1565-1606) exactly: one lane adds the weights in index
// order, then every entry is multiplied by 1/cum.  The bisection on each
// host-supplied uniform is the reference's, so the chosen kernel indices are
// bit-identical.  The draw is x = theta_i + s U_i^T (h z) with s = 1 (Gauss,
// ncm_stats_dist_kernel_gauss.c:335-355) or sqrt(nu/chi2) (Student-t,
// ncm_stats_dist_kernel_st.c:388-414), in the serial dtrmv order.
#include "nckde_internal.h"

#pragma clang fp contract(off)

namespace nck
{

__global__ void
k_wcum (const double *__restrict__ w, int n, double *__restrict__ wcum)
{
  // the serial prefix sum of the reference, one thread, over chunks staged
  // in LDS by the whole block (the loads are off the dependency chain)
  __shared__ double wl[4096];
  __shared__ double s_scale;
  double cum = 0.0;

  if (threadIdx.x == 0)
    wcum[0] = 0.0;

  for (int i0 = 0; i0 < n; i0 += 4096)
  {
    const int cnt = (n - i0 < 4096) ? n - i0 : 4096;

    for (int i = threadIdx.x; i < cnt; i += blockDim.x)
      wl[i] = w[i0 + i];

    __syncthreads ();

    if (threadIdx.x == 0)
      for (int i = 0; i < cnt; i++)
      {
        cum   += wl[i];
        wl[i] = cum;
      }

    __syncthreads ();

    for (int i = threadIdx.x; i < cnt; i += blockDim.x)
      wcum[i0 + i + 1] = wl[i];

    __syncthreads ();
  }

  if (threadIdx.x == 0)
    s_scale = 1.0 / cum;

  __syncthreads ();

  const double s = s_scale;

  for (int i = threadIdx.x; i <= n; i += blockDim.x)
    wcum[i] = wcum[i] * s;
}

hipError_t
launch_wcum (const double *w, int n, double *wcum, hipStream_t s)
{
  hipLaunchKernelGGL (k_wcum, dim3 (1), dim3 (1024), 0, s, w, n, wcum);

  return hipGetLastError ();
}

__device__ inline int
choose (const double *wcum, int n, double p)
{
  int ilo = 0, ihi = n;

  while (ihi > ilo + 1)
  {
    const int mi = (ihi + ilo) / 2;

    if (wcum[mi] > p)
      ihi = mi;
    else
      ilo = mi;
  }

  return ilo;
}

__global__ void
k_choose (const double *__restrict__ wcum, int n, const double *__restrict__ u, int m, uint32_t *__restrict__ idx)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;

  if (p < m)
    idx[p] = (uint32_t) choose (wcum, n, u[p]);
}

hipError_t
launch_choose (const double *wcum, int n, const double *u, int m, uint32_t *idx, hipStream_t s)
{
  hipLaunchKernelGGL (k_choose, dim3 ((m + 255) / 256), dim3 (256), 0, s, wcum, n, u, m, idx);

  return hipGetLastError ();
}

__global__ void
k_sample (const double *__restrict__ wcum, int n, const double *__restrict__ u, const double *__restrict__ z,
          const double *__restrict__ chi2, int m, int d, int kind, int kernel, double nu, double href,
          const double *__restrict__ Uall, const double *__restrict__ X, double *__restrict__ out,
          uint32_t *__restrict__ idx)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;

  if (p >= m)
    return;

  const int i      = choose (wcum, n, u[p]);
  const double *U  = (kind == NCKDE_VKDE) ? Uall + (size_t) i * d * d : Uall;
  const double *mu = X + (size_t) i * d;
  const double *zp = z + (size_t) p * d;
  double *x        = out + (size_t) p * d;

  for (int l = 0; l < d; l++)
    x[l] = zp[l] * href;

  for (int l = d - 1; l >= 0; l--)
  {
    double t = U[l * d + l] * x[l];

    for (int q = 0; q < l; q++)
      t = t + U[q * d + l] * x[q];

    x[l] = t;
  }

  if (kernel == NCKDE_KERNEL_ST)
  {
    const double chi_scale = sqrt (nu / chi2[p]);

    for (int l = 0; l < d; l++)
      x[l] = x[l] * chi_scale;
  }

  for (int l = 0; l < d; l++)
    x[l] = x[l] + mu[l];

  if (idx)
    idx[p] = (uint32_t) i;
}

hipError_t
launch_sample (const double *wcum, int n, const double *u, const double *z, const double *chi2, int m, int d,
               int kind, int kernel, double nu, double href, const double *U, const double *X, double *out,
               uint32_t *idx, hipStream_t s)
{
  hipLaunchKernelGGL (k_sample, dim3 ((m + 127) / 128), dim3 (128), 0, s, wcum, n, u, z, chi2, m, d, kind, kernel,
                      nu, href, U, X, out, idx);

  return hipGetLastError ();
}

} // namespace nck
