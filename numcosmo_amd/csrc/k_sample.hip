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
          uint32_t *__restrict__ idx, const uint32_t *__restrict__ idx_in)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;

  if (p >= m)
    return;

  const int i      = idx_in ? (int) idx_in[p] : choose (wcum, n, u[p]);
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

// --------------------------------------------------------------------------
// Counter-based draws for a batch of m samples (Philox4x32-10 of Salmon,
// Moraes, Dror & Shaw, SC'11 "Parallel random numbers: as easy as 1, 2, 3").
// Sample j uses counters (j, q, c0, c1) for q = 0, 1, ...: one 53-bit uniform
// u, d standard normals z (Box-Muller on pairs of uniforms) and, for the
// Student-t kernel with integer nu, chi2 = sum of nu squared normals.
// --------------------------------------------------------------------------
struct u32x4
{
  uint32_t x, y, z, w;
};

__device__ inline u32x4
philox4x32_10 (u32x4 c, uint32_t k0, uint32_t k1)
{
  for (int r = 0; r < 10; ++r)
  {
    const uint64_t p0 = (uint64_t) 0xD2511F53u * c.x;
    const uint64_t p1 = (uint64_t) 0xCD9E8D57u * c.z;
    const u32x4 n     = {(uint32_t) (p1 >> 32) ^ c.y ^ k0, (uint32_t) p1, (uint32_t) (p0 >> 32) ^ c.w ^ k1,
                         (uint32_t) p0};

    c = n;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }

  return c;
}

__device__ inline double
u53 (uint32_t a, uint32_t b)
{
  return ((double) (a >> 5) * 67108864.0 + (double) (b >> 6)) * (1.0 / 9007199254740992.0);
}

struct DrawStream
{
  uint32_t j, q, c0, c1, k0, k1;
  double spare;
  bool has_spare;

  __device__ u32x4
  next ()
  {
    return philox4x32_10 ({j, q++, c0, c1}, k0, k1);
  }

  __device__ double
  uniform ()
  {
    const u32x4 r = next ();

    return u53 (r.x, r.y);
  }

  __device__ double
  normal ()
  {
    if (has_spare)
    {
      has_spare = false;
      return spare;
    }

    const u32x4 r  = next ();
    const double a = 1.0 - u53 (r.x, r.y);   // (0, 1]
    const double b = u53 (r.z, r.w);
    const double s = sqrt (-2.0 * log (a));
    double sn, cs;

    sincospi (2.0 * b, &sn, &cs);
    spare     = s * sn;
    has_spare = true;

    return s * cs;
  }
};

__global__ void
k_draws (int m, int d, int nu_int, uint64_t key, uint64_t ctr, double *__restrict__ u, double *__restrict__ z,
         double *__restrict__ chi2)
{
  const int j = blockIdx.x * blockDim.x + threadIdx.x;

  if (j >= m)
    return;

  DrawStream st {(uint32_t) j, 0u, (uint32_t) ctr, (uint32_t) (ctr >> 32), (uint32_t) key, (uint32_t) (key >> 32),
                 0.0, false};

  u[j] = st.uniform ();

  for (int i = 0; i < d; ++i)
    z[(size_t) j * d + i] = st.normal ();

  if (chi2)
  {
    double c = 0.0;

    for (int k = 0; k < nu_int; ++k)
    {
      const double g = st.normal ();

      c += g * g;
    }

    chi2[j] = c;
  }
}

hipError_t
launch_draws (int m, int d, int nu_int, uint64_t key, uint64_t ctr, double *u, double *z, double *chi2,
              hipStream_t s)
{
  if (m > 0)
    hipLaunchKernelGGL (k_draws, dim3 ((m + 127) / 128), dim3 (128), 0, s, m, d, nu_int, key, ctr, u, z, chi2);

  return hipGetLastError ();
}

hipError_t
launch_sample (const double *wcum, int n, const double *u, const double *z, const double *chi2, int m, int d,
               int kind, int kernel, double nu, double href, const double *U, const double *X, double *out,
               uint32_t *idx, hipStream_t s)
{
  hipLaunchKernelGGL (k_sample, dim3 ((m + 127) / 128), dim3 (128), 0, s, wcum, n, u, z, chi2, m, d, kind, kernel,
                      nu, href, U, X, out, idx, (const uint32_t *) nullptr);

  return hipGetLastError ();
}

hipError_t
launch_sample_at (const uint32_t *idx_in, const double *z, const double *chi2, int m, int d, int kind, int kernel,
                  double nu, double href, const double *U, const double *X, double *out, hipStream_t s)
{
  if (m > 0)
    hipLaunchKernelGGL (k_sample, dim3 ((m + 127) / 128), dim3 (128), 0, s, (const double *) nullptr, 0,
                        (const double *) nullptr, z, chi2, m, d, kind, kernel, nu, href, U, X, out,
                        (uint32_t *) nullptr, idx_in);

  return hipGetLastError ();
}

// Antithetic kernel pairs of ncm_stats_dist_sample2 (ncm_stats_dist.c:1629-1651):
// the first draw of pair t chooses i from its uniform; the pair samples the
// kernels order[i] and order[n-1-i].
__global__ void
k_pair_index (const double *__restrict__ wcum, int n, const double *__restrict__ u, int npairs,
              const uint32_t *__restrict__ order, uint32_t *__restrict__ idx)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;

  if (t >= npairs)
    return;

  const int i = choose (wcum, n, u[2 * t]);

  idx[2 * t]     = order[i];
  idx[2 * t + 1] = order[n - 1 - i];
}

hipError_t
launch_pair_index (const double *wcum, int n, const double *u, int npairs, const uint32_t *order, uint32_t *idx,
                   hipStream_t s)
{
  if (npairs > 0)
    hipLaunchKernelGGL (k_pair_index, dim3 ((npairs + 255) / 256), dim3 (256), 0, s, wcum, n, u, npairs, order,
                        idx);

  return hipGetLastError ();
}

} // namespace nck
