// k_eval.hip -- FP64 eval_m2lnp / eval for KDE and VKDE.
//
// One lane per evaluation point; every lane of a block walks the same kernel
// range, so the kernel factors (packed W_i = U_i^{-T}, centre, b_i) are
// wave-uniform and come in through scalar loads straight into SGPR operands
// of the FP64 FMAs.  Per pair (VKDE, padded dim D): D subtractions,
// D(D+1)/2 FMAs for z = W_i (x - theta_i), D FMAs for |z|^2, one exp (two
// transcendentals for Student-t).  The log-sum-exp runs online in FP64 with
// a running maximum.  The kernel range is split over blockIdx.y so that small
// point counts still fill the 256 CUs; partial (max, sum) pairs are merged in
// split order by k_eval_merge.
//
// Reference: ncm_stats_dist_vkde.c:681-723 and ncm_stats_dist_kde.c:639-681
// (eval_weights_m2lnp), the sum0/sum1 reductions of
// ncm_stats_dist_kernel_gauss.c:246-333 / ncm_stats_dist_kernel_st.c:295-386
// and the linear eval_weights (vkde.c:631-679, kde.c:596-637).
#include "nckde_internal.h"

namespace nck
{

template <int D, int KERNEL, int KIND>
__global__ void __launch_bounds__ (256)
k_eval_f64 (EvalArgs a)
{
  const int p     = blockIdx.x * 256 + threadIdx.x;
  const int chunk = (a.n + a.nsplit - 1) / a.nsplit;
  const int i0    = blockIdx.y * chunk;
  const int i1    = min (a.n, i0 + chunk);
  const int pp    = (p < a.m) ? p : (a.m - 1);
  const int d     = a.d;
  // VKDE at D = 64: the point is read again for every kernel (from L1/L2)
  // instead of being held beside the differences, which left one wave per
  // SIMD (1.5x as long at n = 8000, m = 20000; at D = 32, and in the FP32
  // rung, the re-reads cost more than the occupancy gains).  KDE keeps its
  // whitened point in registers.
  constexpr bool XREG = (KIND == NCKDE_KDE) || (D < 64);
  const double *xrow  = a.X + (size_t) pp * d;
  double x[XREG ? D : 1];
  auto xv = [&] (int k) -> double {
    if constexpr (XREG)
      return x[k];
    else
      return (k < d) ? xrow[k] : 0.0;
  };

  if constexpr (XREG)
  {
#pragma unroll
    for (int k = 0; k < D; ++k)
      x[k] = (k < d) ? xrow[k] : 0.0;
  }

  if constexpr (KIND == NCKDE_KDE)
  {
    // v = U_g^{-T} x (gsl_blas_dtrsv Upper/Trans, kde.c:648-652)
    double v[D];

#pragma unroll
    for (int l = 0; l < D; ++l)
    {
      if (l < d)
      {
        double t = x[l];

#pragma unroll
        for (int m = 0; m < l; ++m)
          t = fma (-v[m], a.Ug[m * d + l], t);

        v[l] = t / a.Ug[l * d + l];
      }
      else
      {
        v[l] = 0.0;
      }
    }

#pragma unroll
    for (int l = 0; l < D; ++l)
      x[l] = v[l];
  }

  double M = -INFINITY, S = 0.0;

  for (int i = i0; i < i1; ++i)
  {
    const double *th = a.Th + (size_t) i * D;
    double chi2      = 0.0;

    if (KIND == NCKDE_VKDE)
    {
      const double *W = a.Wp + (size_t) i * (D * (D + 1) / 2);
      double delta[D];

#pragma unroll
      for (int k = 0; k < D; ++k)
        delta[k] = xv (k) - th[k];

#pragma unroll
      for (int r = 0; r < D; ++r)
      {
        double z = 0.0;

#pragma unroll
        for (int c = 0; c <= r; ++c)
          z = fma (W[r * (r + 1) / 2 + c], delta[c], z);

        chi2 = fma (z, z, chi2);
      }
    }
    else
    {
#pragma unroll
      for (int k = 0; k < D; ++k)
      {
        const double t = th[k] - xv (k);

        chi2 = fma (t, t, chi2);
      }
    }

    chi2 *= a.inv_h2;

    double lnt;

    if (KERNEL == NCKDE_KERNEL_GAUSS)
      lnt = fma (-0.5, chi2, a.b[i]);
    else
      lnt = fma (a.kappa, log1p (chi2 / a.nu), a.b[i]);

    if (lnt > M)
    {
      S = S * exp (M - lnt);
      M = lnt;
    }

    {
      const double e = exp (lnt - M);

      S += (lnt > -INFINITY) ? e : 0.0;
    }
  }

  if (p < a.m)
  {
    a.partM[(size_t) blockIdx.y * a.m + p] = M;
    a.partS[(size_t) blockIdx.y * a.m + p] = S;
  }
}

// The FP32 rung of the precision ladder: the same walk in single precision.
// x - theta_i is formed in FP64 and rounded to FP32; W_i comes from its FP32
// copy; |z|^2 and the kernel run in FP32 against a fixed reference,
// lnt = b2_i + ln K with b2_i = b_i - max_j b_j (FP32), with an online FP32
// maximum and the sum in FP64.  The partial maximum is stored back in
// natural units (+ max_j b_j) so that k_eval_merge serves both forms.
template <int D, int KERNEL, int KIND>
__global__ void __launch_bounds__ (256)
k_eval_f32 (EvalArgs a)
{
  const int p     = blockIdx.x * 256 + threadIdx.x;
  const int chunk = (a.n + a.nsplit - 1) / a.nsplit;
  const int i0    = blockIdx.y * chunk;
  const int i1    = min (a.n, i0 + chunk);
  const int pp    = (p < a.m) ? p : (a.m - 1);
  const int d     = a.d;
  double x[D];

#pragma unroll
  for (int k = 0; k < D; ++k)
    x[k] = (k < d) ? a.X[(size_t) pp * d + k] : 0.0;

  if (KIND == NCKDE_KDE)
  {
    double v[D];

#pragma unroll
    for (int l = 0; l < D; ++l)
    {
      if (l < d)
      {
        double t = x[l];

#pragma unroll
        for (int m = 0; m < l; ++m)
          t = fma (-v[m], a.Ug[m * d + l], t);

        v[l] = t / a.Ug[l * d + l];
      }
      else
      {
        v[l] = 0.0;
      }
    }

#pragma unroll
    for (int l = 0; l < D; ++l)
      x[l] = v[l];
  }

  const float inv_h2 = (float) a.inv_h2;
  const float kappa  = (float) a.kappa;
  const float inv_nu = (float) (1.0 / a.nu);
  float M            = -INFINITY;
  double S           = 0.0;

  for (int i = i0; i < i1; ++i)
  {
    const double *th = a.Th + (size_t) i * D;
    float chi2       = 0.0f;

    if (KIND == NCKDE_VKDE)
    {
      const float *W = a.Wf + (size_t) i * (D * (D + 1) / 2);
      float delta[D];

#pragma unroll
      for (int k = 0; k < D; ++k)
        delta[k] = (float) (x[k] - th[k]);

#pragma unroll
      for (int r = 0; r < D; ++r)
      {
        float z = 0.0f;

#pragma unroll
        for (int c = 0; c <= r; ++c)
          z = fmaf (W[r * (r + 1) / 2 + c], delta[c], z);

        chi2 = fmaf (z, z, chi2);

        // one row's factor entries live at a time: without the fence the
        // scalar loads of every row are hoisted together and spill at D >= 32
        if (D >= 32)
          __builtin_amdgcn_sched_barrier (0);
      }
    }
    else
    {
#pragma unroll
      for (int k = 0; k < D; ++k)
      {
        const float t = (float) (th[k] - x[k]);

        chi2 = fmaf (t, t, chi2);
      }
    }

    chi2 *= inv_h2;

    float lnt;

    if (KERNEL == NCKDE_KERNEL_GAUSS)
      lnt = fmaf (-0.5f, chi2, a.b2f[i]);
    else
      lnt = fmaf (kappa, log1pf (chi2 * inv_nu), a.b2f[i]);

    if (lnt > M)
    {
      S = S * (double) expf (M - lnt);
      M = lnt;
    }

    S += (lnt > -INFINITY) ? (double) expf (lnt - M) : 0.0;
  }

  if (p < a.m)
  {
    a.partM[(size_t) blockIdx.y * a.m + p] = (M > -INFINITY) ? (double) M + a.bref : -INFINITY;
    a.partS[(size_t) blockIdx.y * a.m + p] = S;
  }
}

// FP32 copies of the packed factors and the shifted b_i
__global__ void
k_f32_factors (const double *__restrict__ Wp, size_t count, float *__restrict__ Wf)
{
  const size_t e = blockIdx.x * (size_t) blockDim.x + threadIdx.x;

  if (e < count)
    Wf[e] = (float) Wp[e];
}

__global__ void
k_f32_b2 (const double *__restrict__ b, double bref, int n, float *__restrict__ b2f)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;

  if (i < n)
    b2f[i] = (float) (b[i] - bref);
}

hipError_t
launch_f32_factors (const double *Wp, size_t count, float *Wf, hipStream_t s)
{
  if (count > 0)
    hipLaunchKernelGGL (k_f32_factors, dim3 ((unsigned) ((count + 255) / 256)), dim3 (256), 0, s, Wp, count, Wf);

  return hipGetLastError ();
}

hipError_t
launch_f32_b2 (const double *b, double bref, int n, float *b2f, hipStream_t s)
{
  hipLaunchKernelGGL (k_f32_b2, dim3 ((n + 255) / 256), dim3 (256), 0, s, b, bref, n, b2f);

  return hipGetLastError ();
}

// gamma = max over splits, 1 + lambda = sum_s S_s exp(M_s - gamma);
// m2lnp = -2 (gamma - lnnorm_shared + log1p(lambda) - d ln h)
__global__ void
k_eval_merge (EvalArgs a)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;

  if (p >= a.m)
    return;

  double M = -INFINITY;

  for (int s = 0; s < a.nsplit; ++s)
    M = fmax (M, a.partM[(size_t) s * a.m + p]);

  double S = 0.0;

  for (int s = 0; s < a.nsplit; ++s)
  {
    const double Ms = a.partM[(size_t) s * a.m + p];

    if (Ms > -INFINITY)
      S += a.partS[(size_t) s * a.m + p] * exp (Ms - M);
  }

  const double m2lnp = -2.0 * ((M - a.lnnorm_shared) + log1p (S - 1.0) - a.lnh_d);

  a.out[p] = a.linear ? exp (-0.5 * m2lnp) : m2lnp;
}

template <int D, int KERNEL, int KIND>
static hipError_t
launch_t (const EvalArgs &a, hipStream_t s)
{
  dim3 grid ((a.m + 255) / 256, a.nsplit);

  if (a.b2f != nullptr)
    hipLaunchKernelGGL ((k_eval_f32<D, KERNEL, KIND>), grid, dim3 (256), 0, s, a);
  else
    hipLaunchKernelGGL ((k_eval_f64<D, KERNEL, KIND>), grid, dim3 (256), 0, s, a);

  return hipGetLastError ();
}

template <int D>
static hipError_t
launch_d (const EvalArgs &a, hipStream_t s)
{
  if (a.kind == NCKDE_VKDE)
    return (a.kernel == NCKDE_KERNEL_GAUSS) ? launch_t<D, NCKDE_KERNEL_GAUSS, NCKDE_VKDE> (a, s)
                                            : launch_t<D, NCKDE_KERNEL_ST, NCKDE_VKDE> (a, s);

  return (a.kernel == NCKDE_KERNEL_GAUSS) ? launch_t<D, NCKDE_KERNEL_GAUSS, NCKDE_KDE> (a, s)
                                          : launch_t<D, NCKDE_KERNEL_ST, NCKDE_KDE> (a, s);
}

int
padded_dim (int d)
{
  static const int dims[] = {4, 8, 12, 16, 24, 32, 48, 64};

  for (int D : dims)
    if (d <= D)
      return D;

  return -1;
}

hipError_t
launch_eval_f64 (const EvalArgs &a, hipStream_t s)
{
  switch (a.D)
  {
    case 4: return launch_d<4> (a, s);
    case 8: return launch_d<8> (a, s);
    case 12: return launch_d<12> (a, s);
    case 16: return launch_d<16> (a, s);
    case 24: return launch_d<24> (a, s);
    case 32: return launch_d<32> (a, s);
    case 48: return launch_d<48> (a, s);
    case 64: return launch_d<64> (a, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t
launch_eval_merge (const EvalArgs &a, hipStream_t s)
{
  hipLaunchKernelGGL (k_eval_merge, dim3 ((a.m + 255) / 256), dim3 (256), 0, s, a);

  return hipGetLastError ();
}

} // namespace nck
