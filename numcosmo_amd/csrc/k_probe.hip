// k_probe.hip -- device side of the check of freshly built mixed-precision
// factors (see mixed_probe in api.cpp): the kernels with the largest risk
// score, the probe points' variates, and the largest relative difference
// between the two precisions' results.  Everything stays on the device; the
// host reads back one number.
#include "nckde_internal.h"

namespace nck
{

// Indices of the np largest scores, ties by smaller index: a radix select of
// the np-th smallest key ~bits(score) (the scores are non-negative floats, so
// their bit patterns are monotone), then the keys below it and the smallest
// indices equal to it.  One 1024-thread block; np <= 1024.
__global__ void __launch_bounds__ (1024)
k_probe_select (const float *__restrict__ risk, int n, int np, uint32_t *__restrict__ idx)
{
  __shared__ unsigned hist[256];
  __shared__ unsigned s_prefix;
  __shared__ int s_kk, s_nlt, s_eq;
  const int tid = threadIdx.x;

  if (tid == 0)
  {
    s_prefix = 0;
    s_kk     = np;
    s_nlt    = 0;
    s_eq     = 0;
  }

  unsigned mask = 0;

  for (int pass = 0; pass < 4; ++pass)
  {
    const int shift = 24 - 8 * pass;

    if (tid < 256)
      hist[tid] = 0;

    __syncthreads ();

    const unsigned prefix = s_prefix;

    for (int i = tid; i < n; i += 1024)
    {
      const unsigned key = ~__float_as_uint (fmaxf (risk[i], 0.0f));

      if ((key & mask) == prefix)
        atomicAdd (&hist[(key >> shift) & 255u], 1u);
    }

    __syncthreads ();

    if (tid == 0)
    {
      // the digit holding rank s_kk (serial over 256 bins, once per pass)
      unsigned below = 0;
      const int kk   = s_kk;

      for (int b = 0; b < 256; ++b)
      {
        if (below + hist[b] >= (unsigned) kk)
        {
          s_kk     = kk - (int) below;
          s_prefix = prefix | ((unsigned) b << shift);
          break;
        }

        below += hist[b];
      }
    }

    mask |= 255u << shift;
    __syncthreads ();
  }

  const unsigned T  = s_prefix;
  const int need_eq = s_kk;

  // keys below T in any order, then the need_eq smallest indices equal to T,
  // chunk by chunk in index order (ballot ranks within a wave, wave totals
  // through LDS)
  __shared__ int s_wcnt[16];

  for (int i = tid; i < n; i += 1024)
  {
    const unsigned key = ~__float_as_uint (fmaxf (risk[i], 0.0f));

    if (key < T)
    {
      const int slot = atomicAdd (&s_nlt, 1);

      if (slot < np)
        idx[slot] = (uint32_t) i;
    }
  }

  __syncthreads ();

  const int nlt  = s_nlt;
  const int lane = tid & 63, wave = tid >> 6;

  for (int base = 0; base < n; base += 1024)
  {
    const int taken = s_eq;

    if (taken >= need_eq)
      break;

    const int i        = base + tid;
    const bool eq      = (i < n) && (~__float_as_uint (fmaxf (risk[i], 0.0f)) == T);
    const uint64_t bal = __ballot (eq);
    const int rank_w   = __popcll (bal & ((1ull << lane) - 1ull));

    if (lane == 0)
      s_wcnt[wave] = __popcll (bal);

    __syncthreads ();

    int before = 0, total = 0;

    for (int w = 0; w < 16; ++w)
    {
      before += (w < wave) ? s_wcnt[w] : 0;
      total += s_wcnt[w];
    }

    if (eq && taken + before + rank_w < need_eq)
      idx[nlt + taken + before + rank_w] = (uint32_t) i;

    __syncthreads ();

    if (tid == 0)
      s_eq = taken + total;

    __syncthreads ();
  }
}

// probe p uses kernel idx[p / 2] with z = c (1, s, 1, s, ...), s = +1 for
// even p and -1 for odd p, and chi2 = chi2v (the sample transform's
// Student-t scale sqrt(nu / chi2) is then 1 for chi2v = nu)
__global__ void
k_probe_fill (const uint32_t *__restrict__ sel, int m, int d, double c, double chi2v, uint32_t *__restrict__ idx,
              double *__restrict__ z, double *__restrict__ chi2)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;

  if (p >= m)
    return;

  idx[p]  = sel[p / 2];
  chi2[p] = chi2v;

  for (int l = 0; l < d; ++l)
    z[(size_t) p * d + l] = ((p & 1) && (l & 1)) ? -c : c;
}

// largest |a - b| / max(|b|, 1) over m results (NaN counts as infinite)
__global__ void __launch_bounds__ (256)
k_probe_err (const double *__restrict__ a, const double *__restrict__ b, int m, double *__restrict__ out)
{
  __shared__ double red[256];
  double e = 0.0;

  for (int p = threadIdx.x; p < m; p += 256)
  {
    const double r = fabs (a[p] - b[p]) / fmax (fabs (b[p]), 1.0);

    e = (r == r) ? fmax (e, r) : INFINITY;
  }

  red[threadIdx.x] = e;
  __syncthreads ();

  for (int o = 128; o > 0; o >>= 1)
  {
    if (threadIdx.x < o)
      red[threadIdx.x] = fmax (red[threadIdx.x], red[threadIdx.x + o]);

    __syncthreads ();
  }

  if (threadIdx.x == 0)
    out[0] = red[0];
}

hipError_t
launch_probe_points (const float *risk, int n, int np, int d, double c, double chi2v, uint32_t *sel, uint32_t *idx,
                     double *z, double *chi2, hipStream_t s)
{
  const int m = 2 * np;

  hipLaunchKernelGGL (k_probe_select, dim3 (1), dim3 (1024), 0, s, risk, n, np, sel);
  hipLaunchKernelGGL (k_probe_fill, dim3 ((m + 127) / 128), dim3 (128), 0, s, sel, m, d, c, chi2v, idx, z, chi2);

  return hipGetLastError ();
}

hipError_t
launch_probe_err (const double *a, const double *b, int m, double *out, hipStream_t s)
{
  hipLaunchKernelGGL (k_probe_err, dim3 (1), dim3 (256), 0, s, a, b, m, out);

  return hipGetLastError ();
}

} // namespace nck
