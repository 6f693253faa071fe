// k_eval_mixed.hip -- mixed-precision VKDE eval_m2lnp on the MFMA units.
//
// For kernel i and point x (natural log units)
//   chi2_i(x) = | U_i^{-T} (x - theta_i) |^2 / h^2 = | B_i (y(x) - y_i) |^2,
//   y(x) = G (x - mu), G = U_g^{-T} (global whitening), B_i = U_i^{-T} U_g^T / h.
// With power-of-two scales ysc (points) and s_i (kernel),
//   z'_i(x) = A_i (ysc y(x)) - c_i,  A_i = (s_i / ysc) B_i,  c_i = A_i (ysc y_i),
//   chi2_i = |z'_i|^2 / s_i^2.
// A_i and ysc y(x) are split into FP16 hi + lo parts (22 significant bits);
// z' is accumulated in FP32 by v_mfma_f32_32x32x16_f16 from three products
// (A_hi y_hi + A_hi y_lo + A_lo y_hi), with -c_i (stored negated) as the
// accumulator input.
// Each 32-row MFMA tile holds 32/DP kernels (DP = padded dimension 4, 8, 16)
// or one kernel over 1 or 2 row tiles (DP = 32, 64); for DP < 16 the hi/lo
// operands are packed along K so no K lanes are wasted.
//
// The kernel sum runs on the VALU in base 2: lnt2 = log2(e) lnt - b_ref with
//   Gauss  lnt2 = q_i |z'|^2 + b2_i,               q_i = -log2(e) / (2 s_i^2)
//   ST     lnt2 = kappa log2(1 + r_i |z'|^2) + b2_i, r_i = 1 / (nu s_i^2)
// and an online maximum M (FP32) with the scaled sum S in FP64.  Partial
// (M, S) per kernel-range split are merged by k_mixed_merge in a fixed order.
//
// Reference: ncm_stats_dist_vkde.c:681-723 with the sum0 reduction of
// ncm_stats_dist_kernel_gauss.c:246-290 / ncm_stats_dist_kernel_st.c:295-341.
#include "nckde_internal.h"

#include <cfloat>
#include <cstring>

namespace nck
{

typedef _Float16 half8 __attribute__ ((ext_vector_type (8)));
typedef _Float16 half4 __attribute__ ((ext_vector_type (4)));
typedef float floatx16 __attribute__ ((ext_vector_type (16)));

constexpr double LOG2E = 1.4426950408889634073599246810018921;
constexpr double LN2   = 0.69314718055994530941723212145817657;

template <int DP>
struct Plan
{
  static constexpr int KPT    = (DP >= 32) ? 1 : 32 / DP;  // kernels per 32-row tile
  static constexpr int NRT    = (DP == 64) ? 2 : 1;        // row tiles per kernel group
  static constexpr int NKB    = (DP >= 32) ? DP / 16 : 1;  // 16-wide K blocks
  static constexpr int NLOCAL = (DP == 4) ? 4 : (DP == 8 ? 2 : 1);   // kernels per lane
};

int
mfma_dim (int d)
{
  if (d <= 4)
    return 4;

  if (d <= 8)
    return 8;

  if (d <= 16)
    return 16;

  if (d <= 32)
    return 32;

  if (d <= 64)
    return 64;

  return -1;
}

// --------------------------------------------------------------------------
// Points: y' = U_g^{-T} (x - mu) in FP64, scaled by ysc (and, for points
// beyond the representable range, by a per-point power of two t_p), split
// into FP16 hi/lo.  Layout: Yh[p][DP], Yl[p][DP]; tp[p] = (t_p, 1/t_p^2).
// --------------------------------------------------------------------------
template <int DP>
__global__ void __launch_bounds__ (128)
k_mixed_points (const double *__restrict__ X, int m, int d, const double *__restrict__ mu,
                const double *__restrict__ Ug, double ysc, _Float16 *__restrict__ Yh, _Float16 *__restrict__ Yl,
                float2 *__restrict__ tp)
{
  const long p = blockIdx.x * (long) blockDim.x + threadIdx.x;

  if (p >= m)
    return;

  // y = U_g^{-T} (x - mu) by forward substitution, in registers (DP is the
  // padded dimension; the loops are unrolled, entries past d stay 0)
  double y[DP];
  double ymax = 0.0;

#pragma unroll
  for (int l = 0; l < DP; ++l)
  {
    y[l] = 0.0;

    if (l < d)
    {
      double t = X[p * d + l] - mu[l];

#pragma unroll
      for (int q = 0; q < l; ++q)
        t = fma (-y[q], Ug[q * d + l], t);

      y[l] = t / Ug[l * d + l];
      ymax = fmax (ymax, fabs (y[l]));
    }
  }

  double sc = ysc;

  // keep |sc y| below 2^14 (FP16 max is 65504); bounded for inf/NaN input
  for (int it = 0; it < 80 && ymax * sc >= 16384.0; ++it)
    sc *= 0.0625;

  const double t_p = sc / ysc;
  _Float16 hi[DP], lo[DP];

#pragma unroll
  for (int l = 0; l < DP; ++l)
  {
    const double v = y[l] * sc;

    hi[l] = (_Float16) v;
    lo[l] = (_Float16) (v - (double) hi[l]);
  }

  // whole-row vector stores (DP halves = 8 or more bytes)
  if constexpr (DP % 8 == 0)
  {
#pragma unroll
    for (int c = 0; c < DP / 8; ++c)
    {
      half8 vh, vl;

#pragma unroll
      for (int k = 0; k < 8; ++k)
      {
        vh[k] = hi[c * 8 + k];
        vl[k] = lo[c * 8 + k];
      }

      *(half8 *) (Yh + p * DP + c * 8) = vh;
      *(half8 *) (Yl + p * DP + c * 8) = vl;
    }
  }
  else
  {
#pragma unroll
    for (int l = 0; l < DP; ++l)
    {
      Yh[p * DP + l] = hi[l];
      Yl[p * DP + l] = lo[l];
    }
  }

  tp[p] = make_float2 ((float) t_p, (float) (1.0 / (t_p * t_p)));
}

hipError_t
launch_mixed_points (const MixedArgs &a, hipStream_t s)
{
  _Float16 *Yh = (_Float16 *) a.Yh;
  _Float16 *Yl = Yh + (size_t) a.m * a.f.DP;
  float2 *tp   = (float2 *) (Yl + (size_t) a.m * a.f.DP);
  const dim3 grid ((a.m + 127) / 128), block (128);

  switch (a.f.DP)
  {
    case 4: hipLaunchKernelGGL (k_mixed_points<4>, grid, block, 0, s, a.X, a.m, a.d, a.mu, a.G, a.ysc, Yh, Yl, tp); break;
    case 8: hipLaunchKernelGGL (k_mixed_points<8>, grid, block, 0, s, a.X, a.m, a.d, a.mu, a.G, a.ysc, Yh, Yl, tp); break;
    case 16: hipLaunchKernelGGL (k_mixed_points<16>, grid, block, 0, s, a.X, a.m, a.d, a.mu, a.G, a.ysc, Yh, Yl, tp); break;
    case 32: hipLaunchKernelGGL (k_mixed_points<32>, grid, block, 0, s, a.X, a.m, a.d, a.mu, a.G, a.ysc, Yh, Yl, tp); break;
    case 64: hipLaunchKernelGGL (k_mixed_points<64>, grid, block, 0, s, a.X, a.m, a.d, a.mu, a.G, a.ysc, Yh, Yl, tp); break;
    default: return hipErrorInvalidValue;
  }

  return hipGetLastError ();
}

// --------------------------------------------------------------------------
// Kernel factors (FP64 -> split), one wave per kernel slot (padded to a whole
// number of tiles; padding kernels get A = 0, c = 0, b2 = -inf).
//   A (half): [n_pad][2][DP][DP] (hi, lo);  c (float): [n_pad][DP];
//   coef (float4): (q_i or r_i, b2_i, 1/s_i^2, 0).
// Wp holds W_i = U_i^{-T} packed by rows with padded dimension D (VKDE; NULL
// for KDE, whose kernels all share U_g); Yc holds y'_i = U_g^{-T} (theta_i - mu).
// Lanes take the entries of the DP x DP block in row-major order (coalesced
// half stores); lane r < d sums row r of c serially in q, so every value is
// the one a serial loop over (r, q) gives.
// --------------------------------------------------------------------------
__global__ void __launch_bounds__ (256)
k_mixed_factors (const double *__restrict__ Wp, int D, const double *__restrict__ Yc, const double *__restrict__ Ug,
                 const double *__restrict__ b, double bref, int n, int n_pad, int d, int DP, double inv_h,
                 double ysc, int kernel, double nu, _Float16 *__restrict__ A, float *__restrict__ c,
                 float4 *__restrict__ coef, float *__restrict__ risk)
{
  const long i   = blockIdx.x * 4L + (threadIdx.x >> 6);   // wave-uniform
  const int lane = threadIdx.x & 63;

  if (i >= n_pad)
    return;

  _Float16 *Ah = A + i * 2 * (size_t) DP * DP;
  _Float16 *Al = Ah + (size_t) DP * DP;
  float *ci    = c + i * (size_t) DP;

  if (i >= n)
  {
    for (int e = lane; e < DP * DP; e += 64)
    {
      Ah[e] = (_Float16) 0.0f;
      Al[e] = (_Float16) 0.0f;
    }

    if (lane < DP)
      ci[lane] = 0.0f;

    if (lane == 0)
      coef[i] = make_float4 (0.0f, -INFINITY, 0.0f, 0.0f);

    return;
  }

  // VKDE: B[r][q] = sum_{k=q..r} W[r][k] U_g[q][k] (W = U_i^{-T}); KDE
  // (Wp == NULL): every kernel shares U_g, so B is the identity
  const double *W  = Wp ? Wp + i * (size_t) (D * (D + 1) / 2) : nullptr;
  const double *yi = Yc + i * (size_t) d;
  auto bval        = [&] (int r, int q) {
    if (W == nullptr)
      return (r == q) ? 1.0 : 0.0;

    double t = 0.0;

    for (int k = q; k <= r; ++k)
      t = fma (W[r * (r + 1) / 2 + k], Ug[q * d + k], t);

    return t;
  };
  double bmax = 0.0;

  for (int e = lane; e < d * d; e += 64)
    bmax = fmax (bmax, fabs (bval (e / d, e % d) * inv_h));

#pragma unroll
  for (int o = 32; o >= 1; o >>= 1)
    bmax = fmax (bmax, __shfl_xor (bmax, o, 64));

  // risk score of the split products: their absolute error in z grows with
  // |B_i| |y| while what matters near the kernel is |B_i (y - y_i)|, so
  // narrow kernels far from the sample mean are the ones to probe
  double yabs = 0.0;

  for (int q = lane; q < d; q += 64)
    yabs = fmax (yabs, fabs (yi[q]));

#pragma unroll
  for (int o = 32; o >= 1; o >>= 1)
    yabs = fmax (yabs, __shfl_xor (yabs, o, 64));

  if (lane == 0)
    risk[i] = (float) (bmax * (yabs + 1.0));

  // s_i / ysc brings max |A_i| into [2^13, 2^14)
  const int e      = (bmax > 0.0) ? ilogb (bmax / ysc) : 0;
  const double s_i = ldexp (1.0, 13 - e);
  const double sy  = s_i / ysc;

  for (int f = lane; f < DP * DP; f += 64)
  {
    const int r    = f / DP;
    const int q    = f % DP;
    const double v = (r < d && q < d) ? bval (r, q) * inv_h * sy : 0.0;
    const _Float16 hi = (_Float16) v;
    const _Float16 lo = (_Float16) (v - (double) hi);

    Ah[f] = hi;
    Al[f] = lo;
  }

  if (lane < DP)
  {
    double cr = 0.0;

    if (lane < d)
      for (int q = 0; q < d; ++q)
        cr = fma (bval (lane, q) * inv_h * sy, yi[q] * ysc, cr);

    // stored negated: the MFMA accumulator starts from -B y_i directly
    ci[lane] = (float) -cr;
  }

  if (lane == 0)
  {
    const double inv_s2 = 1.0 / (s_i * s_i);
    const double qr     = (kernel == NCKDE_KERNEL_GAUSS) ? -0.5 * LOG2E * inv_s2 : inv_s2 / nu;

    coef[i] = make_float4 ((float) qr, (float) ((b[i] - bref) * LOG2E), (float) inv_s2, 0.0f);
  }
}

// b2_i refresh after the weights change
__global__ void
k_mixed_coef_b (const double *__restrict__ b, double bref, int n, float4 *__restrict__ coef)
{
  const long i = blockIdx.x * (long) blockDim.x + threadIdx.x;

  if (i < n)
    coef[i].y = (float) ((b[i] - bref) * LOG2E);
}

// y'_i = U_g^{-T} (theta_i - mu) for the kernel centres, and max |y'|
__global__ void
k_mixed_centres (const double *__restrict__ X, int n, int d, const double *__restrict__ mu,
                 const double *__restrict__ Ug, double *__restrict__ Yc, unsigned long long *__restrict__ ymax_bits)
{
  const long i = blockIdx.x * (long) blockDim.x + threadIdx.x;

  if (i >= n)
    return;

  double ym = 0.0;

  for (int l = 0; l < d; ++l)
  {
    double t = X[i * d + l] - mu[l];

    for (int q = 0; q < l; ++q)
      t = fma (-Yc[i * d + q], Ug[q * d + l], t);

    t            = t / Ug[l * d + l];
    Yc[i * d + l] = t;
    ym           = fmax (ym, fabs (t));
  }

  atomicMax (ymax_bits, (unsigned long long) __double_as_longlong (ym));
}

// column means of the sample: one block per column, strided partial sums
// and a fixed-order tree reduction (deterministic)
__global__ void __launch_bounds__ (256)
k_colmean (const double *__restrict__ X, int n, int d, double *__restrict__ mu)
{
  __shared__ double red[256];
  const int l = blockIdx.x;
  double s    = 0.0;

  for (int i = threadIdx.x; i < n; i += 256)
    s += X[(size_t) i * d + l];

  red[threadIdx.x] = s;
  __syncthreads ();

  for (int o = 128; o > 0; o >>= 1)
  {
    if (threadIdx.x < o)
      red[threadIdx.x] += red[threadIdx.x + o];

    __syncthreads ();
  }

  if (threadIdx.x == 0)
    mu[l] = red[0] / n;
}

// --------------------------------------------------------------------------
// The evaluation kernel.  4 waves per block (8 at DP = 64); each wave owns NPT tiles of
// 32 points whose B fragments stay in registers; every wave of the grid walks
// a range of kernel groups.
// --------------------------------------------------------------------------
// The group's factors start at a wave-uniform address (scalar registers);
// each lane adds a 32-bit offset, so the loads take the scalar-base form and
// need no 64-bit vector address arithmetic.
//
// Row map of a 32-row tile.  The 32x32 accumulator gives lane l the rows
// 8 j + 4 (l / 32) + r (j, r < 4).  Kernel rows are placed so that these
// rows hold whole kernels and no lane needs a partner's partial sum:
//   DP = 4:  row R = kernel R / 4, dim R % 4 (kernels 2 j + g);
//   DP = 8:  row 8 j + 4 g + r = kernel 2 g + j / 2, dim 4 (j % 2) + r;
//   DP = 16: row 8 j + 4 g + r = kernel g, dim 4 j + r.
// At DP >= 32 one kernel fills the tile and the half-waves add partials.
template <int DP>
__device__ inline void
tile_row (unsigned R, unsigned &kl, unsigned &dim)
{
  if (DP == 4)
  {
    kl  = R >> 2;
    dim = R & 3;
  }
  else if (DP == 8)
  {
    kl  = 2 * ((R >> 2) & 1) + (R >> 4);
    dim = ((R >> 3) & 1) * 4 + (R & 3);
  }
  else if (DP == 16)
  {
    kl  = (R >> 2) & 1;
    dim = (R >> 3) * 4 + (R & 3);
  }
  else
  {
    kl  = 0;
    dim = R;
  }
}
template <int DP>
__device__ inline half8
load_a_frag (const _Float16 *__restrict__ A, int tg, int rt, int kb, int lohi, int lane)
{
  using P            = Plan<DP>;
  const unsigned g   = lane >> 5;
  const unsigned R   = lane & 31;
  const _Float16 *Ag = A + (size_t) tg * P::KPT * 2 * DP * DP;
  half8 v;

  unsigned kl, row;

  tile_row<DP> (R, kl, row);

  if (DP >= 16)
  {
    row += (DP == 64) ? rt * 32 : 0;
    v = *(const half8 *) (Ag + ((kl * 2 + lohi) * DP * DP + row * DP + kb * 16 + g * 8));
  }
  else if (DP == 8)
  {
    v = *(const half8 *) (Ag + ((kl * 2 + g) * 64 + row * 8));  // g = 0: hi, 1: lo
  }
  else
  {
    const _Float16 *bh = Ag + (kl * 2 * 16 + row * 4);
    const half4 h      = *(const half4 *) bh;
    const half4 l      = *(const half4 *) (bh + 16);

    v = (half8) {h[0], h[1], h[2], h[3], l[0], l[1], l[2], l[3]};
  }

  return v;
}

template <int DP>
__device__ inline half8
load_b_frag (const _Float16 *__restrict__ Yh, const _Float16 *__restrict__ Yl, int p, int kb, int lohi, int lane)
{
  const int g = lane >> 5;
  half8 v;

  if (DP >= 16)
  {
    const _Float16 *src = (lohi ? Yl : Yh) + (size_t) p * DP + kb * 16 + g * 8;

    v = *(const half8 *) src;
  }
  else if (DP == 8)
  {
    v = *(const half8 *) ((lohi ? Yl : Yh) + (size_t) p * 8);
  }
  else
  {
    const half4 q = *(const half4 *) ((g ? Yl : Yh) + (size_t) p * 4);

    v = (half8) {q[0], q[1], q[2], q[3], q[0], q[1], q[2], q[3]};
  }

  return v;
}

// Pair reduction across the half-waves with one v_permlane32_swap (it swaps
// lanes 32-63 of its first operand with lanes 0-31 of its second).  With
// (a, b) the partial sums of two kernel slots, lanes 0-31 get a[L] + a[L+32]
// and lanes 32-63 get b[L] + b[L-32]: each half-wave receives the total of
// the slot it owns.
__device__ inline float
halfwave_pair_sum (float a, float b)
{
  const auto r = __builtin_amdgcn_permlane32_swap (__float_as_uint (a), __float_as_uint (b), false, false);

  return __uint_as_float (r[0]) + __uint_as_float (r[1]);
}

// All lnt2 <= 0: b2_i is taken relative to max_j b_j and the kernel factor is
// <= 1.  The sum therefore runs against a fixed reference M0 (0 in the main
// pass) with no rescaling; Mb tracks the largest term so that points whose
// terms all fall below 2^-100 can be recomputed against their own maximum.
template <int KERNEL, bool OFFSET>
__device__ inline void
lse_push (float c2, float4 cf, float kappa, float M0, float &Mb, float &S32)
{
  float lnt;

  if (KERNEL == NCKDE_KERNEL_GAUSS)
    lnt = fmaf (cf.x, c2, cf.y);
  else
    lnt = fmaf (kappa, __builtin_amdgcn_logf (fmaf (c2, cf.x, 1.0f)), cf.y);

  // the running maximum as an unsigned minimum of the bit patterns: for
  // floats <= 0 (every lnt here, and Mb from its -inf start) the larger
  // value has the smaller pattern (+0 is 0, -inf is 0xff800000).  One
  // v_min_u32; fmaxf costs two more v_max_f32 that quiet signalling NaNs.
  Mb = __uint_as_float (min (__float_as_uint (Mb), __float_as_uint (lnt)));

  if (OFFSET)
    lnt -= M0;

  S32 += __builtin_amdgcn_exp2f (lnt);
}

template <int DP>
struct GroupData
{
  half8 Ah[Plan<DP>::NRT][Plan<DP>::NKB];
  half8 Al[Plan<DP>::NRT][Plan<DP>::NKB];
  floatx16 cinit[Plan<DP>::NRT];
  float4 cf[4];
};

template <int DP>
__device__ inline void
load_group (GroupData<DP> &G, const _Float16 *__restrict__ A, const float *__restrict__ cv,
            const float4 *__restrict__ coef, int tg, int lane)
{
  using P          = Plan<DP>;
  const unsigned g = lane >> 5;
  const float *cg  = cv + (size_t) tg * P::KPT * DP;   // wave-uniform group bases
  const float4 *fg = coef + (size_t) tg * P::KPT;

#pragma unroll
  for (int rt = 0; rt < P::NRT; ++rt)
  {
#pragma unroll
    for (int kb = 0; kb < P::NKB; ++kb)
    {
      G.Ah[rt][kb] = load_a_frag<DP> (A, tg, rt, kb, 0, lane);

      if (DP >= 16)
        G.Al[rt][kb] = load_a_frag<DP> (A, tg, rt, kb, 1, lane);
    }

    // the (negated) centre terms of the 16 rows held by this lane: rows 8 j + 4 g + r
#pragma unroll
    for (int j = 0; j < 4; ++j)
    {
      unsigned kl, row;

      tile_row<DP> (8 * j + 4 * g, kl, row);

      const float4 c4 = *(const float4 *) (cg + kl * DP + ((DP == 64) ? rt * 32 + row : row));

      G.cinit[rt][4 * j + 0] = c4.x;
      G.cinit[rt][4 * j + 1] = c4.y;
      G.cinit[rt][4 * j + 2] = c4.z;
      G.cinit[rt][4 * j + 3] = c4.w;
    }
  }

  // coefficients of the kernel slots this lane owns
#pragma unroll
  for (int q = 0; q < 4; ++q)
  {
    int slot = 0;

    if (DP == 4)
      slot = 2 * q + g;
    else if (DP == 8)
      slot = 2 * g + (q & 1);
    else if (DP == 16)
      slot = g;

    G.cf[q] = fg[slot];
  }
}

// One kernel group against the NPT point tiles of a wave: all MFMA chains
// first (separate accumulators), then the VALU reduction and sums, so the
// matrix cores of tile t+1 overlap the vector work of tile t.
template <int DP, int KERNEL, int NPT, bool OFFSET, bool SCALED, int NST>
__device__ inline void
group_step (const GroupData<DP> &cur, const half8 (&Bh)[NPT][Plan<DP>::NKB], const half8 (&Bl)[NPT][Plan<DP>::NKB],
            const float (&tscale)[NPT], const float (&it2)[NPT], const float (&M0)[NST], float (&Mb)[NST],
            float (&S32)[NST], float kappa, int g)
{
  using P = Plan<DP>;
  floatx16 acc[NPT][P::NRT];

#pragma unroll
  for (int t = 0; t < NPT; ++t)
  {
#pragma unroll
    for (int rt = 0; rt < P::NRT; ++rt)
    {
      floatx16 c0 = cur.cinit[rt];

      if (SCALED)
      {
#pragma unroll
        for (int r = 0; r < 16; ++r)
          c0[r] *= tscale[t];
      }

      if (DP == 4)
      {
        acc[t][rt] = __builtin_amdgcn_mfma_f32_32x32x16_f16 (cur.Ah[rt][0], Bh[t][0], c0, 0, 0, 0);
      }
      else if (DP == 8)
      {
        acc[t][rt] = __builtin_amdgcn_mfma_f32_32x32x16_f16 (cur.Ah[rt][0], Bh[t][0], c0, 0, 0, 0);
        acc[t][rt] = __builtin_amdgcn_mfma_f32_32x32x16_f16 (cur.Ah[rt][0], Bl[t][0], acc[t][rt], 0, 0, 0);
      }
      else
      {
        acc[t][rt] = __builtin_amdgcn_mfma_f32_32x32x16_f16 (cur.Ah[rt][0], Bh[t][0], c0, 0, 0, 0);
        acc[t][rt] = __builtin_amdgcn_mfma_f32_32x32x16_f16 (cur.Ah[rt][0], Bl[t][0], acc[t][rt], 0, 0, 0);
        acc[t][rt] = __builtin_amdgcn_mfma_f32_32x32x16_f16 (cur.Al[rt][0], Bh[t][0], acc[t][rt], 0, 0, 0);

#pragma unroll
        for (int kb = 1; kb < P::NKB; ++kb)
        {
          acc[t][rt] = __builtin_amdgcn_mfma_f32_32x32x16_f16 (cur.Ah[rt][kb], Bh[t][kb], acc[t][rt], 0, 0, 0);
          acc[t][rt] = __builtin_amdgcn_mfma_f32_32x32x16_f16 (cur.Ah[rt][kb], Bl[t][kb], acc[t][rt], 0, 0, 0);
          acc[t][rt] = __builtin_amdgcn_mfma_f32_32x32x16_f16 (cur.Al[rt][kb], Bh[t][kb], acc[t][rt], 0, 0, 0);
        }
      }
    }
  }

  float tot_prev = 0.0f;

#pragma unroll
  for (int t = 0; t < NPT; ++t)
  {
    // squared norms per kernel slot held by this lane
    float part[P::NLOCAL];

#pragma unroll
    for (int j = 0; j < P::NLOCAL; ++j)
      part[j] = 0.0f;

#pragma unroll
    for (int rt = 0; rt < P::NRT; ++rt)
#pragma unroll
      for (int r = 0; r < 16; ++r)
      {
        const int j = (DP == 4) ? r / 4 : ((DP == 8) ? r / 8 : 0);

        part[j] = fmaf (acc[t][rt][r], acc[t][rt][r], part[j]);
      }

    const float sc2 = SCALED ? it2[t] : 1.0f;

    if (DP == 4)
    {
      // each kernel's 4 rows sit in one lane: slots 2 j + g
#pragma unroll
      for (int j = 0; j < 4; ++j)
        lse_push<KERNEL, OFFSET> (part[j] * sc2, cur.cf[j], kappa, M0[t], Mb[t], S32[t]);
    }
    else if (DP == 8)
    {
      // whole kernels per lane: slots 2 g + q
#pragma unroll
      for (int q = 0; q < 2; ++q)
        lse_push<KERNEL, OFFSET> (part[q] * sc2, cur.cf[q], kappa, M0[t], Mb[t], S32[t]);
    }
    else if (DP == 16)
    {
      // one whole kernel per lane: slot g
      lse_push<KERNEL, OFFSET> (part[0] * sc2, cur.cf[0], kappa, M0[t], Mb[t], S32[t]);
    }
    else
    {
      // one kernel per group: the lower half keeps even point tiles, the
      // upper half odd ones; state q of a lane is its tile 2 q + g
      const float tot = halfwave_pair_sum (part[0], part[0]) * sc2;

      if (t & 1)
      {
        const float mine = g ? tot : tot_prev;

        lse_push<KERNEL, OFFSET> (mine, cur.cf[0], kappa, M0[t / 2], Mb[t / 2], S32[t / 2]);
      }
      else
      {
        tot_prev = tot;
      }
    }
  }
}

// Tiling chosen by measurement on MI355X: up to DP = 16 a wave keeps more
// point tiles and loads each kernel group just before its MFMA chains (the
// occupancy of the smaller register set hides the loads).  At DP = 32 the
// next kernel is prefetched into a second register set.  At DP = 64 one
// kernel's factors (16 KiB) feed all four waves of a block: they are staged
// once per block in LDS, double-buffered, the next kernel's global loads in
// flight while the current one is computed (1.5x faster than the register
// prefetch at d = 64; at d = 32 the barrier per kernel costs more than the
// shared loads save).
constexpr int
tiles_per_wave (int DP)
{
  return (DP == 64) ? 2 : ((DP == 16) ? 8 : 4);
}

constexpr bool
lds_groups (int DP)
{
  return DP == 64;
}

// waves per block: at DP = 64 eight waves share each staged kernel
constexpr int
waves_per_block (int DP)
{
  return (DP == 64) ? 8 : 4;
}

// LDS stage of one kernel (DP = 64): A hi rows then lo rows, each padded to
// DP + 8 halves (16 B) so that the 32 rows read by a fragment load spread
// over the banks; then c (DP floats) and the coefficients.
template <int DP>
struct Stage
{
  static constexpr int RS  = DP + 8;                  // row stride, halves
  static constexpr int ASZ = 2 * DP * RS;             // halves per stage
  static constexpr int NT  = 64 * waves_per_block (DP);   // threads per block
  static constexpr int CPT = 2 * DP * DP / 8 / NT;        // 16 B chunks per thread
};

typedef unsigned int u32x4 __attribute__ ((ext_vector_type (4)));
typedef float f32x4 __attribute__ ((ext_vector_type (4)));

// staging registers (plain vector types, so they stay in registers)
template <int DP>
struct StageFetch
{
  u32x4 a[(Stage<DP>::CPT > 0) ? Stage<DP>::CPT : 1];
  f32x4 c, f;
};

template <int DP>
__device__ inline void
stage_load (StageFetch<DP> &s, const _Float16 *__restrict__ A, const float *__restrict__ cv,
            const float4 *__restrict__ coef, int tg, int tid)
{
  const u32x4 *src = (const u32x4 *) (A + (size_t) tg * 2 * DP * DP);

#pragma unroll
  for (int q = 0; q < Stage<DP>::CPT; ++q)
    s.a[q] = src[tid + q * Stage<DP>::NT];

  // every thread loads (in bounds) so the stage stays in registers
  s.c = ((const f32x4 *) (cv + (size_t) tg * DP))[tid % (DP / 4)];
  s.f = ((const f32x4 *) coef)[tg];
}

template <int DP>
__device__ inline void
stage_store (const StageFetch<DP> &s, _Float16 *sA, float *sC, float4 *sF, int tid)
{
  constexpr int CPR = DP / 8;   // chunks per row

#pragma unroll
  for (int q = 0; q < Stage<DP>::CPT; ++q)
  {
    const int e = tid + q * Stage<DP>::NT;

    *(u32x4 *) (sA + (e / CPR) * Stage<DP>::RS + (e % CPR) * 8) = s.a[q];
  }

  if (tid < DP / 4)
    ((f32x4 *) sC)[tid] = s.c;

  if (tid == 0)
    *(f32x4 *) sF = s.f;
}

// the fragments of a staged kernel, as load_group reads them from global memory
template <int DP>
__device__ inline void
load_group_lds (GroupData<DP> &G, const _Float16 *sA, const float *sC, const float4 *sF, int lane)
{
  using P            = Plan<DP>;
  constexpr int RS   = Stage<DP>::RS;
  const unsigned g   = lane >> 5;
  const unsigned R   = lane & 31;

#pragma unroll
  for (int rt = 0; rt < P::NRT; ++rt)
  {
    const unsigned row = (DP == 64) ? rt * 32 + R : R;

#pragma unroll
    for (int kb = 0; kb < P::NKB; ++kb)
    {
      G.Ah[rt][kb] = *(const half8 *) (sA + row * RS + kb * 16 + g * 8);
      G.Al[rt][kb] = *(const half8 *) (sA + (DP + row) * RS + kb * 16 + g * 8);
    }

#pragma unroll
    for (int j = 0; j < 4; ++j)
    {
      const unsigned Rr = 8 * j + 4 * g;
      const float4 c4   = *(const float4 *) (sC + ((DP == 64) ? rt * 32 + Rr : Rr));

      G.cinit[rt][4 * j + 0] = c4.x;
      G.cinit[rt][4 * j + 1] = c4.y;
      G.cinit[rt][4 * j + 2] = c4.z;
      G.cinit[rt][4 * j + 3] = c4.w;
    }
  }

#pragma unroll
  for (int q = 0; q < 4; ++q)
    G.cf[q] = *sF;
}

template <int DP, int KERNEL, int NPT, bool OFFSET, bool SCALED, int NST>
__device__ inline void
group_loop (int g0, int g1, const _Float16 *__restrict__ A, const float *__restrict__ cv,
            const float4 *__restrict__ coef, const half8 (&Bh)[NPT][Plan<DP>::NKB],
            const half8 (&Bl)[NPT][Plan<DP>::NKB], const float (&tscale)[NPT], const float (&it2)[NPT],
            const float (&M0)[NST], float (&Mb)[NST], double (&S64)[NST], float kappa, int lane,
            _Float16 *sA, float *sC, float4 *sF)
{
  const int g = lane >> 5;
  float S32[NST];

#pragma unroll
  for (int q = 0; q < NST; ++q)
    S32[q] = 0.0f;

  if constexpr (DP <= 16)
  {
    for (int tg = g0; tg < g1; ++tg)
    {
      GroupData<DP> gd;

      load_group<DP> (gd, A, cv, coef, tg, lane);
      group_step<DP, KERNEL, NPT, OFFSET, SCALED, NST> (gd, Bh, Bl, tscale, it2, M0, Mb, S32, kappa, g);

      // fold the FP32 partial sums into FP64 every 16 groups
      if (((tg + 1 - g0) & 15) == 0)
      {
#pragma unroll
        for (int q = 0; q < NST; ++q)
        {
          S64[q] += (double) S32[q];
          S32[q] = 0.0f;
        }
      }
    }

#pragma unroll
    for (int q = 0; q < NST; ++q)
      S64[q] += (double) S32[q];
  }
  else if constexpr (!lds_groups (DP))
  {
    // two register sets alternate roles: while one group is computed the
    // next one is loaded into the other set (no register copies)
    GroupData<DP> ga, gb;
    int tg = g0;

    if (tg < g1)
      load_group<DP> (ga, A, cv, coef, tg, lane);

    while (tg < g1)
    {
      if (tg + 1 < g1)
        load_group<DP> (gb, A, cv, coef, tg + 1, lane);

      group_step<DP, KERNEL, NPT, OFFSET, SCALED, NST> (ga, Bh, Bl, tscale, it2, M0, Mb, S32, kappa, g);
      ++tg;

      if (tg >= g1)
        break;

      if (tg + 1 < g1)
        load_group<DP> (ga, A, cv, coef, tg + 1, lane);

      group_step<DP, KERNEL, NPT, OFFSET, SCALED, NST> (gb, Bh, Bl, tscale, it2, M0, Mb, S32, kappa, g);
      ++tg;

      // fold the FP32 partial sums into FP64 every 16 groups
      if (((tg - g0) & 15) == 0)
      {
#pragma unroll
        for (int q = 0; q < NST; ++q)
        {
          S64[q] += (double) S32[q];
          S32[q] = 0.0f;
        }
      }
    }

#pragma unroll
    for (int q = 0; q < NST; ++q)
      S64[q] += (double) S32[q];
  }
  else
  {
  // staged groups: g0 and g1 are block-uniform, so every wave of the block
  // runs the same iterations and meets the same barriers.  Stage st is read
  // in iteration tg while the other one is written; the barrier at the end of
  // each iteration orders those reads before the next overwrite.
  const int tid = threadIdx.x;
  StageFetch<DP> nx;

  if (g0 < g1)
  {
    stage_load<DP> (nx, A, cv, coef, g0, tid);
    stage_store<DP> (nx, sA, sC, sF, tid);
  }

  __syncthreads ();

  for (int tg = g0; tg < g1; ++tg)
  {
    const int st = (tg - g0) & 1;

    if (tg + 1 < g1)
      stage_load<DP> (nx, A, cv, coef, tg + 1, tid);

    {
      GroupData<DP> gd;

      load_group_lds<DP> (gd, sA + st * Stage<DP>::ASZ, sC + st * DP, sF + st, lane);
      group_step<DP, KERNEL, NPT, OFFSET, SCALED, NST> (gd, Bh, Bl, tscale, it2, M0, Mb, S32, kappa, g);
    }

    if (tg + 1 < g1)
      stage_store<DP> (nx, sA + (st ^ 1) * Stage<DP>::ASZ, sC + (st ^ 1) * DP, sF + (st ^ 1), tid);

    __syncthreads ();

    // fold the FP32 partial sums into FP64 every 16 groups
    if (((tg + 1 - g0) & 15) == 0)
    {
#pragma unroll
      for (int q = 0; q < NST; ++q)
      {
        S64[q] += (double) S32[q];
        S32[q] = 0.0f;
      }
    }
  }

#pragma unroll
  for (int q = 0; q < NST; ++q)
    S64[q] += (double) S32[q];
  }
}

// The evaluation of one block: block (bx, by) takes the point tiles of
// column bx and the kernel groups of split by.
template <int DP, int KERNEL, int NPT, bool OFFSET>
__device__ inline void
eval_mixed_block (const MixedArgs &a, int bx, int by)
{
  using P        = Plan<DP>;
  constexpr int NST = (DP >= 32) ? NPT / 2 : NPT;   // LSE states per lane
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int g    = lane >> 5;
  const int R    = lane & 31;

  const _Float16 *Yh = (const _Float16 *) a.Yh;
  const _Float16 *Yl = Yh + (size_t) a.m * DP;
  const float2 *tpv  = (const float2 *) (Yl + (size_t) a.m * DP);
  const _Float16 *A  = (const _Float16 *) a.f.A;
  const float *cv    = a.f.c;
  const float4 *coef = (const float4 *) a.f.coef;
  const float kappa  = (float) a.kappa;

  const int tile0   = (bx * waves_per_block (DP) + wave) * NPT;
  const int ngroups = (a.n + P::KPT - 1) / P::KPT;
  const int chunk   = (ngroups + a.nsplit - 1) / a.nsplit;
  const int g0      = by * chunk;
  const int g1      = min (ngroups, g0 + chunk);

  // second pass: only waves holding a flagged point do any work
  float M0[NST];

#pragma unroll
  for (int q = 0; q < NST; ++q)
    M0[q] = 0.0f;

  if (OFFSET)
  {
    int any = 0;

#pragma unroll
    for (int t = 0; t < NPT; ++t)
    {
      const int p = (tile0 + t) * 32 + R;

      if (p < a.m)
        any |= a.flag[p];
    }

    // staged kernels keep the whole block (its barriers need every wave)
    if (lds_groups (DP) ? !__syncthreads_or (any) : !__any (any))
      return;

#pragma unroll
    for (int q = 0; q < NST; ++q)
    {
      const int t = (DP >= 32) ? 2 * q + g : q;
      const int p = min ((tile0 + t) * 32 + R, a.m - 1);

      M0[q] = a.M0[p];
    }
  }

  // B fragments (points) and per-point scales
  half8 Bh[NPT][P::NKB], Bl[NPT][P::NKB];
  float tscale[NPT], it2[NPT];
  bool scaled = false;

#pragma unroll
  for (int t = 0; t < NPT; ++t)
  {
    int p = (tile0 + t) * 32 + R;

    p = (p < a.m) ? p : a.m - 1;

#pragma unroll
    for (int kb = 0; kb < P::NKB; ++kb)
    {
      Bh[t][kb] = load_b_frag<DP> (Yh, Yl, p, kb, 0, lane);
      Bl[t][kb] = load_b_frag<DP> (Yh, Yl, p, kb, 1, lane);
    }

    const float2 tv = tpv[p];

    tscale[t] = tv.x;
    it2[t]    = tv.y;
    scaled    = scaled || (tv.x != 1.0f);
  }

  float Mb[NST];
  double S64[NST];

#pragma unroll
  for (int q = 0; q < NST; ++q)
  {
    Mb[q]  = -INFINITY;
    S64[q] = 0.0;
  }

  // LDS stages of the factors (DP >= 32), two per block
  __shared__ __attribute__ ((aligned (16))) _Float16 sA[lds_groups (DP) ? 2 * Stage<DP>::ASZ : 8];
  __shared__ __attribute__ ((aligned (16))) float sC[lds_groups (DP) ? 2 * DP : 4];
  __shared__ float4 sF[2];

  // the scaled/unscaled choice is block-wide where the loop has barriers
  const bool any_scaled = lds_groups (DP) ? (bool) __syncthreads_or (scaled) : (bool) __any (scaled);

  if (any_scaled)
    group_loop<DP, KERNEL, NPT, OFFSET, true, NST> (g0, g1, A, cv, coef, Bh, Bl, tscale, it2, M0, Mb, S64, kappa,
                                                    lane, sA, sC, sF);
  else
    group_loop<DP, KERNEL, NPT, OFFSET, false, NST> (g0, g1, A, cv, coef, Bh, Bl, tscale, it2, M0, Mb, S64, kappa,
                                                     lane, sA, sC, sF);

  if (DP >= 32)
  {
#pragma unroll
    for (int q = 0; q < NST; ++q)
    {
      const int p = (tile0 + 2 * q + g) * 32 + R;

      if (p < a.m)
      {
        a.partM[(size_t) by * a.m + p] = Mb[q];
        a.partS[(size_t) by * a.m + p] = S64[q];
      }
    }

    return;
  }

  // the two half-waves hold the same points (disjoint kernel slots) and the
  // same reference: add the sums, take the larger maximum
#pragma unroll
  for (int t = 0; t < NPT; ++t)
  {
    const float Mo  = __shfl_xor (Mb[t], 32, 64);
    const double So = __shfl_xor (S64[t], 32, 64);
    const int p     = (tile0 + t) * 32 + R;

    if (g == 0 && p < a.m)
    {
      a.partM[(size_t) by * a.m + p] = fmaxf (Mb[t], Mo);
      a.partS[(size_t) by * a.m + p] = S64[t] + So;
    }
  }
}

template <int DP, int KERNEL, int NPT, bool OFFSET>
__global__ void __launch_bounds__ (64 * waves_per_block (DP))
k_eval_mixed (MixedArgs a)
{
  eval_mixed_block<DP, KERNEL, NPT, OFFSET> (a, blockIdx.x, blockIdx.y);
}

// Several evaluations in one launch (sessions with the same DP and kernel
// type): segment s owns the blocks [block0[s], block0[s+1]) of the flattened
// grid, laid out as its own (column, split) grid, so every block computes
// exactly what it computes in the segment's own launch.
__device__ inline int
segment_of (const int *__restrict__ block0, int nseg, int b)
{
  int lo = 0, hi = nseg - 1;

  while (lo < hi)
  {
    const int mid = (lo + hi + 1) / 2;

    if (block0[mid] <= b)
      lo = mid;
    else
      hi = mid - 1;
  }

  return lo;
}

template <int DP, int KERNEL, int NPT, bool OFFSET>
__global__ void __launch_bounds__ (64 * waves_per_block (DP))
k_eval_mixed_seg (const MixedArgs *__restrict__ segs, const int *__restrict__ block0, int nseg)
{
  const int sg      = segment_of (block0, nseg, blockIdx.x);
  const MixedArgs a = segs[sg];
  const int pts     = waves_per_block (DP) * NPT * 32;
  const int cols    = (a.m + pts - 1) / pts;
  const int local   = blockIdx.x - block0[sg];

  eval_mixed_block<DP, KERNEL, NPT, OFFSET> (a, local % cols, local / cols);
}

// Below 2^-100 the FP32 terms lose range: such points are recomputed against
// their own largest term.  Against a pilot reference the largest term may
// also lie up to 2^96 above it (the FP32 lane sums stay finite).
constexpr float MIXED_FLOOR = -100.0f;
constexpr float MIXED_CEIL  = 96.0f;

// Merge of the nsplit partial (max, sum) pairs of every point: a block takes
// 32 points with 8 lanes each; lane l sums the splits s = l, l+8, ... in
// order (reads of consecutive points are coalesced), then the 8 lane
// partials combine in lane order, so the result depends only on nsplit.
// log sum = b_ref + ln2 (M0 + log2 S), M0 the terms' reference (0 for
// MERGE_PLAIN); mode is a MixedMerge.
constexpr int MG_PTS = 32, MG_LANES = 8;

__device__ inline void
mixed_merge_block (const MixedArgs &a, double bref, int mode, int bx)
{
  __shared__ float Ms[MG_LANES][MG_PTS];
  __shared__ double Ss[MG_LANES][MG_PTS];
  const int pl  = threadIdx.x % MG_PTS;
  const int sl  = threadIdx.x / MG_PTS;
  const int p   = bx * MG_PTS + pl;
  const bool on = (p < a.m) && !(mode == MERGE_FLAGGED && !a.flag[p]);
  const int ns  = (mode == MERGE_PILOT) ? a.pnsplit : a.nsplit;
  float Mx      = -INFINITY;
  double S      = 0.0;

  if (on)
    for (int s = sl; s < ns; s += MG_LANES)
    {
      Mx = fmaxf (Mx, a.partM[(size_t) s * a.m + p]);
      S += a.partS[(size_t) s * a.m + p];
    }

  Ms[sl][pl] = Mx;
  Ss[sl][pl] = S;
  __syncthreads ();

  if (sl != 0 || !on)
    return;

  for (int l = 1; l < MG_LANES; ++l)
  {
    Mx = fmaxf (Mx, Ms[l][pl]);
    S += Ss[l][pl];
  }

  const float Mref = (Mx > -INFINITY) ? Mx : 0.0f;

  if (mode == MERGE_PILOT)
  {
    a.flag[p] = 1;
    a.M0[p]   = Mref;

    return;
  }

  const float M0 = (mode == MERGE_PLAIN) ? 0.0f : a.M0[p];

  if (mode == MERGE_PLAIN || mode == MERGE_CHECK)
  {
    const int bad = (Mx - M0 < MIXED_FLOOR) || (Mx - M0 > MIXED_CEIL) || !(S > 0.0) || !(S <= DBL_MAX);

    a.flag[p] = bad;

    if (bad || mode == MERGE_PLAIN)
      a.M0[p] = Mref;
  }

  const double lnsum = bref + LN2 * ((double) M0 + log2 (S));
  const double m2lnp = -2.0 * ((lnsum - a.lnnorm_shared) - a.lnh_d);

  a.out[p] = a.linear ? exp (-0.5 * m2lnp) : m2lnp;
}

__global__ void __launch_bounds__ (MG_PTS * MG_LANES)
k_mixed_merge (MixedArgs a, double bref, int mode)
{
  mixed_merge_block (a, bref, mode, blockIdx.x);
}

// the merges of several evaluations in one launch (segment s owns the point
// blocks [block0[s], block0[s+1])); each uses its own b_ref
__global__ void __launch_bounds__ (MG_PTS * MG_LANES)
k_mixed_merge_seg (const MixedArgs *__restrict__ segs, const int *__restrict__ block0, int nseg, int mode)
{
  const int sg      = segment_of (block0, nseg, blockIdx.x);
  const MixedArgs a = segs[sg];

  mixed_merge_block (a, a.bref, mode, blockIdx.x - block0[sg]);
}

template <int DP, int KERNEL, bool OFFSET>
static hipError_t
eval_t (const MixedArgs &a, hipStream_t s)
{
  constexpr int NPT = tiles_per_wave (DP);
  const int pts     = waves_per_block (DP) * NPT * 32;
  dim3 grid ((a.m + pts - 1) / pts, a.nsplit);

  hipLaunchKernelGGL ((k_eval_mixed<DP, KERNEL, NPT, OFFSET>), grid, dim3 (64 * waves_per_block (DP)), 0, s, a);

  return hipGetLastError ();
}

// Without a pilot: the main pass against b_ref, then the points whose terms
// underflowed again against their own largest term.  With one: the pilot
// pass gives every point a reference, the main pass runs against it, and
// the few points whose largest term lies too far from it run again.
template <int DP, int KERNEL>
static hipError_t
eval_dk (const MixedArgs &a, hipStream_t s)
{
  hipError_t e = hipSuccess;
  const bool pilot = (a.pn > 0);

  if (pilot)
  {
    MixedArgs ap = a;

    ap.f      = a.pf;
    ap.n      = a.pn;
    ap.nsplit = a.pnsplit;
    e         = eval_t<DP, KERNEL, false> (ap, s);

    if (e == hipSuccess)
      e = launch_mixed_merge (a, a.bref, MERGE_PILOT, s);
  }

  if (a.ev0)
    (void) hipEventRecord (a.ev0, s);

  if (e == hipSuccess)
    e = pilot ? eval_t<DP, KERNEL, true> (a, s) : eval_t<DP, KERNEL, false> (a, s);

  if (a.ev1)
    (void) hipEventRecord (a.ev1, s);

  if (e == hipSuccess)
    e = launch_mixed_merge (a, a.bref, pilot ? MERGE_CHECK : MERGE_PLAIN, s);

  if (e == hipSuccess)
    e = eval_t<DP, KERNEL, true> (a, s);

  if (e == hipSuccess)
    e = launch_mixed_merge (a, a.bref, MERGE_FLAGGED, s);

  return e;
}

template <int DP>
static hipError_t
eval_d (const MixedArgs &a, hipStream_t s)
{
  return (a.kernel == NCKDE_KERNEL_GAUSS) ? eval_dk<DP, NCKDE_KERNEL_GAUSS> (a, s)
                                          : eval_dk<DP, NCKDE_KERNEL_ST> (a, s);
}

template <int DP, int KERNEL>
static hipError_t
eval_multi_dk (const MixedArgs *dsegs, const int *dblock0, int nseg, int eval_blocks, int merge_blocks,
               hipStream_t s)
{
  constexpr int NPT = tiles_per_wave (DP);
  const dim3 eb (64 * waves_per_block (DP)), mb (MG_PTS * MG_LANES);

  hipLaunchKernelGGL ((k_eval_mixed_seg<DP, KERNEL, NPT, false>), dim3 (eval_blocks), eb, 0, s, dsegs, dblock0,
                      nseg);
  hipLaunchKernelGGL (k_mixed_merge_seg, dim3 (merge_blocks), mb, 0, s, dsegs, dblock0 + nseg + 1, nseg,
                      (int) MERGE_PLAIN);
  hipLaunchKernelGGL ((k_eval_mixed_seg<DP, KERNEL, NPT, true>), dim3 (eval_blocks), eb, 0, s, dsegs, dblock0,
                      nseg);
  hipLaunchKernelGGL (k_mixed_merge_seg, dim3 (merge_blocks), mb, 0, s, dsegs, dblock0 + nseg + 1, nseg,
                      (int) MERGE_FLAGGED);

  return hipGetLastError ();
}

int
mixed_eval_blocks (const MixedArgs &a)
{
  const int pts = mixed_points_per_block (a.f.DP);

  return ((a.m + pts - 1) / pts) * a.nsplit;
}

int
mixed_merge_blocks (const MixedArgs &a)
{
  return (a.m + MG_PTS - 1) / MG_PTS;
}

hipError_t
launch_eval_mixed_multi (int DP, int kernel, const MixedArgs *dsegs, const int *dblock0, int nseg, int eval_blocks,
                         int merge_blocks, hipStream_t s)
{
  const bool st = (kernel == NCKDE_KERNEL_ST);

  switch (DP)
  {
    case 4: return st ? eval_multi_dk<4, NCKDE_KERNEL_ST> (dsegs, dblock0, nseg, eval_blocks, merge_blocks, s)
                      : eval_multi_dk<4, NCKDE_KERNEL_GAUSS> (dsegs, dblock0, nseg, eval_blocks, merge_blocks, s);
    case 8: return st ? eval_multi_dk<8, NCKDE_KERNEL_ST> (dsegs, dblock0, nseg, eval_blocks, merge_blocks, s)
                      : eval_multi_dk<8, NCKDE_KERNEL_GAUSS> (dsegs, dblock0, nseg, eval_blocks, merge_blocks, s);
    case 16: return st ? eval_multi_dk<16, NCKDE_KERNEL_ST> (dsegs, dblock0, nseg, eval_blocks, merge_blocks, s)
                       : eval_multi_dk<16, NCKDE_KERNEL_GAUSS> (dsegs, dblock0, nseg, eval_blocks, merge_blocks, s);
    case 32: return st ? eval_multi_dk<32, NCKDE_KERNEL_ST> (dsegs, dblock0, nseg, eval_blocks, merge_blocks, s)
                       : eval_multi_dk<32, NCKDE_KERNEL_GAUSS> (dsegs, dblock0, nseg, eval_blocks, merge_blocks, s);
    case 64: return st ? eval_multi_dk<64, NCKDE_KERNEL_ST> (dsegs, dblock0, nseg, eval_blocks, merge_blocks, s)
                       : eval_multi_dk<64, NCKDE_KERNEL_GAUSS> (dsegs, dblock0, nseg, eval_blocks, merge_blocks, s);
    default: return hipErrorInvalidValue;
  }
}

int
mixed_points_per_block (int DP)
{
  return waves_per_block (DP) * tiles_per_wave (DP) * 32;
}

hipError_t
launch_eval_mixed (const MixedArgs &a, hipStream_t s)
{
  switch (a.f.DP)
  {
    case 4: return eval_d<4> (a, s);
    case 8: return eval_d<8> (a, s);
    case 16: return eval_d<16> (a, s);
    case 32: return eval_d<32> (a, s);
    case 64: return eval_d<64> (a, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t
launch_mixed_merge (const MixedArgs &a, double bref, int mode, hipStream_t s)
{
  hipLaunchKernelGGL (k_mixed_merge, dim3 ((a.m + MG_PTS - 1) / MG_PTS), dim3 (MG_PTS * MG_LANES), 0, s, a, bref,
                      mode);

  return hipGetLastError ();
}

// factor build pipeline (called by the host at the first mixed eval after a
// prepare, and when the weights change)
hipError_t
mixed_build (const double *X, int n, int d, int D, int DP, const double *Wp, const double *Ug, const double *b,
             double bref, double inv_h, int kernel, double nu, double *mu, double *Yc,
             unsigned long long *ymax_bits, double *ysc_out, _Float16 *A, float *c, float *coef, float *risk,
             hipStream_t s)
{
  hipError_t e;

  hipLaunchKernelGGL (k_colmean, dim3 (d), dim3 (256), 0, s, X, n, d, mu);
  e = hipMemsetAsync (ymax_bits, 0, sizeof (unsigned long long), s);

  if (e != hipSuccess)
    return e;

  hipLaunchKernelGGL (k_mixed_centres, dim3 ((n + 127) / 128), dim3 (128), 0, s, X, n, d, mu, Ug, Yc, ymax_bits);

  unsigned long long bits = 0;

  e = hipMemcpyAsync (&bits, ymax_bits, sizeof (bits), hipMemcpyDeviceToHost, s);

  if (e != hipSuccess)
    return e;

  e = hipStreamSynchronize (s);

  if (e != hipSuccess)
    return e;

  double ymax;

  memcpy (&ymax, &bits, sizeof (double));

  // the sample's largest |y'| maps to about 2^11: eval points up to 8x
  // farther out need no per-point rescaling
  const double ysc = (ymax > 0.0) ? ldexp (1.0, 11 - ilogb (ymax) - 1) : 1.0;
  const int kpt    = (DP >= 32) ? 1 : 32 / DP;
  const int n_pad  = ((n + kpt - 1) / kpt) * kpt;

  *ysc_out = ysc;
  hipLaunchKernelGGL (k_mixed_factors, dim3 ((n_pad + 3) / 4), dim3 (256), 0, s, Wp, D, Yc, Ug, b, bref, n, n_pad,
                      d, DP, inv_h, ysc, kernel, nu, A, c, (float4 *) coef, risk);

  return hipGetLastError ();
}

hipError_t
mixed_refresh_b (const double *b, double bref, int n, float *coef, hipStream_t s)
{
  hipLaunchKernelGGL (k_mixed_coef_b, dim3 ((n + 255) / 256), dim3 (256), 0, s, b, bref, n, (float4 *) coef);

  return hipGetLastError ();
}

} // namespace nck
