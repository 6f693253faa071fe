// cv_min.cpp -- host-side pieces of the cross-validated bandwidth fits.
//
// * nm_minimize: the Nelder-Mead simplex that the reference drives through
//   gsl_multimin_fminimizer_nmsimplex2 (ncm_stats_dist.c:175, 664-701).
//   GSL 2.8 is not part of the reference tree; this restates the published
//   nmsimplex2 algorithm (multimin/simplex2.c of GSL 2.8): corner moves
//   x' = a c + b x_h with a = (1-t) P/(P-1), b = (P t - 1)/(P-1) for
//   t = -1 (reflection), -2 (expansion), 1/2 (contraction), contraction of
//   the whole simplex towards the best corner, and the size
//   sqrt(sum |x_i - c|^2 / P) kept incrementally between moves.
// * lm_dif1: the forward-difference Levenberg-Marquardt of levmar 2.6
//   (external/levmar/lm_core.c:439-850 with misc_core.c:137-171), for one
//   parameter and a zero measurement vector, as called by
//   ncm_stats_dist.c:1063-1065.  The 1x1 augmented normal equation is
//   solved by division; the Jacobian is refreshed by finite differences
//   when the damping grows past 16 after an accepted step or after 10
//   Broyden rank-one updates.
// * pair_stats_update: the weighted Welford / West update of NcmStatsVec
//   for two variables (ncm_stats_vec.c:510-565), unit weights.
// * host_normal: one standard normal from the counter-based generator
//   (Philox4x32-10, Salmon et al., SC'11) used for the random restarts of
//   the split fit.
#include "nckde_internal.h"

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <vector>

namespace nck
{

// ---- Nelder-Mead ----------------------------------------------------------

namespace
{

struct Simplex
{
  int n, P;
  std::vector<double> x1, y1, c;
  double S2 = 0.0;

  double *row (int i) { return x1.data () + (size_t) i * n; }

  void
  center ()
  {
    for (int k = 0; k < n; ++k)
      c[k] = 0.0;

    for (int i = 0; i < P; ++i)
      for (int k = 0; k < n; ++k)
        c[k] = c[k] + row (i)[k];

    for (int k = 0; k < n; ++k)
      c[k] = c[k] * (1.0 / P);
  }

  double
  size ()
  {
    double ss = 0.0;

    for (int i = 0; i < P; ++i)
    {
      double t2 = 0.0;

      for (int k = 0; k < n; ++k)
      {
        const double t = row (i)[k] - c[k];

        t2 += t * t;
      }

      ss += t2;
    }

    S2 = ss / P;

    return std::sqrt (ss / P);
  }

  // x' = a c + b x_corner
  void
  corner (double coeff, int i, std::vector<double> &xc)
  {
    const double a = (1.0 - coeff) * P / (P - 1.0);
    const double b = (P * coeff - 1.0) / (P - 1.0);

    for (int k = 0; k < n; ++k)
      xc[k] = a * c[k] + b * row (i)[k];
  }

  void
  update (int i, const std::vector<double> &x, double val)
  {
    double dd = 0.0, xmcd = 0.0;

    for (int k = 0; k < n; ++k)
    {
      const double delta = x[k] - row (i)[k];
      const double xmc   = row (i)[k] - c[k];

      dd += delta * delta;
      xmcd += xmc * delta;
    }

    S2 += (2.0 / P) * xmcd + ((P - 1.0) / P) * (dd / P);

    for (int k = 0; k < n; ++k)
    {
      c[k] = c[k] - row (i)[k] / P;
      c[k] = c[k] + x[k] / P;
      row (i)[k] = x[k];
    }

    y1[i] = val;
  }
};

} // namespace

int
nm_minimize (const ObjFn &f, std::vector<double> &x, const std::vector<double> &step, double epsabs, int maxiter,
             NMInfo *info)
{
  const int n = (int) x.size ();
  Simplex s;

  s.n = n;
  s.P = n + 1;
  s.x1.assign ((size_t) s.P * n, 0.0);
  s.y1.assign (s.P, 0.0);
  s.c.assign (n, 0.0);

  std::vector<double> xc (n), xc2 (n);
  double val, val2;

  // set: the start point, then one step along each axis
  TRY_OK (f (x.data (), val));

  if (!std::isfinite (val))
  {
    set_error ("nm_minimize: non-finite function value at the start point");
    return NCKDE_ENUMERIC;
  }

  for (int k = 0; k < n; ++k)
    s.row (0)[k] = x[k];

  s.y1[0] = val;

  for (int i = 0; i < n; ++i)
  {
    xc = x;
    xc[i] = x[i] + step[i];
    TRY_OK (f (xc.data (), val));

    if (!std::isfinite (val))
    {
      set_error ("nm_minimize: non-finite function value at a start corner");
      return NCKDE_ENUMERIC;
    }

    for (int k = 0; k < n; ++k)
      s.row (i + 1)[k] = xc[k];

    s.y1[i + 1] = val;
  }

  s.center ();

  double size = s.size ();
  double fval = s.y1[0];
  int iter = 0, status = 1;

  do
  {
    iter++;

    // highest, second highest and lowest corners
    int hi = 0, lo = 0, s_hi = 1;
    double dhi = s.y1[0], dlo = s.y1[0], ds_hi = s.y1[1];

    for (int i = 1; i < s.P; ++i)
    {
      const double v = s.y1[i];

      if (v < dlo)
      {
        dlo = v;
        lo  = i;
      }
      else if (v > dhi)
      {
        ds_hi = dhi;
        s_hi  = hi;
        dhi   = v;
        hi    = i;
      }
      else if (v > ds_hi)
      {
        ds_hi = v;
        s_hi  = i;
      }
    }

    bool failed = false;

    s.corner (-1.0, hi, xc);
    TRY_OK (f (xc.data (), val));

    if (std::isfinite (val) && val < s.y1[lo])
    {
      s.corner (-2.0, hi, xc2);
      TRY_OK (f (xc2.data (), val2));

      if (std::isfinite (val2) && val2 < s.y1[lo])
        s.update (hi, xc2, val2);
      else
        s.update (hi, xc, val);
    }
    else if (!std::isfinite (val) || val > s.y1[s_hi])
    {
      if (std::isfinite (val) && val <= s.y1[hi])
        s.update (hi, xc, val);

      s.corner (0.5, hi, xc2);
      TRY_OK (f (xc2.data (), val2));

      if (std::isfinite (val2) && val2 <= s.y1[hi])
      {
        s.update (hi, xc2, val2);
      }
      else
      {
        // contract every corner halfway towards the best one
        for (int i = 0; i < s.P; ++i)
        {
          if (i == lo)
            continue;

          for (int k = 0; k < n; ++k)
            s.row (i)[k] = 0.5 * (s.row (i)[k] + s.row (lo)[k]);

          for (int k = 0; k < n; ++k)
            xc[k] = s.row (i)[k];

          TRY_OK (f (xc.data (), val));
          s.y1[i] = val;

          if (!std::isfinite (val))
            failed = true;
        }

        s.center ();
        s.size ();
      }
    }
    else
    {
      s.update (hi, xc, val);
    }

    if (failed)
    {
      status = -1;
      break;
    }

    // the first lowest corner is the current estimate
    lo = 0;

    for (int i = 1; i < s.P; ++i)
      if (s.y1[i] < s.y1[lo])
        lo = i;

    for (int k = 0; k < n; ++k)
      x[k] = s.row (lo)[k];

    fval = s.y1[lo];
    size = (s.S2 > 0.0) ? std::sqrt (s.S2) : s.size ();
    status = (size < epsabs) ? 0 : 1;
  } while (status == 1 && iter < maxiter);

  if (info)
  {
    info->iter   = iter;
    info->status = status;
    info->fval   = fval;
    info->size   = size;
  }

  return NCKDE_OK;
}

// ---- Levenberg-Marquardt, one parameter ------------------------------------

int
lm_dif1 (const ResFn &func, double &p, int n, int itmax, const double opts[5], LMInfo *info)
{
  const double tau = opts[0], eps1 = opts[1], eps2 = opts[2], eps2_sq = opts[2] * opts[2], eps3 = opts[3];
  const double delta = opts[4];
  const int K        = 10;
  std::vector<double> e (n), hx (n), jac (n), wrk (n), wrk2 (n);
  double mu = 0.0, jacTe_inf = 0.0, p_L2 = 0.0, Dp_L2 = DBL_MAX, jacTjac = 0.0, jacTe = 0.0, diag = 0.0;
  int nu = 20, updjac = 0, updp = 1, newjac = 0, stop = 0, nfev = 0, njap = 0, k;

  TRY_OK (func (p, hx.data ()));
  nfev = 1;

  double p_eL2 = 0.0;

  for (int i = 0; i < n; ++i)
  {
    e[i] = -hx[i];
    p_eL2 += e[i] * e[i];
  }

  const double init_p_eL2 = p_eL2;

  if (!std::isfinite (p_eL2))
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
      // forward difference with step max(1e-4 |p|, delta)
      double dstep = std::fabs (1e-4 * p);

      if (dstep < delta)
        dstep = delta;

      TRY_OK (func (p + dstep, wrk.data ()));

      const double inv = 1.0 / dstep;

      for (int i = 0; i < n; ++i)
        jac[i] = (wrk[i] - hx[i]) * inv;

      ++njap;
      ++nfev;
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

      for (int l = n; l-- > 0;)
      {
        jacTjac += jac[l] * jac[l];
        jacTe += jac[l] * e[l];
      }

      jacTe_inf = std::fabs (jacTe);
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

      TRY_OK (func (pDp, wrk.data ()));
      ++nfev;

      double pDp_eL2 = 0.0;

      for (int i = 0; i < n; ++i)
      {
        wrk2[i] = -wrk[i];
        pDp_eL2 += wrk2[i] * wrk2[i];
      }

      if (!std::isfinite (pDp_eL2))
      {
        stop = 7;
        break;
      }

      const double dF = p_eL2 - pDp_eL2;

      if (updp || dF > 0.0)
      {
        // Broyden rank-one update of the Jacobian
        for (int i = 0; i < n; ++i)
        {
          const double t = (wrk[i] - hx[i] - jac[i] * Dp) / Dp_L2;

          jac[i] += t * Dp;
        }

        ++updjac;
        newjac = 1;
      }

      const double dL = Dp * (mu * Dp + jacTe);

      if (dL > 0.0 && dF > 0.0)
      {
        double t = 2.0 * dF / dL - 1.0;

        t  = 1.0 - t * t * t;
        mu = mu * ((t >= 1.0 / 3.0) ? t : 1.0 / 3.0);
        nu = 2;
        p  = pDp;

        for (int i = 0; i < n; ++i)
        {
          e[i]  = wrk2[i];
          hx[i] = wrk[i];
        }

        p_eL2 = pDp_eL2;
        updp  = 1;
        continue;
      }
    }

    // rejected step (or singular system): more damping
    mu *= nu;

    const int nu2 = nu << 1;

    if (nu2 <= nu)
    {
      stop = 5;
      break;
    }

    nu      = nu2;
    jacTjac = diag;
  }

  if (k >= itmax)
    stop = 3;

  if (info)
  {
    info->init_eL2 = init_p_eL2;
    info->eL2      = p_eL2;
    info->iter     = k;
    info->stop     = stop;
    info->nfev     = nfev;
    info->njac     = njap;
  }

  return NCKDE_OK;
}

// ---- two-variable Welford statistics ----------------------------------------

void
pair_stats_update (PairStats &st, double x0, double x1)
{
  const double cw = st.W + 1.0;
  const double x[2] = {x0, x1};

  for (int i = 0; i < 2; ++i)
  {
    const double delta = x[i] - st.mean[i];
    const double R     = delta * 1.0 / cw;
    const double dvar  = st.W * delta * R;

    st.mean[i] += R;
    st.var[i] += dvar;

    if (i == 0)
      st.cov01 += 1.0 * (x[0] - st.mean[0]) * (x[1] - st.mean[1]);
  }

  st.W = cw;
  st.W2 += 1.0;
  st.bias = 1.0 / (st.W - st.W2 / st.W);
}

// ---- host normals -----------------------------------------------------------

static inline void
philox_host (uint32_t c[4], uint32_t k0, uint32_t k1)
{
  for (int r = 0; r < 10; ++r)
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
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

static inline double
u53_host (uint32_t a, uint32_t b)
{
  return ((double) (a >> 5) * 67108864.0 + (double) (b >> 6)) * (1.0 / 9007199254740992.0);
}

double
host_normal (uint64_t key, uint64_t ctr)
{
  uint32_t c[4] = {(uint32_t) ctr, 0u, (uint32_t) (ctr >> 32), HOST_NORMAL_TAG};

  philox_host (c, (uint32_t) key, (uint32_t) (key >> 32));

  const double a = 1.0 - u53_host (c[0], c[1]);
  const double b = u53_host (c[2], c[3]);

  return std::sqrt (-2.0 * std::log (a)) * std::cos (2.0 * M_PI * b);
}

} // namespace nck
