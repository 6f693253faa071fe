// api_apes.cpp -- APES ensemble sampler driving the kernel-density engine.
//
// Restates NcmFitESMCMCWalkerAPES (ncm_fit_esmcmc_walker_apes.c:509-919)
// and the ensemble loop of NcmFitESMCMC (ncm_fit_esmcmc.c:2136-2288):
//   per step: one uniform "jump" per walker;
//   per half-ensemble H (first, then second):
//     setup: a VKDE is built from the other half's current positions and
//            prepare_interp'ed on their m2lnL; the random-walk widths are
//            0.25 sqrt(diag(full cov)); one proposal per walker of H, drawn
//            from the VKDE (or, with probability rw_prob, a bounded Gaussian
//            random walk), redrawn until inside the bounds;
//     step:  m2lnp_star = q(theta*), m2lnp_cur = q(theta) (VKDE mixed with
//            the truncated-Gaussian random walk);
//     accept if jump < exp(-0.5 [(m2lnL* - m2lnL) + (m2lnp_cur - m2lnp_star)]).
// Batching: the proposals of a half are drawn in one nckde_sample call and
// both transition densities of all its walkers come from one eval call; the
// likelihood is called once per half with every proposal.
#include "nckde_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

using nck::set_error;

namespace
{

constexpr double LN2PI = 1.8378770664093454835606594728112353;

// log of the standard normal probability of [zl, zu] (zl <= zu), evaluated in
// the tail where it is small (ncm_util.c:646-690 computes the same quantity)
double
log_normal_interval (double zl, double zu)
{
  if (zl >= zu)
    return -INFINITY;

  const double s = M_SQRT1_2;

  if (zl > 0.0)
    return std::log (0.5 * (std::erfc (zl * s) - std::erfc (zu * s)));

  if (zu < 0.0)
    return std::log (0.5 * (std::erfc (-zu * s) - std::erfc (-zl * s)));

  return std::log1p (-0.5 * (std::erfc (zu * s) + std::erfc (-zl * s)));
}

} // namespace

// probe tolerance of the walker's mixed-precision densities: half of the
// 3e-6 relative bar the transition densities are held to
constexpr double APES_MIXED_TOL = 1.5e-6;

struct nckde_apes
{
  unsigned n, n2, d;
  int method, ktype;
  double over_smooth = 1.0, shrink = 0.01, local_frac = 0.05, rw_prob = 0.02;
  int use_interp = 1, precision = NCKDE_PREC_FP64;
  unsigned exploration = 0;
  std::vector<double> lb, ub;
  std::vector<double> theta, m2lnL;  // n x d, n
  std::mt19937_64 rng;
  uint64_t seed = 0, draw_ctr = 0;   // counter-based device draws of the KDE proposals
  nckde *sd[2] = {nullptr, nullptr};
  uint64_t accepted = 0, proposed = 0, offboard = 0;
  // optional record of every move (nckde_apes_set_trace)
  double *trace = nullptr;
  size_t trace_cap = 0, trace_len = 0;
  double t_prepare = 0, t_sample = 0, t_eval = 0, t_like = 0;
  bool initialised = false;
};

static int
apes_kernel (const nckde_apes *a, int &kernel, double &nu)
{
  switch (a->ktype)
  {
    case NCKDE_APES_KTYPE_CAUCHY: kernel = NCKDE_KERNEL_ST; nu = 1.0; return NCKDE_OK;
    case NCKDE_APES_KTYPE_ST3: kernel = NCKDE_KERNEL_ST; nu = 3.0; return NCKDE_OK;
    case NCKDE_APES_KTYPE_GAUSS: kernel = NCKDE_KERNEL_GAUSS; nu = 0.0; return NCKDE_OK;
    default:
      set_error ("nckde_apes: unknown kernel type %d", a->ktype);
      return NCKDE_EINVAL;
  }
}

static double
now_s ()
{
  return std::chrono::duration<double> (std::chrono::steady_clock::now ().time_since_epoch ()).count ();
}

static bool
in_bounds (const nckde_apes *a, const double *x)
{
  for (unsigned i = 0; i < a->d; ++i)
    if (!(x[i] >= a->lb[i] && x[i] <= a->ub[i]))
      return false;

  return true;
}

#define TRYA(call)             \
  do {                         \
    int r_ = (call);           \
    if (r_ != NCKDE_OK)        \
      return r_;               \
  } while (0)

// one half-ensemble: walkers [k0, k1) move, the other half builds the KDE
static int
apes_half (nckde_apes *a, unsigned k0, unsigned k1, unsigned other0, int which, const double *jumps,
           nckde_m2lnL_batch_fn like, void *user)
{
  const unsigned d  = a->d;
  const unsigned nh = k1 - k0;
  nckde *sd         = a->sd[which];
  int kernel;
  double nu;

  TRYA (apes_kernel (a, kernel, nu));

  // ---- setup: KDE from the other half ---------------------------------
  double t0 = now_s ();

  TRYA (nckde_reset (sd));
  TRYA (nckde_add_obs_batch (sd, a->theta.data () + (size_t) other0 * d, a->n2, d));

  if (a->use_interp)
    TRYA (nckde_prepare_interp (sd, a->m2lnL.data () + other0, a->n2));
  else
    TRYA (nckde_prepare (sd));

  std::vector<double> rw_std (d, 0.0);

  if (a->rw_prob > 0.0)
  {
    std::vector<double> cov ((size_t) d * d);

    TRYA (nckde_get_full_cov (sd, cov.data ()));

    for (unsigned i = 0; i < d; ++i)
    {
      const double var = cov[(size_t) i * d + i];

      if (!(var > 0.0))
      {
        set_error ("Invalid covariance matrix: diagonal element %u is non-positive.", i);
        return NCKDE_ENUMERIC;
      }

      rw_std[i] = std::sqrt (var) * 0.25;
    }
  }

  // per coordinate: 0.5 ln 2 pi + ln sigma_i, and whether it is unbounded
  std::vector<double> rw_ln0 (d, 0.0);
  std::vector<char> rw_open (d, 0);

  for (unsigned i = 0; i < d; ++i)
  {
    rw_ln0[i]  = 0.5 * LN2PI + std::log (rw_std[i]);
    rw_open[i] = std::isinf (a->lb[i]) && a->lb[i] < 0 && std::isinf (a->ub[i]) && a->ub[i] > 0;
  }

  a->t_prepare += now_s () - t0;
  t0 = now_s ();

  // ---- proposals ------------------------------------------------------
  std::uniform_real_distribution<double> U01 (0.0, 1.0);
  std::normal_distribution<double> N01 (0.0, 1.0);
  std::vector<double> prop ((size_t) nh * d);
  std::vector<char> todo (nh, 1);
  unsigned left = nh;

  for (int round = 0; left > 0; ++round)
  {
    if (round > 10000)
    {
      set_error ("nckde_apes: no proposal inside the bounds after 10000 rounds");
      return NCKDE_ENUMERIC;
    }

    std::vector<unsigned> kde_w;

    for (unsigned j = 0; j < nh; ++j)
    {
      if (!todo[j])
        continue;

      const double *th = a->theta.data () + (size_t) (k0 + j) * d;
      const bool rw    = (a->rw_prob > 0.0) && (U01 (a->rng) < a->rw_prob);

      if (rw)
      {
        // truncated Gaussian random walk, per coordinate
        double *x = prop.data () + (size_t) j * d;

        for (unsigned i = 0; i < d; ++i)
        {
          double v;

          do
          {
            v = th[i] + rw_std[i] * N01 (a->rng);
          } while (v < a->lb[i] || v > a->ub[i]);

          x[i] = v;
        }
      }
      else
      {
        kde_w.push_back (j);
      }
    }

    if (!kde_w.empty ())
    {
      // the uniform, normals and chi-squared variates of the KDE draws come
      // from the device's counter-based stream (one counter per batch)
      std::vector<double> out (kde_w.size () * d);

      TRYA (nck::sample_counter_rng (sd, kde_w.size (), a->seed, a->draw_ctr++, out.data ()));

      for (size_t q = 0; q < kde_w.size (); ++q)
        std::memcpy (prop.data () + (size_t) kde_w[q] * d, out.data () + q * d, sizeof (double) * d);
    }

    left = 0;

    for (unsigned j = 0; j < nh; ++j)
    {
      if (!todo[j])
        continue;

      if (in_bounds (a, prop.data () + (size_t) j * d))
        todo[j] = 0;
      else
        ++left;
    }
  }

  a->t_sample += now_s () - t0;
  t0 = now_s ();

  // the exploration countdown runs at the end of each half's setup, before
  // its acceptance (ncm_fit_esmcmc_walker_apes.c:815-816), so N steps of
  // exploration cover the first N - 1 half-steps
  if (a->exploration > 0)
    a->exploration--;

  // ---- transition densities: q(theta*) and q(theta) in one eval ----------
  std::vector<double> pts ((size_t) 2 * nh * d), q (2 * nh);

  std::memcpy (pts.data (), prop.data (), sizeof (double) * nh * d);
  std::memcpy (pts.data () + (size_t) nh * d, a->theta.data () + (size_t) k0 * d, sizeof (double) * nh * d);
  TRYA (nckde_eval_m2lnp (sd, pts.data (), 2 * nh, d, q.data (), a->precision));

  std::vector<double> m2lnp_star (nh), m2lnp_cur (nh);

  for (unsigned j = 0; j < nh; ++j)
  {
    const double *th = a->theta.data () + (size_t) (k0 + j) * d;
    const double *ts = prop.data () + (size_t) j * d;
    double m_star    = q[j];
    double m_cur     = q[nh + j];

    if (a->rw_prob > 0.0)
    {
      // ncm_fit_esmcmc_walker_apes.c:819-859
      double rw_star = 0.0, rw_cur = 0.0;

      for (unsigned i = 0; i < d; ++i)
      {
        const double sd_i = rw_std[i];
        // the truncation integral is exactly 1 (log 0) for an unbounded coordinate
        const double ln_s = rw_ln0[i] + (rw_open[i] ? 0.0 : log_normal_interval ((a->lb[i] - th[i]) / sd_i,
                                                                                (a->ub[i] - th[i]) / sd_i));
        const double ln_c = rw_ln0[i] + (rw_open[i] ? 0.0 : log_normal_interval ((a->lb[i] - ts[i]) / sd_i,
                                                                                (a->ub[i] - ts[i]) / sd_i));
        const double dz   = (ts[i] - th[i]) / sd_i;

        rw_star += dz * dz + 2.0 * ln_s;
        rw_cur  += dz * dz + 2.0 * ln_c;
      }

      rw_star += -2.0 * std::log (a->rw_prob);
      rw_cur  += -2.0 * std::log (a->rw_prob);
      m_star  += -2.0 * std::log1p (-a->rw_prob);
      m_cur   += -2.0 * std::log1p (-a->rw_prob);

      m_star = (m_star < rw_star) ? m_star - 2.0 * std::log1p (std::exp (-0.5 * (rw_star - m_star)))
                                  : rw_star - 2.0 * std::log1p (std::exp (-0.5 * (m_star - rw_star)));
      m_cur  = (m_cur < rw_cur) ? m_cur - 2.0 * std::log1p (std::exp (-0.5 * (rw_cur - m_cur)))
                                : rw_cur - 2.0 * std::log1p (std::exp (-0.5 * (m_cur - rw_cur)));
    }

    if (!(std::isfinite (m_star) && std::isfinite (m_cur)))
    {
      set_error ("nckde_apes: non-finite transition density");
      return NCKDE_ENUMERIC;
    }

    m2lnp_star[j] = m_star;
    m2lnp_cur[j]  = m_cur;
  }

  a->t_eval += now_s () - t0;
  t0 = now_s ();

  // ---- likelihood at the proposals and acceptance -----------------------
  std::vector<double> like_star (nh);

  if (like (prop.data (), nh, d, like_star.data (), user) != 0)
  {
    set_error ("nckde_apes: the likelihood callback failed");
    return NCKDE_EINVAL;
  }

  a->t_like += now_s () - t0;

  for (unsigned j = 0; j < nh; ++j)
  {
    const unsigned k = k0 + j;
    double prob      = 0.0;

    if (std::isfinite (like_star[j]))
    {
      const double m2lnq = (a->exploration > 0) ? 0.0 : (m2lnp_cur[j] - m2lnp_star[j]);
      const double m2lnp = like_star[j] - a->m2lnL[k] + m2lnq;

      prob = std::min (std::exp (-0.5 * m2lnp), 1.0);
    }

    ++a->proposed;

    const bool acc = jumps[k] < prob;

    if (a->trace != nullptr && a->trace_len < a->trace_cap)
    {
      double *r = a->trace + a->trace_len * (2 * (size_t) d + 9);

      r[0] = which;
      r[1] = k;
      r[2] = (a->exploration > 0) ? 1.0 : 0.0;
      r[3] = a->m2lnL[k];
      r[4] = like_star[j];
      r[5] = m2lnp_star[j];
      r[6] = m2lnp_cur[j];
      r[7] = jumps[k];
      r[8] = acc ? 1.0 : 0.0;
      std::memcpy (r + 9, a->theta.data () + (size_t) k * d, sizeof (double) * d);
      std::memcpy (r + 9 + d, prop.data () + (size_t) j * d, sizeof (double) * d);
      a->trace_len++;
    }

    if (acc)
    {
      std::memcpy (a->theta.data () + (size_t) k * d, prop.data () + (size_t) j * d, sizeof (double) * d);
      a->m2lnL[k] = like_star[j];
      ++a->accepted;
    }
  }

  return NCKDE_OK;
}

extern "C" {

int
nckde_apes_create (nckde_apes **out, unsigned nwalkers, unsigned dim, int method, int ktype, uint64_t seed)
{
  if (out == nullptr || dim == 0 || nwalkers < 4 || (nwalkers % 2) != 0)
  {
    set_error ("nckde_apes_create: nwalkers must be even and >= 4, dim > 0");
    return NCKDE_EINVAL;
  }

  if (method != NCKDE_APES_METHOD_KDE && method != NCKDE_APES_METHOD_VKDE)
  {
    set_error ("nckde_apes_create: unknown method");
    return NCKDE_EINVAL;
  }

  nckde_apes *a = new nckde_apes ();
  int kernel;
  double nu;

  a->n      = nwalkers;
  a->n2     = nwalkers / 2;
  a->d      = dim;
  a->method = method;
  a->ktype  = ktype;
  a->lb.assign (dim, -INFINITY);
  a->ub.assign (dim, INFINITY);
  a->rng.seed (seed);
  a->seed = seed;

  if (apes_kernel (a, kernel, nu) != NCKDE_OK)
  {
    delete a;
    return NCKDE_EINVAL;
  }

  // both methods build a VKDE (ncm_fit_esmcmc_walker_apes.c:561-576)
  for (int w = 0; w < 2; ++w)
  {
    int r = nckde_create (&a->sd[w], NCKDE_VKDE, kernel, nu, dim, NCKDE_CV_NONE);

    if (r != NCKDE_OK)
    {
      nckde_destroy (a->sd[0]);
      delete a;
      return r;
    }
  }

  *out = a;

  return NCKDE_OK;
}

int
nckde_apes_destroy (nckde_apes *a)
{
  if (a == nullptr)
    return NCKDE_OK;

  nckde_destroy (a->sd[0]);
  nckde_destroy (a->sd[1]);
  delete a;

  return NCKDE_OK;
}

int
nckde_apes_set_over_smooth (nckde_apes *a, double v)
{
  a->over_smooth = v;
  TRYA (nckde_set_over_smooth (a->sd[0], v));

  return nckde_set_over_smooth (a->sd[1], v);
}

int
nckde_apes_set_shrink (nckde_apes *a, double v)
{
  a->shrink = v;
  TRYA (nckde_set_shrink (a->sd[0], v));

  return nckde_set_shrink (a->sd[1], v);
}

int
nckde_apes_set_local_frac (nckde_apes *a, double v)
{
  a->local_frac = v;
  TRYA (nckde_set_local_frac (a->sd[0], v));

  return nckde_set_local_frac (a->sd[1], v);
}

int
nckde_apes_set_random_walk_prob (nckde_apes *a, double v)
{
  if (!(v >= 0.0 && v < 1.0))
  {
    set_error ("nckde_apes_set_random_walk_prob: probability must be in [0, 1)");
    return NCKDE_EINVAL;
  }

  a->rw_prob = v;

  return NCKDE_OK;
}

int
nckde_apes_set_use_interp (nckde_apes *a, int use_interp)
{
  a->use_interp = use_interp ? 1 : 0;

  return NCKDE_OK;
}

int
nckde_apes_set_exploration (nckde_apes *a, unsigned steps)
{
  a->exploration = steps;

  return NCKDE_OK;
}

int
nckde_apes_set_precision (nckde_apes *a, int precision)
{
  if (precision != NCKDE_PREC_FP64 && precision != NCKDE_PREC_MIXED)
  {
    set_error ("nckde_apes_set_precision: unknown precision");
    return NCKDE_EINVAL;
  }

  a->precision = precision;

  // the walker holds its mixed-precision transition densities to 3e-6
  // relative (half of it for the probe of the factors), not to the 1e-6 bar
  // of a plain eval
  for (int w = 0; w < 2; ++w)
  {
    const int r = nckde_set_mixed_tolerance (a->sd[w], APES_MIXED_TOL);

    if (r != NCKDE_OK)
      return r;
  }

  return NCKDE_OK;
}

int
nckde_apes_set_bounds (nckde_apes *a, const double *lb, const double *ub)
{
  for (unsigned i = 0; i < a->d; ++i)
  {
    if (!(lb[i] < ub[i]))
    {
      set_error ("nckde_apes_set_bounds: lower bound must be below the upper bound");
      return NCKDE_EINVAL;
    }
  }

  a->lb.assign (lb, lb + a->d);
  a->ub.assign (ub, ub + a->d);

  return NCKDE_OK;
}

int
nckde_apes_init (nckde_apes *a, const double *theta, const double *m2lnL)
{
  if (a == nullptr || theta == nullptr || m2lnL == nullptr)
  {
    set_error ("nckde_apes_init: invalid arguments");
    return NCKDE_EINVAL;
  }

  // ncm_fit_esmcmc_walker_apes.c:490-507: enough walkers per half for the
  // local covariances
  if ((unsigned) (a->local_frac * a->n2) < 2)
  {
    set_error ("Number of walkers per block (%u) is too low for the current dimension (%u).\n"
               "\tToo few points (%u) to estimate local covariances.",
               a->n2, a->d, (unsigned) (a->local_frac * a->n2));
    return NCKDE_ESAMPLE;
  }

  a->theta.assign (theta, theta + (size_t) a->n * a->d);
  a->m2lnL.assign (m2lnL, m2lnL + a->n);
  a->initialised = true;

  return NCKDE_OK;
}

}  // extern "C"

// nsteps ensemble steps, the first of which moves only walkers
// [k_first, n): the loop of ncm_fit_esmcmc.c:2244-2286.  A step that starts
// inside the first half moves [k_first, n/2) with a KDE of the second half
// and then the whole second half; one that starts inside the second half
// moves [k_first, n) only.  The jump uniforms are drawn for the walkers that
// move, in walker order (_ncm_fit_esmcmc_get_jumps (esmcmc, ki, nwalkers)).
static int
apes_steps (nckde_apes *a, unsigned k_first, unsigned nsteps, nckde_m2lnL_batch_fn like, void *user,
            double *chain)
{
  std::uniform_real_distribution<double> U01 (0.0, 1.0);
  std::vector<double> jumps (a->n);
  const size_t row = 1 + a->d;

  for (unsigned s = 0; s < nsteps; ++s)
  {
    const unsigned k0 = (s == 0) ? k_first : 0;

    for (unsigned k = k0; k < a->n; ++k)
      jumps[k] = U01 (a->rng);

    // first half moves with a KDE of the second half, then the reverse
    if (k0 < a->n2)
      TRYA (apes_half (a, k0, a->n2, a->n2, 0, jumps.data (), like, user));

    TRYA (apes_half (a, std::max (k0, a->n2), a->n, 0, 1, jumps.data (), like, user));

    if (chain)
    {
      double *c = chain + (size_t) s * a->n * row;

      for (unsigned k = 0; k < a->n; ++k)
      {
        c[k * row] = a->m2lnL[k];
        std::memcpy (c + k * row + 1, a->theta.data () + (size_t) k * a->d, sizeof (double) * a->d);
      }
    }
  }

  return NCKDE_OK;
}

extern "C" {

int
nckde_apes_run (nckde_apes *a, unsigned nsteps, nckde_m2lnL_batch_fn like, void *user, double *chain)
{
  if (a == nullptr || like == nullptr || !a->initialised)
  {
    set_error ("nckde_apes_run: sampler not initialised or no likelihood");
    return NCKDE_ESTATE;
  }

  return apes_steps (a, 0, nsteps, like, user, chain);
}

int
nckde_apes_run_from (nckde_apes *a, unsigned k_first, unsigned nsteps, nckde_m2lnL_batch_fn like, void *user,
                     double *chain)
{
  if (a == nullptr || like == nullptr || !a->initialised)
  {
    set_error ("nckde_apes_run_from: sampler not initialised or no likelihood");
    return NCKDE_ESTATE;
  }

  if (k_first >= a->n)
  {
    set_error ("nckde_apes_run_from: first walker %u out of range (%u walkers)", k_first, a->n);
    return NCKDE_EINVAL;
  }

  return apes_steps (a, k_first, nsteps, like, user, chain);
}

int
nckde_apes_set_trace (nckde_apes *a, double *buf, size_t max_rows)
{
  if (a == nullptr)
  {
    set_error ("nckde_apes: NULL handle");
    return NCKDE_EINVAL;
  }

  a->trace     = buf;
  a->trace_cap = (buf != nullptr) ? max_rows : 0;
  a->trace_len = 0;

  return NCKDE_OK;
}

int
nckde_apes_get_trace_rows (nckde_apes *a, size_t *rows)
{
  if (a == nullptr || rows == nullptr)
  {
    set_error ("nckde_apes: NULL argument");
    return NCKDE_EINVAL;
  }

  *rows = a->trace_len;

  return NCKDE_OK;
}

int
nckde_apes_get_state (nckde_apes *a, double *theta, double *m2lnL)
{
  if (a == nullptr || !a->initialised)
  {
    set_error ("nckde_apes_get_state: sampler not initialised");
    return NCKDE_ESTATE;
  }

  if (theta)
    std::memcpy (theta, a->theta.data (), sizeof (double) * a->n * a->d);

  if (m2lnL)
    std::memcpy (m2lnL, a->m2lnL.data (), sizeof (double) * a->n);

  return NCKDE_OK;
}

int
nckde_apes_get_stats (nckde_apes *a, uint64_t *accepted, uint64_t *proposed, double *seconds)
{
  if (a == nullptr)
  {
    set_error ("nckde_apes_get_stats: NULL sampler");
    return NCKDE_EINVAL;
  }

  if (accepted)
    *accepted = a->accepted;

  if (proposed)
    *proposed = a->proposed;

  if (seconds)
  {
    seconds[0] = a->t_prepare;
    seconds[1] = a->t_sample;
    seconds[2] = a->t_eval;
    seconds[3] = a->t_like;
  }

  return NCKDE_OK;
}

} // extern "C"
