// api.cpp -- host side of the nckde C-ABI.
//
// Control flow follows NcmStatsDist (ncm_stats_dist.c:703-1094), the KDE /
// VKDE prepare_kernel methods (ncm_stats_dist_kde.c:378-490,
// ncm_stats_dist_vkde.c:362-515) and ncm_nnls_solve (ncm_nnls.c:767-871).
// Every arithmetic step on sample-sized data runs in a HIP kernel; the host
// keeps the index-set bookkeeping of the active-set loop and the O(n)
// decisions of prepare_interp (range check, cut, fallback weights).
#include "nckde_internal.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

namespace nck
{

static thread_local char g_err[1024];

void
set_error (const char *fmt, ...)
{
  va_list ap;

  va_start (ap, fmt);
  vsnprintf (g_err, sizeof (g_err), fmt, ap);
  va_end (ap);
}

hipError_t
DBuf::ensure (size_t nbytes)
{
  if (nbytes <= bytes && p != nullptr)
    return hipSuccess;

  release ();

  if (nbytes == 0)
    nbytes = 8;

  hipError_t e = hipMalloc (&p, nbytes);

  if (e != hipSuccess)
  {
    p     = nullptr;
    bytes = 0;
    return e;
  }

  bytes = nbytes;

  return hipSuccess;
}

void
DBuf::release ()
{
  if (p)
    (void) hipFree (p);

  p     = nullptr;
  bytes = 0;
}

} // namespace nck

using nck::DBuf;
using nck::set_error;

struct nckde
{
  int kind, kernel, cv;
  double nu;
  int d;

  double over_smooth      = 1.0;
  double shrink           = 0.01;
  double local_frac       = 0.05;
  double split_frac       = 0.5;     // "split-frac" property default (ncm_stats_dist.c:430-434)
  bool print_fit          = false;
  uint64_t cv_draws       = 0;       // normals used so far by the split fit's restarts
  unsigned cv_nobj        = 0;       // objective evaluations / NNLS fits of the last CV fit
  double cv_seconds       = 0.0;     // wall time of the last CV fit
  int use_rot_href        = 0;
  bool use_threads        = false;   // "use-threads" property: stored, the device path ignores it
  int cov_type            = NCKDE_COV_SAMPLE;
  unsigned nearPD_maxiter = 200;

  bool has_cov_fixed = false;
  std::vector<double> cov_fixed, cov_fixed_decomp;
  double cov_fixed_lnnorm = 0.0;

  std::vector<double> X;  // sample array (copies), row-major
  size_t n_samples = 0;

  bool prepared = false;
  size_t n_obs = 0, n_k = 0, k = 0;
  int D = 0;
  double href = 0.0, kernel_lnnorm = 0.0, rnorm = 0.0;
  // last NNLS fit: solution (before shrink), gradient A^T r and residual f - A x
  std::vector<double> nnls_x, nnls_g, nnls_r;
  // unconstrained-solve method (NcmNNLSUMethod, ncm_nnls.h:51-60; NORMAL in
  // ncm_stats_dist.c:861), refinement steps after a Cholesky solve, and the
  // solves of the last prepare_interp per rung: Cholesky, symmetric solve,
  // QR, refinement steps
  int nnls_umethod      = NCKDE_NNLS_NORMAL;
  unsigned nnls_refine  = 0;
  unsigned nnls_rungs[4] = {0, 0, 0, 0};
  bool wcum_ready = false;

  int device              = 0;        // device current at nckde_create
  hipStream_t stream      = nullptr;
  hipStream_t last_stream = nullptr;  // stream of the last device-form eval
  hipEvent_t ready_ev     = nullptr;  // marks the handle stream's work for evals on other streams
  bool profiling          = false;    // events around the main eval kernel
  bool pev_recorded       = false;
  hipEvent_t pev[2]       = {nullptr, nullptr};
  std::mutex mtx;

  // prepared device state
  DBuf dX, dY, dcovg, dUg, dlng, dcov, dU, dinfo, dlnnorm, dw, db, dWp, dTh, dknn, dkeys, dscratch, dstatus,
       dwcum, dwhich;
  // mixed-precision factors (built at the first mixed eval after a prepare)
  DBuf dmu, dYc, dymax, dA16, dc32, dcoef, dYw, dpartMf, dflag;
  bool mixed_ready = false, mixed_b_ready = false;
  // pilot of the mixed eval (d > 16 and at least 1024 kernels): copies of
  // the factors of every pilot_stride-th kernel, pilot_n of them
  DBuf dpA16, dpc32, dpcoef;
  int pilot_n         = 0;
  size_t pilot_stride = 0;
  // captured eval launch sequences: an eval of the same shape on the same
  // buffers and stream is replayed as one graph launch from its second
  // repetition on; graph_gen counts the changes (prepare, weights, factor
  // builds, bandwidth) that invalidate every captured graph
  struct EvalGraph
  {
    const double *dX;
    double *dout;
    size_t m;
    int precision, linear;
    hipStream_t s;
    unsigned gen;
    int seen;
    hipGraphExec_t exec;
  };
  unsigned split_override = 0;   // tuning hook: forced split count of evals
  bool graphs        = false;
  unsigned graph_gen = 0;
  std::vector<EvalGraph> graph_cache;
  // FP32 rung: FP32 copies of the packed factors and of b_i - max b
  DBuf dWf, db2f;
  bool f32_ready = false, f32_b_ready = false;
  double bref32  = 0.0;
  // probe of the mixed factors (built with them): risk score per kernel,
  // probe points and their results; the largest relative m2lnp error the
  // probes found, and whether mixed evals run in FP64 for that reason
  DBuf drisk, dprobe;
  DBuf dmulti;   // argument table of fused multi-handle evals led by this handle
  double mixed_probe_err = 0.0;
  double mixed_tol       = 5.0e-7;   // probe error above which mixed evals run in FP64
  bool mixed_fallback    = false;
  double ysc = 1.0, bref = 0.0;
  // work buffers
  DBuf dXe, dout, dpartM, dpartS, du, dz, dchi2, didx;
  DBuf dAT, dM, dS, dbv, dbs, dx, dxt, dr, drt, dg, dP, dpart, dscal, dm2lnp, dlu, dtv, drw, drobust, dpd;
  // NNLS refinement vectors, the symmetric solve's pivot column and the QR
  // rung's work (matrix, row map)
  DBuf dref, dcolv, dqr, dqri;
  // pinned host side of the NNLS loop: solution and gradient read-backs (n
  // doubles each), the active list upload (n ints), solve statistics
  nck::HBuf hnnls;
  // flags of the chained triangular sweeps (status word, then the panel
  // flags; flags_words of them zeroed), the last epoch used, the epoch of a
  // chained solve whose status is still to be read, and whether chained
  // solves are off for this handle (after one gave up)
  DBuf dtsflags;
  size_t flags_words     = 0;
  unsigned trsv_epoch    = 0;
  unsigned trsv_pending  = 0;
  bool trsv_chain_off    = false;
  // cross-validation work buffers
  DBuf dcv1, dcv2, dorder;
};

#define HIPCHK(call)                                                                 \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
    {                                                                                \
      set_error ("%s failed: %s", #call, hipGetErrorString (e_));                    \
      return NCKDE_EHIP;                                                             \
    }                                                                                \
  } while (0)

#define TRY(call)                                                                    \
  do {                                                                               \
    int r_ = (call);                                                                 \
    if (r_ != NCKDE_OK)                                                              \
      return r_;                                                                     \
  } while (0)

static double
rot_bandwidth (const nckde *h, double n)
{
  // ncm_stats_dist_kernel_gauss.c:193-199, ncm_stats_dist_kernel_st.c:225-239
  const unsigned d = (unsigned) h->d;

  if (h->kernel == NCKDE_KERNEL_GAUSS)
    return std::pow (4.0 / (n * (d + 2.0)), 1.0 / (d + 4.0));

  const double nu = (h->nu >= 3.0) ? h->nu : 3.0;

  return std::pow (16.0 * ((nu - 2) * (nu - 2)) * (1.0 + d + nu) * (3.0 + d + nu) /
                   ((2.0 + d) * (d + nu) * (2.0 + d + nu) * (d + 2.0 * nu) * (2.0 + d + 2.0 * nu) * n),
                   1.0 / (d + 4.0));
}

static double
st_lnnorm_const (const nckde *h)
{
  // lgamma(nu/2) - lgamma((nu+d)/2) + (d/2)(ln pi + ln nu)
  // (ncm_stats_dist_kernel_st.c:241-253)
  const unsigned d = (unsigned) h->d;
  const double lg  = std::lgamma (h->nu / 2.0) - std::lgamma ((h->nu + d) / 2.0);
  const double nc  = (d / 2.0) * (nck::LNPI + std::log (h->nu));

  return lg + nc;
}

static double
get_href (const nckde *h)
{
  // ncm_stats_dist.c:476-482, ncm_stats_dist_vkde.c:316-335
  if (h->kind == NCKDE_KDE)
    return h->over_smooth * rot_bandwidth (h, (double) h->n_k);

  if (h->use_rot_href)
    return h->over_smooth * rot_bandwidth (h, (double) h->n_k) / h->local_frac;

  return h->over_smooth;
}

static int
check_handle (nckde *h)
{
  if (h == nullptr)
  {
    set_error ("nckde: NULL handle");
    return NCKDE_EINVAL;
  }

  return NCKDE_OK;
}

static int
check_prepared (nckde *h)
{
  TRY (check_handle (h));

  if (!h->prepared)
  {
    set_error ("nckde: object not prepared (call prepare or prepare_interp first)");
    return NCKDE_ESTATE;
  }

  return NCKDE_OK;
}

// Batched Cholesky with the reference's nearPD / diagonal fallback
// (ncm_stats_dist_kde.c:344-367): factor nmats matrices of dcov into dU,
// then repair the failures on the device.
static int
factor_batch (nckde *h, const double *dcov, double *dU, double *dlnnorm, size_t nmats)
{
  const int d          = h->d;
  const double lconst  = (h->kernel == NCKDE_KERNEL_ST) ? st_lnnorm_const (h) : 0.0;
  hipStream_t s        = h->stream;

  HIPCHK (h->dinfo.ensure (sizeof (int) * nmats));
  HIPCHK (nck::launch_chol_batched (dcov, dU, h->dinfo.as<int> (), dlnnorm, d, (int) nmats, h->kernel, lconst, s));

  std::vector<int> info (nmats);

  HIPCHK (hipMemcpyAsync (info.data (), h->dinfo.p, sizeof (int) * nmats, hipMemcpyDeviceToHost, s));
  HIPCHK (hipStreamSynchronize (s));

  std::vector<int> which;

  for (size_t i = 0; i < nmats; ++i)
    if (info[i] != 0)
      which.push_back ((int) i);

  if (which.empty ())
    return NCKDE_OK;

  // the projection's cm, D_S and eigenvectors of each failing matrix
  const size_t per = 3 * (size_t) d * d * sizeof (double);
  int zero         = 0;

  HIPCHK (h->dwhich.ensure (sizeof (int) * which.size ()));
  HIPCHK (h->dscratch.ensure (per * which.size ()));
  HIPCHK (h->dstatus.ensure (sizeof (int) * 4));
  HIPCHK (hipMemcpyAsync (h->dwhich.p, which.data (), sizeof (int) * which.size (), hipMemcpyHostToDevice, s));
  HIPCHK (hipMemcpyAsync (h->dstatus.p, &zero, sizeof (int), hipMemcpyHostToDevice, s));
  HIPCHK (nck::launch_nearpd_fix (h->dwhich.as<int> (), (int) which.size (), dcov, dU, dlnnorm,
                                  h->dscratch.as<double> (), d, (int) h->nearPD_maxiter, h->kernel, lconst,
                                  h->dstatus.as<int> (), s));

  int status = 0;

  HIPCHK (hipMemcpyAsync (&status, h->dstatus.p, sizeof (int), hipMemcpyDeviceToHost, s));
  HIPCHK (hipStreamSynchronize (s));

  if (status != 0)
  {
    set_error ("_cholesky_decomp: the diagonal of the covariance is not positive.");
    return NCKDE_ENUMERIC;
  }

  return NCKDE_OK;
}

// robust covariance (ncm_stats_vec.c:1834-2064) of nmats row sets of nrows
// rows each (rows == NULL: the first nrows rows, one set) into dcov
static int
robust_cov (nckde *h, int type, int nrows, const int *rows, int nmats, double *dcov)
{
  if (nrows < 4)
  {
    set_error ("ncm_stats_vec_compute_cov_robust_diag: too few points to estimate the covariance [%d].", nrows);
    return NCKDE_ESAMPLE;
  }

  if (h->d > 64)
  {
    set_error ("nckde: robust covariances are implemented up to dimension 64");
    return NCKDE_EUNSUPPORTED;
  }

  HIPCHK (h->drobust.ensure (sizeof (double) * nck::robust_work_doubles (type, h->d, nrows, nmats)));
  HIPCHK (nck::launch_robust_cov (type, h->dX.as<double> (), h->d, nrows, rows, nrows, nmats, dcov,
                                  h->drobust.as<double> (), h->stream));

  return NCKDE_OK;
}

// KDE prepare_kernel (ncm_stats_dist_kde.c:378-490)
static int
kde_prepare_kernel (nckde *h)
{
  const int d   = h->d;
  hipStream_t s = h->stream;

  HIPCHK (h->dX.ensure (sizeof (double) * h->n_obs * d));
  HIPCHK (h->dY.ensure (sizeof (double) * h->n_obs * d));
  HIPCHK (h->dcovg.ensure (sizeof (double) * d * d));
  HIPCHK (h->dUg.ensure (sizeof (double) * d * d));
  HIPCHK (h->dlng.ensure (sizeof (double)));
  HIPCHK (hipMemcpyAsync (h->dX.p, h->X.data (), sizeof (double) * h->n_obs * d, hipMemcpyHostToDevice, s));

  if (h->cov_type == NCKDE_COV_SAMPLE || h->cov_type == NCKDE_COV_FIXED)
    HIPCHK (nck::launch_welford_cov (h->dX.as<double> (), d, (int) h->n_k, nullptr, 0, 1, h->dcovg.as<double> (),
                                     s));
  else
    TRY (robust_cov (h, h->cov_type, (int) h->n_k, nullptr, 1, h->dcovg.as<double> ()));

  if (h->cov_type != NCKDE_COV_FIXED)
  {
    TRY (factor_batch (h, h->dcovg.as<double> (), h->dUg.as<double> (), h->dlng.as<double> (), 1));
    HIPCHK (hipMemcpyAsync (&h->kernel_lnnorm, h->dlng.p, sizeof (double), hipMemcpyDeviceToHost, s));
  }
  else if (h->cov_type == NCKDE_COV_FIXED)
  {
    if (!h->has_cov_fixed)
    {
      set_error ("_ncm_stats_dist_kde_prepare_kernel: cov_type is FIXED but a fixed covariance matrix was not "
                 "provided, use ncm_stats_dist_kde_set_cov_fixed to set one.");
      return NCKDE_ESTATE;
    }

    HIPCHK (hipMemcpyAsync (h->dcovg.p, h->cov_fixed.data (), sizeof (double) * d * d, hipMemcpyHostToDevice, s));
    HIPCHK (hipMemcpyAsync (h->dUg.p, h->cov_fixed_decomp.data (), sizeof (double) * d * d, hipMemcpyHostToDevice, s));
    h->kernel_lnnorm = h->cov_fixed_lnnorm;
  }

  HIPCHK (nck::launch_whiten (h->dX.as<double> (), h->dUg.as<double> (), h->dY.as<double> (), (int) h->n_obs, d, s));
  HIPCHK (hipStreamSynchronize (s));

  return NCKDE_OK;
}

// VKDE _build_cov_array_kdtree (ncm_stats_dist_vkde.c:362-496)
static int
vkde_build_cov_array (nckde *h)
{
  const int d   = h->d;
  hipStream_t s = h->stream;
  const double kk = (h->local_frac * h->n_obs > 2) ? h->local_frac * h->n_obs : 2;

  h->k = (size_t) kk;

  if (h->k > nck::knn_max_k ())
  {
    set_error ("nckde: k = %zu nearest neighbours exceeds the supported maximum %zu", h->k, nck::knn_max_k ());
    return NCKDE_EUNSUPPORTED;
  }

  HIPCHK (h->dknn.ensure (sizeof (int) * h->n_k * h->k));
  HIPCHK (h->dstatus.ensure (sizeof (int) * 4));
  HIPCHK (hipMemsetAsync (h->dstatus.p, 0, sizeof (int) * 4, s));

  // queries in batches so that the distance keys stay under ~512 MiB (none
  // when each query's keys fit in LDS)
  const size_t budget = (size_t) 512 << 20;
  const bool in_lds   = nck::knn_keys_in_lds ((int) h->n_obs, (int) h->k);
  size_t nq           = in_lds ? h->n_k : budget / (sizeof (uint64_t) * h->n_obs);

  nq = std::max<size_t> (1, std::min (nq, h->n_k));

  if (!in_lds)
    HIPCHK (h->dkeys.ensure (sizeof (uint64_t) * h->n_obs * nq));

  // the whitened sample transposed for unit-stride distance reads
  HIPCHK (h->dscratch.ensure (sizeof (double) * h->n_obs * d));
  HIPCHK (nck::launch_transpose (h->dY.as<double> (), (int) h->n_obs, d, h->dscratch.as<double> (), s));

  for (size_t q0 = 0; q0 < h->n_k; q0 += nq)
  {
    const size_t cnt = std::min (nq, h->n_k - q0);

    HIPCHK (nck::launch_knn (h->dscratch.as<double> (), (int) h->n_obs, d, (int) q0, (int) cnt, (int) h->k,
                             h->dkeys.as<uint64_t> (), h->dknn.as<int> (), h->dstatus.as<int> (), s));
  }

  HIPCHK (h->dcov.ensure (sizeof (double) * h->n_k * d * d));
  HIPCHK (h->dU.ensure (sizeof (double) * h->n_k * d * d));
  HIPCHK (h->dlnnorm.ensure (sizeof (double) * h->n_k));
  // local covariances: the sample estimate for SAMPLE and FIXED
  // (ncm_stats_dist_vkde.c:460-473), the robust ones otherwise
  if (h->cov_type == NCKDE_COV_SAMPLE || h->cov_type == NCKDE_COV_FIXED)
    HIPCHK (nck::launch_welford_cov (h->dX.as<double> (), d, (int) h->k, h->dknn.as<int> (), (int) h->k,
                                     (int) h->n_k, h->dcov.as<double> (), s));
  else
    TRY (robust_cov (h, h->cov_type, (int) h->k, h->dknn.as<int> (), (int) h->n_k, h->dcov.as<double> ()));

  int status = 0;

  HIPCHK (hipMemcpyAsync (&status, h->dstatus.p, sizeof (int), hipMemcpyDeviceToHost, s));
  HIPCHK (hipStreamSynchronize (s));

  if (status != 0)
  {
    set_error ("nckde: kNN selection inconsistency (status %d)", status);
    return NCKDE_ENUMERIC;
  }

  return factor_batch (h, h->dcov.as<double> (), h->dU.as<double> (), h->dlnnorm.as<double> (), h->n_k);
}

static int
refresh_bvec (nckde *h)
{
  HIPCHK (h->db.ensure (sizeof (double) * h->n_k));
  HIPCHK (nck::launch_bvec (h->dw.as<double> (), (h->kind == NCKDE_VKDE) ? h->dlnnorm.as<double> () : nullptr,
                            (int) h->n_k, h->db.as<double> (), h->stream));
  h->wcum_ready    = false;
  h->mixed_b_ready = false;
  h->f32_b_ready   = false;
  h->graph_gen++;

  return NCKDE_OK;
}

static int cv_prepare_minimize (nckde *h);
static int cv_split_fit (nckde *h, const double *m2lnp, double min_m2lnp);

// _ncm_stats_dist_prepare (ncm_stats_dist.c:703-789)
static int
do_prepare (nckde *h)
{
  const int d = h->d;

  // a device-form eval still running on a caller's stream reads the factors
  // this prepare overwrites
  if (h->last_stream != nullptr && h->last_stream != h->stream)
    HIPCHK (hipStreamSynchronize (h->last_stream));

  h->last_stream = nullptr;
  h->prepared    = false;
  h->nnls_x.clear ();
  h->nnls_g.clear ();
  h->nnls_r.clear ();
  h->n_obs     = h->n_samples;
  // the split cross-validations keep the first ceil(n split_frac) points as
  // kernels and the rest as held-out observations
  h->n_k       = (h->cv == NCKDE_CV_SPLIT || h->cv == NCKDE_CV_SPLIT_NOFIT)
                   ? (size_t) std::ceil (h->n_samples * h->split_frac)
                   : h->n_samples;

  if (h->n_obs <= (size_t) d)
  {
    set_error ("_ncm_stats_dist_prepare: the sample is too small.");
    return NCKDE_ESAMPLE;
  }

  if (h->kind == NCKDE_VKDE && h->local_frac * h->n_obs < 2)
  {
    set_error ("Too few observations.\n\tThe number of observations is too small to use the local covariance "
               "method.\n\tThe local fraction = %f times number of observations %d is less than 2.",
               h->local_frac, (int) h->n_obs);
    return NCKDE_ESAMPLE;
  }

  if (h->kind == NCKDE_VKDE && h->n_obs <= 2)
  {
    set_error ("_ncm_stats_dist_vkde_build_cov_array_kdtree: n_obs must exceed 2.");
    return NCKDE_ESAMPLE;
  }

  h->D = nck::padded_dim (d);

  if (h->D < 0)
  {
    set_error ("nckde: dimension %d above the supported maximum 64", d);
    return NCKDE_EUNSUPPORTED;
  }

  TRY (kde_prepare_kernel (h));

  if (h->kind == NCKDE_VKDE)
    TRY (vkde_build_cov_array (h));

  h->href = get_href (h);

  {
    std::vector<double> w (h->n_k, 1.0 / (1.0 * h->n_k));

    HIPCHK (h->dw.ensure (sizeof (double) * h->n_k));
    HIPCHK (hipMemcpyAsync (h->dw.p, w.data (), sizeof (double) * h->n_k, hipMemcpyHostToDevice, h->stream));
  }

  HIPCHK (h->dTh.ensure (sizeof (double) * h->n_k * h->D));

  if (h->kind == NCKDE_VKDE)
  {
    HIPCHK (h->dWp.ensure (sizeof (double) * h->n_k * (h->D * (h->D + 1) / 2)));
    HIPCHK (nck::launch_vkde_factors (h->dU.as<double> (), h->dX.as<double> (), (int) h->n_k, d, h->D,
                                      h->dWp.as<double> (), h->dTh.as<double> (), h->stream));
  }
  else
  {
    HIPCHK (nck::launch_kde_factors (h->dY.as<double> (), (int) h->n_k, d, h->D, h->dTh.as<double> (), h->stream));
  }

  TRY (refresh_bvec (h));
  HIPCHK (h->dwcum.ensure (sizeof (double) * (h->n_k + 1)));
  HIPCHK (hipStreamSynchronize (h->stream));

  h->rnorm       = 0.0;
  h->prepared    = true;
  h->mixed_ready = false;
  h->f32_ready   = false;
  h->graph_gen++;

  // bandwidth by minimisation for CV_SPLIT_NOFIT and CV_LOO
  // (ncm_stats_dist.c:769-788)
  if (h->cv == NCKDE_CV_SPLIT_NOFIT || h->cv == NCKDE_CV_LOO)
    TRY (cv_prepare_minimize (h));

  return NCKDE_OK;
}

// fills dAT (n_k x n_obs) with IM^T, optionally scaled by 1/f_j
static int
build_IM (nckde *h, const double *dm2lnp, double min_m2lnp)
{
  nck::IMArgs a;

  HIPCHK (h->dAT.ensure (sizeof (double) * h->n_k * h->n_obs));

  a.kind        = h->kind;
  a.kernel      = h->kernel;
  a.d           = h->d;
  a.D           = h->D;
  a.n_obs       = (int) h->n_obs;
  a.n_k         = (int) h->n_k;
  a.nu          = h->nu;
  a.h2          = h->href * h->href;
  a.inv_h2      = 1.0 / a.h2;
  a.lnnorm_href = h->d * std::log (h->href);
  a.kde_scale   = std::exp (-(h->kernel_lnnorm + h->d * std::log (h->href)));
  a.X           = h->dX.as<double> ();
  a.Y           = h->dY.as<double> ();
  a.Th          = h->dTh.as<double> ();
  a.Wp          = h->dWp.as<double> ();
  a.lnnorm      = h->dlnnorm.as<double> ();
  a.m2lnp       = dm2lnp;
  a.min_m2lnp   = min_m2lnp;
  a.AT          = h->dAT.as<double> ();

  HIPCHK (nck::launch_build_IM (a, h->stream));

  return NCKDE_OK;
}

// ---- active-set NNLS (ncm_nnls.c:464-871) --------------------------------

struct NNLS
{
  nckde *h;
  int n, nobs;
  std::vector<int> Pidx;
};

// unconstrained solve on the sorted index list P (normal equations:
// Cholesky, else Gaussian elimination); solution scattered into dxo
// NCKDE_PROGRESS=1 in the environment: one stderr line per normal-equation
// solve of the NNLS (long fits report as they go)
static bool
progress_enabled ()
{
  static const bool on = [] {
    const char *v = std::getenv ("NCKDE_PROGRESS");
    return v != nullptr && v[0] != '\0' && v[0] != '0';
  }();

  return on;
}

// A Cholesky solve of the normal equations is refined when its pivots span
// more than this ratio (squared extreme diagonal entries of U, a lower
// bound on the condition number of the normal matrix): the normal-equation
// solution then carries relative errors of 1e-6 and more.
constexpr double REFINE_PIVOT_RATIO = 1.0e10;

// Corrected semi-normal refinement of x_P (in dbs) against A itself:
// r = 1 - A_P x, delta = M_PP^{-1} A_P^T r with the factor in dS, x += delta.
// Each step is kept only if the residual norm it leads to is not larger.
static int
nnls_refine (nckde *h, int n, int np, double *dxo)
{
  hipStream_t s  = h->stream;
  const int nobs = (int) h->n_obs;
  const int nb   = (nobs + 255) / 256;
  double *c      = h->dref.as<double> ();
  double *tmp    = c + n;
  double *xprev  = tmp + n;
  double *stat   = h->dpart.as<double> () + nb + 2;
  double last    = -1.0;

  for (unsigned step = 0;; ++step)
  {
    double hs[3];

    HIPCHK (nck::launch_scatter (h->dbs.as<double> (), h->dP.as<int> (), np, n, dxo, s));
    HIPCHK (nck::launch_residual (h->dAT.as<double> (), (int) h->n_k, nobs, dxo, nullptr, h->drt.as<double> (),
                                  h->dpart.as<double> (), stat + 2, h->drw.as<double> (), s));
    HIPCHK (hipMemcpyAsync (hs, stat, sizeof (double) * 3, hipMemcpyDeviceToHost, s));
    HIPCHK (hipStreamSynchronize (s));

    if (step > 0 && hs[2] > last)
    {
      // the step made the fit worse: keep the previous solution
      HIPCHK (hipMemcpyAsync (h->dbs.p, xprev, sizeof (double) * np, hipMemcpyDeviceToDevice, s));
      break;
    }

    if (step > 0 && hs[0] <= 1.0e-16 * hs[1])
      break;

    if (step == h->nnls_refine)
      break;

    last = hs[2];
    HIPCHK (nck::launch_row_dot_idx (h->dAT.as<double> (), h->dP.as<int> (), np, nobs, h->drt.as<double> (), c, s));
    HIPCHK (nck::launch_trsv_pair (h->dS.as<double> (), np, c, tmp, s));
    HIPCHK (hipMemcpyAsync (xprev, h->dbs.p, sizeof (double) * np, hipMemcpyDeviceToDevice, s));
    HIPCHK (nck::launch_add_update (h->dbs.as<double> (), c, np, stat, s));
    h->nnls_rungs[3]++;
  }

  return NCKDE_OK;
}

// Unconstrained solve on the sorted index list P (ncm_nnls.c:655-706):
// NORMAL: Cholesky of the normal matrix (refined against A when its pivots
// are badly spread), else the symmetric solve, else QR on A_P; NORMAL_LU
// starts at the symmetric solve; QR solves min |1 - A_P x| directly.
// The solution is scattered into dxo.
// the pinned NNLS arena: x (n doubles), g (n doubles), P (n ints), 8 stats
static double *
nnls_hx (nckde *h)
{
  return h->hnnls.as<double> ();
}

static double *
nnls_hg (nckde *h)
{
  return h->hnnls.as<double> () + h->n_k;
}

static int *
nnls_hP (nckde *h)
{
  return (int *) (h->hnnls.as<double> () + 2 * h->n_k);
}

static double *
nnls_hstat (nckde *h)
{
  return h->hnnls.as<double> () + 2 * h->n_k + (h->n_k + 1) / 2;
}

// slot of the pinned statistics that receives the chained sweeps' status
constexpr int TRSV_STATUS_SLOT = 6;

// (U^T U) x = b with the factor in dS and b in dbs: the flag-chained sweeps
// when every panel block fits on the device at once, else one launch per
// panel.  A chained solve's status is read with the caller's next read-back
// (trsv_pending holds its epoch).
static int
nnls_trsv (nckde *h, int np, hipStream_t s)
{
  const int npan = (np + 63) / 64;

  if (h->trsv_chain_off || npan > nck::trsv_chain_max_panels ())
  {
    HIPCHK (nck::launch_trsv_pair (h->dS.as<double> (), np, h->dbs.as<double> (), h->dtv.as<double> (), s));
    return NCKDE_OK;
  }

  const size_t words = 1 + 2 * (size_t) npan;

  if (words > h->flags_words || h->trsv_epoch == 0xffffffffu)
  {
    HIPCHK (h->dtsflags.ensure (sizeof (unsigned) * words));
    HIPCHK (hipMemsetAsync (h->dtsflags.p, 0, sizeof (unsigned) * words, s));
    h->flags_words = words;
    h->trsv_epoch  = 0;
  }

  const unsigned epoch = ++h->trsv_epoch;

  HIPCHK (nck::launch_trsv_chain (h->dS.as<double> (), np, h->dbs.as<double> (), h->dtv.as<double> (),
                                  h->dtsflags.as<unsigned> (), epoch, s));
  HIPCHK (hipMemcpyAsync (nnls_hstat (h) + TRSV_STATUS_SLOT, h->dtsflags.p, sizeof (unsigned), hipMemcpyDeviceToHost,
                          s));
  h->trsv_pending = epoch;

  return NCKDE_OK;
}

// true when the chained solve whose status the last read-back brought gave
// up (chained solves are then off for the handle)
static bool
nnls_trsv_failed (nckde *h)
{
  if (h->trsv_pending == 0)
    return false;

  unsigned stv = 0;

  std::memcpy (&stv, nnls_hstat (h) + TRSV_STATUS_SLOT, sizeof (unsigned));

  const bool failed = (stv == h->trsv_pending);

  h->trsv_pending = 0;

  if (failed)
    h->trsv_chain_off = true;

  return failed;
}

static int
nnls_unconstrained (nckde *h, int n, const std::vector<int> &P, double *dxo)
{
  hipStream_t s  = h->stream;
  const int np   = (int) P.size ();
  const int nobs = (int) h->n_obs;
  bool solved    = false;

  if (progress_enabled ())
  {
    fprintf (stderr, "nckde nnls: solve on %d of %d columns\n", np, n);
    fflush (stderr);
  }

  if (np == 0)
  {
    HIPCHK (hipMemsetAsync (dxo, 0, sizeof (double) * n, s));
    return NCKDE_OK;
  }

  // the previous upload from the pinned list ended at that solve's read-back
  std::memcpy (nnls_hP (h), P.data (), sizeof (int) * np);
  HIPCHK (hipMemcpyAsync (h->dP.p, nnls_hP (h), sizeof (int) * np, hipMemcpyHostToDevice, s));

  if (h->nnls_umethod == NCKDE_NNLS_NORMAL)
  {
    struct
    {
      int info, pad;
      double dmax, dmin;
    } st = {0, 0, 0.0, 0.0};

    HIPCHK (nck::launch_gather_sym (h->dM.as<double> (), n, h->dP.as<int> (), np, h->dS.as<double> (),
                                    h->dbv.as<double> (), h->dbs.as<double> (), s));
    HIPCHK (hipMemsetAsync (h->dscal.p, 0, sizeof (int) * 2, s));
    HIPCHK (h->dpd.ensure (sizeof (double) * nck::potrf_diag_doubles (np)));
    HIPCHK (nck::launch_potrf (h->dS.as<double> (), np, h->dscal.as<int> (), h->dpd.as<double> (), s));
    HIPCHK (nck::launch_diag_range (h->dS.as<double> (), np, h->dscal.as<double> () + 1, s));
    HIPCHK (hipMemcpyAsync (nnls_hstat (h), h->dscal.p, sizeof (st), hipMemcpyDeviceToHost, s));
    HIPCHK (hipStreamSynchronize (s));
    std::memcpy (&st, nnls_hstat (h), sizeof (st));

    if (st.info == 0)
    {
      TRY (nnls_trsv (h, np, s));
      h->nnls_rungs[0]++;

      if (h->nnls_refine > 0 && st.dmax * st.dmax >= REFINE_PIVOT_RATIO * st.dmin * st.dmin)
        TRY (nnls_refine (h, n, np, dxo));

      solved = true;
    }
  }

  if (!solved && h->nnls_umethod != NCKDE_NNLS_QR)
  {
    int info = 0;

    HIPCHK (nck::launch_gather_sym (h->dM.as<double> (), n, h->dP.as<int> (), np, h->dS.as<double> (),
                                    h->dbv.as<double> (), h->dbs.as<double> (), s));
    HIPCHK (hipMemsetAsync (h->dscal.p, 0, sizeof (int) * 2, s));
    HIPCHK (nck::launch_lu_factor (h->dS.as<double> (), np, h->dbs.as<double> (), h->dcolv.as<double> (),
                                   h->dscal.as<int> (), s));
    HIPCHK (hipMemcpyAsync (nnls_hstat (h), h->dscal.p, sizeof (int), hipMemcpyDeviceToHost, s));
    HIPCHK (hipStreamSynchronize (s));
    std::memcpy (&info, nnls_hstat (h), sizeof (int));

    if (info == 0)
    {
      HIPCHK (nck::launch_mirror_lower (h->dS.as<double> (), np, s));
      HIPCHK (nck::launch_trsv_backward (h->dS.as<double> (), np, h->dbs.as<double> (), h->dtv.as<double> (), s));
      HIPCHK (hipMemcpyAsync (h->dbs.p, h->dtv.p, sizeof (double) * np, hipMemcpyDeviceToDevice, s));
      h->nnls_rungs[1]++;
      solved = true;
    }
  }

  if (!solved)
  {
    HIPCHK (h->dqr.ensure (sizeof (double) * nck::qr_work_doubles (nobs, np)));
    HIPCHK (h->dqri.ensure (sizeof (int) * (np + 1)));
    HIPCHK (nck::launch_qr_lstsq (h->dAT.as<double> (), nobs, h->dP.as<int> (), np, nullptr, h->dqr.as<double> (),
                                  h->dqri.as<int> (), h->dbs.as<double> (), s));
    h->nnls_rungs[2]++;
  }

  HIPCHK (nck::launch_scatter (h->dbs.as<double> (), h->dP.as<int> (), np, n, dxo, s));

  return NCKDE_OK;
}

// _ncm_nnls_solve_feasible (ncm_nnls.c:728-751)
static int
nnls_feasible (nckde *h, int n, std::vector<char> &inP, double *dxo, std::vector<double> &x, int max_remove)
{
  std::vector<int> P;

  while (true)
  {
    P.clear ();

    for (int a = 0; a < n; ++a)
      if (inP[a])
        P.push_back (a);

    TRY (nnls_unconstrained (h, n, P, dxo));
    HIPCHK (hipMemcpyAsync (nnls_hx (h), dxo, sizeof (double) * n, hipMemcpyDeviceToHost, h->stream));
    HIPCHK (hipStreamSynchronize (h->stream));

    // a chained triangular solve that gave up: the same solve again with
    // one launch per panel
    if (nnls_trsv_failed (h))
      continue;

    std::memcpy (x.data (), nnls_hx (h), sizeof (double) * n);

    std::vector<int> inv;

    for (int a : P)
      if (x[a] < 1.0e-300)
        inv.push_back (a);

    if (inv.empty ())
      break;

    if ((size_t) max_remove >= inv.size ())
    {
      for (int a : inv)
        inP[a] = 0;
    }
    else
    {
      // ncm_iset_remove_smallest_subset (ncm_iset.c:919-979)
      std::stable_sort (inv.begin (), inv.end (), [&] (int u, int v) { return x[u] < x[v]; });

      for (int q = 0; q < max_remove; ++q)
        inP[inv[q]] = 0;
    }
  }

  return NCKDE_OK;
}

static int
nnls_residual (nckde *h, const double *dxv, double *drv, double &rnorm)
{
  const int nb = ((int) h->n_obs + 255) / 256;
  double r2    = 0.0;

  HIPCHK (nck::launch_residual (h->dAT.as<double> (), (int) h->n_k, (int) h->n_obs, dxv, nullptr, drv,
                                h->dpart.as<double> (), h->dpart.as<double> () + nb, h->drw.as<double> (),
                                h->stream));
  HIPCHK (hipMemcpyAsync (nnls_hstat (h), h->dpart.as<double> () + nb, sizeof (double), hipMemcpyDeviceToHost,
                          h->stream));
  HIPCHK (hipStreamSynchronize (h->stream));
  r2 = nnls_hstat (h)[0];
  rnorm = std::sqrt (r2);

  return NCKDE_OK;
}

static int
nnls_mgrad (nckde *h, const double *drv, std::vector<double> &g)
{
  HIPCHK (nck::launch_row_dot (h->dAT.as<double> (), (int) h->n_k, (int) h->n_obs, drv, h->dg.as<double> (), h->stream));
  HIPCHK (hipMemcpyAsync (nnls_hg (h), h->dg.p, sizeof (double) * h->n_k, hipMemcpyDeviceToHost, h->stream));
  HIPCHK (hipStreamSynchronize (h->stream));
  std::memcpy (g.data (), nnls_hg (h), sizeof (double) * h->n_k);

  return NCKDE_OK;
}

// ncm_nnls_solve (ncm_nnls.c:767-871) with A = dAT^T, f = 1; writes dw
static int
nnls_solve (nckde *h)
{
  const int n      = (int) h->n_k;
  const int nobs   = (int) h->n_obs;
  const double reltol = 2.220446049250313e-16;
  hipStream_t s    = h->stream;

  if (h->nnls_umethod != NCKDE_NNLS_QR)
  {
    HIPCHK (h->dM.ensure (sizeof (double) * n * n));
    HIPCHK (h->dS.ensure (sizeof (double) * n * n));
  }
  HIPCHK (h->dbv.ensure (sizeof (double) * n));
  HIPCHK (h->dbs.ensure (sizeof (double) * n));
  HIPCHK (h->dtv.ensure (sizeof (double) * n));
  HIPCHK (h->drw.ensure (sizeof (double) * nck::residual_work_doubles (n, nobs)));
  HIPCHK (h->dx.ensure (sizeof (double) * n));
  HIPCHK (h->dxt.ensure (sizeof (double) * n));
  HIPCHK (h->dr.ensure (sizeof (double) * nobs));
  HIPCHK (h->drt.ensure (sizeof (double) * nobs));
  HIPCHK (h->dg.ensure (sizeof (double) * n));
  HIPCHK (h->dP.ensure (sizeof (int) * n));
  HIPCHK (h->dpart.ensure (sizeof (double) * ((nobs + 255) / 256 + 8)));
  HIPCHK (h->dscal.ensure (64));
  HIPCHK (h->dref.ensure (sizeof (double) * 3 * n));
  HIPCHK (h->dcolv.ensure (sizeof (double) * n));
  HIPCHK (h->hnnls.ensure (sizeof (double) * (2 * (size_t) n + ((size_t) n + 1) / 2 + 8)));

  // M = A^T A, b = A^T f (the QR method works on A alone, ncm_nnls.c:795-797)
  if (h->nnls_umethod != NCKDE_NNLS_QR)
  {
    HIPCHK (nck::launch_gram_nt (h->dAT.as<double> (), n, nobs, h->dM.as<double> (), s));
    HIPCHK (nck::launch_row_dot (h->dAT.as<double> (), n, nobs, nullptr, h->dbv.as<double> (), s));
  }

  std::vector<char> P (n, 1), Ptry (n);
  std::vector<double> x (n), xt (n), g (n);
  double rnorm, lrnorm;

  TRY (nnls_feasible (h, n, P, h->dx.as<double> (), x, n));
  TRY (nnls_residual (h, h->dx.as<double> (), h->dr.as<double> (), rnorm));
  TRY (nnls_mgrad (h, h->dr.as<double> (), g));

  while (true)
  {
    double add_frac = 1.0;
    bool finish     = false;

    while (true)
    {
      // ncm_iset_add_largest_subset (ncm_iset.c:993-1078)
      const double gmin = rnorm * reltol;
      std::vector<int> cand;

      Ptry = P;

      for (int j = 0; j < n; ++j)
        if (!Ptry[j] && g[j] > gmin)
          cand.push_back (j);

      const unsigned kc = (unsigned) cand.size ();
      unsigned added    = 0;

      if (kc > 0)
      {
        const double kf = kc * add_frac;
        const double mx = (kf > 1) ? kf : 1;

        added = (unsigned) ((kc < mx) ? kc : mx);
      }

      add_frac *= 0.5;

      if (added == 0)
      {
        finish = true;
        break;
      }

      std::stable_sort (cand.begin (), cand.end (), [&] (int u, int v) { return g[u] > g[v]; });

      for (unsigned q = 0; q < added; ++q)
        Ptry[cand[q]] = 1;

      TRY (nnls_feasible (h, n, Ptry, h->dxt.as<double> (), xt, (int) added));
      TRY (nnls_residual (h, h->dxt.as<double> (), h->drt.as<double> (), lrnorm));

      if (rnorm - lrnorm > rnorm * reltol)
      {
        P = Ptry;
        HIPCHK (hipMemcpyAsync (h->dx.p, h->dxt.p, sizeof (double) * n, hipMemcpyDeviceToDevice, s));
        HIPCHK (hipMemcpyAsync (h->dr.p, h->drt.p, sizeof (double) * nobs, hipMemcpyDeviceToDevice, s));
        rnorm = lrnorm;
        break;
      }

      if (added == 1)
      {
        finish = true;
        break;
      }
    }

    if (finish)
      break;

    TRY (nnls_mgrad (h, h->dr.as<double> (), g));
  }

  HIPCHK (hipMemcpyAsync (h->dw.p, h->dx.p, sizeof (double) * n, hipMemcpyDeviceToDevice, s));
  h->rnorm = rnorm;

  // the fit's final solution, gradient and residual (diagnostic peek)
  h->nnls_x.resize (n);
  h->nnls_r.resize (nobs);
  h->nnls_g = g;
  HIPCHK (hipMemcpyAsync (h->nnls_x.data (), h->dx.p, sizeof (double) * n, hipMemcpyDeviceToHost, s));
  HIPCHK (hipMemcpyAsync (h->nnls_r.data (), h->dr.p, sizeof (double) * nobs, hipMemcpyDeviceToHost, s));
  HIPCHK (hipStreamSynchronize (s));

  return NCKDE_OK;
}

// _ncm_stats_dist_prepare_interp (ncm_stats_dist.c:878-1094)
static int
do_prepare_interp (nckde *h, const double *m2lnp, size_t len)
{
  const double limit = -2.0 * 2.0 * nck::LOG_DBL_EPSILON;

  TRY (do_prepare (h));

  if (len != h->n_obs)
  {
    set_error ("ncm_stats_dist_prepare_interp: m2lnp has %zu entries, the sample has %zu.", len, h->n_obs);
    return NCKDE_EINVAL;
  }

  double mn = INFINITY, mx = -INFINITY;

  for (size_t i = 0; i < h->n_k; ++i)
  {
    mn = std::min (mn, m2lnp[i]);
    mx = std::max (mx, m2lnp[i]);
  }

  if (mx - mn > limit)
  {
    // every observation is sorted (gsl_sort_index over the whole m2lnp
    // vector, ncm_stats_dist.c:911-914); the cut looks at the first n_k
    std::vector<size_t> srt (h->n_obs);

    std::iota (srt.begin (), srt.end (), 0);
    std::stable_sort (srt.begin (), srt.end (), [&] (size_t a, size_t b) { return m2lnp[a] < m2lnp[b]; });

    size_t n_cut = 0;

    for (size_t i = 0; i < h->n_k; ++i)
      if (m2lnp[srt[i]] - mn > limit)
      {
        n_cut = i;
        break;
      }

    if (n_cut < (size_t) (unsigned) (0.5 * h->n_obs))
    {
      // 90% of the weight inside the window, 10% outside; no shrink
      std::vector<double> w (h->n_k, 0.1 / (h->n_k - n_cut));

      // with split cross-validation a point inside the window may be a
      // held-out observation; the reference writes past its weight vector
      // there (ncm_stats_dist.c:933-940), this build gives weights to
      // kernels only
      for (size_t i = 0; i < n_cut; ++i)
        if (srt[i] < h->n_k)
          w[srt[i]] = 0.9 / n_cut;

      HIPCHK (hipMemcpyAsync (h->dw.p, w.data (), sizeof (double) * h->n_k, hipMemcpyHostToDevice, h->stream));
      TRY (refresh_bvec (h));
      HIPCHK (hipStreamSynchronize (h->stream));

      return NCKDE_OK;
    }

    // drop the points outside the window and refit on the rest
    std::vector<double> Xc, mc;

    Xc.reserve (n_cut * h->d);
    mc.reserve (n_cut);

    for (size_t i = 0; i < h->n_obs; ++i)
      if (m2lnp[i] - mn <= limit)
      {
        mc.push_back (m2lnp[i]);
        Xc.insert (Xc.end (), h->X.begin () + i * h->d, h->X.begin () + (i + 1) * h->d);
      }

    h->X         = Xc;
    h->n_samples = mc.size ();

    return do_prepare_interp (h, mc.data (), mc.size ());
  }

  if (h->n_k > 20000)
    fprintf (stderr, "** WARNING: _ncm_stats_dist_prepare_interp: very large system n = %u!\n", (unsigned) h->n_k);

  HIPCHK (h->dm2lnp.ensure (sizeof (double) * h->n_obs));
  HIPCHK (hipMemcpyAsync (h->dm2lnp.p, m2lnp, sizeof (double) * h->n_obs, hipMemcpyHostToDevice, h->stream));
  if (h->cv == NCKDE_CV_SPLIT)
  {
    TRY (cv_split_fit (h, m2lnp, mn));
  }
  else
  {
    TRY (build_IM (h, h->dm2lnp.as<double> (), mn));
    TRY (nnls_solve (h));
  }

  HIPCHK (h->dstatus.ensure (sizeof (int) * 4));
  HIPCHK (hipMemsetAsync (h->dstatus.p, 0, sizeof (int) * 4, h->stream));
  HIPCHK (nck::launch_shrink (h->dw.as<double> (), (int) h->n_k, h->shrink, h->dstatus.as<int> (), h->stream));

  int status = 0;

  HIPCHK (hipMemcpyAsync (&status, h->dstatus.p, sizeof (int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK (hipStreamSynchronize (h->stream));

  if (status != 0)
  {
    set_error ("_ncm_stats_dist_prepare_interp: total weight is not positive.");
    return NCKDE_ENUMERIC;
  }

  TRY (refresh_bvec (h));
  HIPCHK (hipStreamSynchronize (h->stream));

  return NCKDE_OK;
}

// largest b_i = ln w_i - lnnorm_i (reference point of the base-2 sums)
static int
max_b (nckde *h, double &bref)
{
  std::vector<double> b (h->n_k);

  HIPCHK (hipMemcpyAsync (b.data (), h->db.p, sizeof (double) * h->n_k, hipMemcpyDeviceToHost, h->stream));
  HIPCHK (hipStreamSynchronize (h->stream));
  bref = -INFINITY;

  for (double v : b)
    bref = std::max (bref, v);

  if (!std::isfinite (bref))
  {
    set_error ("nckde: no kernel has a positive weight");
    return NCKDE_ENUMERIC;
  }

  return NCKDE_OK;
}

static int eval_mixed (nckde *h, const double *dX, size_t m, double *dout, int linear, hipStream_t s);
static int eval_device (nckde *h, const double *dX, size_t m, double *dout, int precision, int linear,
                        hipStream_t s);

// Probe of freshly built mixed factors.  The split products carry an absolute
// error in z that grows with |B_i| |y| (kernel narrowness times distance from
// the sample mean), so narrow kernels far out can exceed the 1e-6 bar in
// m2lnp.  The PROBE_KERNELS kernels with the largest score |B_i| (|y_i| + 1)
// each get two points 1.5 kernel widths from their centre (the sample
// transform with z = 1.5 (1, +-1, 1, ...)/sqrt(d)); both precisions evaluate
// them and a relative m2lnp error above the handle's tolerance (5e-7 unless
// set) sends its mixed evals to the FP64 path until the next prepare.
constexpr int PROBE_KERNELS = 64;

static int
mixed_probe (nckde *h)
{
  const int d   = h->d;
  const int n   = (int) h->n_k;
  const int np  = std::min (PROBE_KERNELS, n);
  const int m   = 2 * np;
  hipStream_t s = h->stream;

  // device layout: x[m*d] | out64[m] | outmx[m] | z[m*d] | chi2[m] | err | sel[np] | idx[m]
  const size_t nd = (size_t) m * d;
  HIPCHK (h->dprobe.ensure (sizeof (double) * (2 * nd + 3 * (size_t) m + 1) + sizeof (uint32_t) * (np + m)));

  double *dx    = h->dprobe.as<double> ();
  double *do64  = dx + nd;
  double *domx  = do64 + m;
  double *dz    = domx + m;
  double *dc2   = dz + nd;
  double *derr  = dc2 + m;
  uint32_t *sel = (uint32_t *) (derr + 1);
  uint32_t *did = sel + np;

  HIPCHK (nck::launch_probe_points (h->drisk.as<float> (), n, np, d, 1.5 / std::sqrt ((double) d),
                                    (h->kernel == NCKDE_KERNEL_ST) ? h->nu : 1.0, sel, did, dz, dc2, s));
  HIPCHK (nck::launch_sample_at (did, dz, dc2, m, d, h->kind, h->kernel, h->nu, h->href,
                                 (h->kind == NCKDE_VKDE) ? h->dU.as<double> () : h->dUg.as<double> (),
                                 h->dX.as<double> (), dx, s));
  TRY (eval_device (h, dx, m, do64, NCKDE_PREC_FP64, 0, s));
  TRY (eval_mixed (h, dx, m, domx, 0, s));
  HIPCHK (nck::launch_probe_err (domx, do64, m, derr, s));

  double err = 0.0;

  HIPCHK (hipMemcpyAsync (&err, derr, sizeof (double), hipMemcpyDeviceToHost, s));
  HIPCHK (hipStreamSynchronize (s));

  h->mixed_probe_err = err;
  h->mixed_fallback  = !(err <= h->mixed_tol);

  return NCKDE_OK;
}

// The mixed eval's pilot: PILOT_KERNELS kernels at a fixed stride over the
// sample.  At d > 16 the terms of most points lie far below b_ref (a
// Student-t term falls as (nu + d)/2 log(1 + chi2/nu)), so the main pass runs
// against the pilot's largest term per point instead; without it most points
// of a d = 64 eval would underflow and run twice.
constexpr int PILOT_KERNELS   = 256;
constexpr size_t PILOT_MIN_N  = 4 * PILOT_KERNELS;

// copies the pilot kernels' factors (coef only when coef_only)
static int
pilot_gather (nckde *h, bool coef_only)
{
  const int DP = nck::mfma_dim (h->d);

  if (DP < 32 || h->n_k < PILOT_MIN_N)
  {
    h->pilot_n = 0;
    return NCKDE_OK;
  }

  const size_t np = PILOT_KERNELS, st = h->n_k / PILOT_KERNELS;
  const size_t wA = sizeof (_Float16) * 2 * DP * DP, wc = sizeof (float) * DP, wf = sizeof (float) * 4;

  if (!coef_only)
  {
    HIPCHK (h->dpA16.ensure (wA * np));
    HIPCHK (h->dpc32.ensure (wc * np));
    HIPCHK (h->dpcoef.ensure (wf * np));
    HIPCHK (hipMemcpy2DAsync (h->dpA16.p, wA, h->dA16.p, wA * st, wA, np, hipMemcpyDeviceToDevice, h->stream));
    HIPCHK (hipMemcpy2DAsync (h->dpc32.p, wc, h->dc32.p, wc * st, wc, np, hipMemcpyDeviceToDevice, h->stream));
  }

  HIPCHK (hipMemcpy2DAsync (h->dpcoef.p, wf, h->dcoef.p, wf * st, wf, np, hipMemcpyDeviceToDevice, h->stream));
  h->pilot_n      = (int) np;
  h->pilot_stride = st;

  return NCKDE_OK;
}

static int
ensure_mixed (nckde *h)
{
  const int d   = h->d;
  const int DP  = nck::mfma_dim (d);
  const int kpt = (DP >= 32) ? 1 : 32 / DP;
  const size_t n_pad = ((h->n_k + kpt - 1) / kpt) * kpt;

  if (!h->mixed_ready)
  {
    TRY (max_b (h, h->bref));
    HIPCHK (h->dmu.ensure (sizeof (double) * d));
    HIPCHK (h->dYc.ensure (sizeof (double) * h->n_k * d));
    HIPCHK (h->dymax.ensure (sizeof (unsigned long long)));
    HIPCHK (h->dA16.ensure (sizeof (_Float16) * n_pad * 2 * DP * DP));
    HIPCHK (h->dc32.ensure (sizeof (float) * n_pad * DP));
    HIPCHK (h->dcoef.ensure (sizeof (float) * 4 * n_pad));
    HIPCHK (h->drisk.ensure (sizeof (float) * n_pad));
    HIPCHK (nck::mixed_build (h->dX.as<double> (), (int) h->n_k, d, h->D, DP,
                              (h->kind == NCKDE_VKDE) ? h->dWp.as<double> () : nullptr,
                              h->dUg.as<double> (), h->db.as<double> (), h->bref, 1.0 / h->href, h->kernel, h->nu,
                              h->dmu.as<double> (), h->dYc.as<double> (),
                              h->dymax.as<unsigned long long> (), &h->ysc, h->dA16.as<_Float16> (),
                              h->dc32.as<float> (), h->dcoef.as<float> (), h->drisk.as<float> (), h->stream));
    TRY (pilot_gather (h, false));
    h->mixed_ready   = true;
    h->mixed_b_ready = true;
    h->graph_gen++;
    TRY (mixed_probe (h));
  }
  else if (!h->mixed_b_ready)
  {
    TRY (max_b (h, h->bref));
    HIPCHK (nck::mixed_refresh_b (h->db.as<double> (), h->bref, (int) h->n_k, h->dcoef.as<float> (), h->stream));
    TRY (pilot_gather (h, true));
    h->graph_gen++;
    h->mixed_b_ready = true;
  }

  return NCKDE_OK;
}

// FP32 copies of the factors (at the first FP32 eval after a prepare) and of
// b_i - max b (after every change of the weights)
static int
ensure_f32 (nckde *h)
{
  if (!h->f32_ready)
  {
    if (h->kind == NCKDE_VKDE)
    {
      const size_t cnt = h->n_k * (size_t) (h->D * (h->D + 1) / 2);

      HIPCHK (h->dWf.ensure (sizeof (float) * cnt));
      HIPCHK (nck::launch_f32_factors (h->dWp.as<double> (), cnt, h->dWf.as<float> (), h->stream));
    }

    h->f32_ready   = true;
    h->f32_b_ready = false;
    h->graph_gen++;
  }

  if (!h->f32_b_ready)
  {
    TRY (max_b (h, h->bref32));
    HIPCHK (h->db2f.ensure (sizeof (float) * h->n_k));
    HIPCHK (nck::launch_f32_b2 (h->db.as<double> (), h->bref32, (int) h->n_k, h->db2f.as<float> (), h->stream));
    h->f32_b_ready = true;
    h->graph_gen++;
  }

  return NCKDE_OK;
}

// Split count and scratch layout of the two eval forms, shared by the
// allocating entry points (handle buffers) and the allocation-free one (the
// device array `scratch` below).
// Kernel-range split counts chosen by the project's tuner (numcosmo_amd/tune.py
// times every candidate on the device and writes eval_tune.inc): precision
// (0 FP64, 1 mixed, 2 FP32), dimension (the MFMA dimension DP for mixed,
// the padded dimension D otherwise), log2 n and log2 m of the tuned case,
// split count.  An eval uses the entry of its precision and dimension
// nearest in (log2 n, log2 m) within one octave in each, else the default
// rule: about 8192 blocks in all and at least 8 kernel groups per split
// (mixed), about 16384 blocks and at least 32 kernels per split (FP64 and
// FP32).  Many short splits keep a split's kernel factors in its XCD's L2
// and even out the blocks' finishing times.
struct SplitTune
{
  int prec, dim, lgn, lgm, nsplit;
};

static const SplitTune split_tune[] = {
#include "eval_tune.inc"
  {-1, 0, 0, 0, 0}
};

static int
tuned_nsplit (int prec, int dim, size_t n, size_t m)
{
  const double ln = std::log2 ((double) std::max<size_t> (n, 1));
  const double lm = std::log2 ((double) std::max<size_t> (m, 1));
  double best     = 1e30;
  int ns          = 0;

  for (const SplitTune &e : split_tune)
  {
    if (e.prec != prec || e.dim != dim)
      continue;

    const double dn = std::fabs (ln - e.lgn), dm = std::fabs (lm - e.lgm);

    if (dn <= 1.0 && dm <= 1.0 && dn + dm < best)
    {
      best = dn + dm;
      ns   = e.nsplit;
    }
  }

  return ns;
}

static int
mixed_nsplit (const nckde *h, size_t m)
{
  const int DP      = nck::mfma_dim (h->d);
  const int kpt     = (DP >= 32) ? 1 : 32 / DP;
  const int ngroups = (int) ((h->n_k + kpt - 1) / kpt);
  int chosen        = (h->split_override > 0) ? (int) h->split_override : tuned_nsplit (1, DP, h->n_k, m);

  if (chosen > 0)
    return std::max (1, std::min (chosen, ngroups));

  const int bx = (int) ((m + nck::mixed_points_per_block (DP) - 1) / nck::mixed_points_per_block (DP));
  int nsplit   = (8192 + bx - 1) / bx;

  return std::max (1, std::min (nsplit, std::max (1, ngroups / 8)));
}

static int
fp64_nsplit (const nckde *h, size_t m, int prec)
{
  int chosen = (h->split_override > 0) ? (int) h->split_override : tuned_nsplit (prec, h->D, h->n_k, m);

  if (chosen > 0)
    return std::max (1, std::min (chosen, (int) h->n_k));

  const int nblk = (int) ((m + 255) / 256);
  int nsplit     = (16384 + nblk - 1) / nblk;
  const int maxs = (int) ((h->n_k + 31) / 32);

  return std::max (1, std::min (nsplit, maxs));
}

static size_t
align256 (size_t b)
{
  return (b + 255) & ~(size_t) 255;
}

// byte sizes of the mixed eval's work: points (hi, lo, scales), flags and
// row maxima, split maxima, split sums
static void
mixed_work_bytes (const nckde *h, size_t m, size_t sz[4])
{
  const int DP     = nck::mfma_dim (h->d);
  const int nsplit = mixed_nsplit (h, m);

  sz[0] = m * (2 * sizeof (_Float16) * DP + sizeof (float) * 2);
  sz[1] = m * (sizeof (int) + sizeof (float));
  sz[2] = sizeof (float) * m * nsplit;
  sz[3] = sizeof (double) * m * nsplit;
}

// ensure_mixed must have run (eval_device does it before ordering the streams).
// ws != NULL: the work arrays are carved from it (laid out as
// mixed_work_bytes, each start aligned to 256 bytes) instead of the handle's
static int
eval_mixed_ws (nckde *h, const double *dX, size_t m, double *dout, int linear, hipStream_t s, char *ws);
static int mixed_args (nckde *h, const double *dX, size_t m, double *dout, int linear, hipStream_t s, char *ws,
                       nck::MixedArgs &a);

static int
eval_mixed (nckde *h, const double *dX, size_t m, double *dout, int linear, hipStream_t s)
{
  return eval_mixed_ws (h, dX, m, dout, linear, s, nullptr);
}

static int
eval_mixed_ws (nckde *h, const double *dX, size_t m, double *dout, int linear, hipStream_t s, char *ws)
{
  nck::MixedArgs a;

  TRY (mixed_args (h, dX, m, dout, linear, s, ws, a));

  if (h->profiling)
  {
    a.ev0           = h->pev[0];
    a.ev1           = h->pev[1];
    h->pev_recorded = true;
  }

  HIPCHK (nck::launch_mixed_points (a, s));
  HIPCHK (nck::launch_eval_mixed (a, s));

  return NCKDE_OK;
}

// the arguments of a mixed eval of m device points on stream s, with its
// work arrays from the handle (ws == NULL, sized here) or carved from ws
static int
mixed_args (nckde *h, const double *dX, size_t m, double *dout, int linear, hipStream_t s, char *ws,
            nck::MixedArgs &a)
{
  const int DP     = nck::mfma_dim (h->d);
  const int nsplit = mixed_nsplit (h, m);
  size_t sz[4];
  void *Yw, *flag, *partMf, *partS;

  mixed_work_bytes (h, m, sz);

  if (ws == nullptr)
  {
    if (h->last_stream != nullptr && h->last_stream != s)
      HIPCHK (hipStreamSynchronize (h->last_stream));

    h->last_stream = s;

    HIPCHK (h->dYw.ensure (sz[0]));
    HIPCHK (h->dflag.ensure (sz[1]));
    HIPCHK (h->dpartMf.ensure (sz[2]));
    HIPCHK (h->dpartS.ensure (sz[3]));
    Yw     = h->dYw.p;
    flag   = h->dflag.p;
    partMf = h->dpartMf.p;
    partS  = h->dpartS.p;
  }
  else
  {
    Yw     = ws;
    flag   = ws + align256 (sz[0]);
    partMf = (char *) flag + align256 (sz[1]);
    partS  = (char *) partMf + align256 (sz[2]);
  }

  a.kernel = h->kernel;
  a.d      = h->d;
  a.n      = (int) h->n_k;
  a.m      = (int) m;
  a.nsplit = nsplit;
  a.nu     = h->nu;
  a.kappa  = -0.5 * (h->nu + (double) (unsigned) h->d);
  a.lnh_d  = h->d * std::log (h->href);
  a.lnnorm_shared = (h->kind == NCKDE_KDE) ? h->kernel_lnnorm : 0.0;
  a.f.DP   = DP;
  a.f.A    = h->dA16.p;
  a.f.c    = h->dc32.as<float> ();
  a.f.coef = h->dcoef.as<float> ();
  a.X      = dX;
  a.mu     = h->dmu.as<double> ();
  a.G      = h->dUg.as<double> ();
  a.ysc    = h->ysc;
  a.Yh     = Yw;
  a.partM  = (float *) partMf;
  a.partS  = (double *) partS;
  a.flag   = (int *) flag;
  a.M0     = (float *) ((int *) flag + m);
  a.bref   = h->bref;
  a.out    = dout;
  a.linear = linear;
  a.pn     = h->pilot_n;

  if (a.pn > 0)
  {
    const int bx = (int) ((m + nck::mixed_points_per_block (DP) - 1) / nck::mixed_points_per_block (DP));

    a.pf.DP    = DP;
    a.pf.A     = h->dpA16.p;
    a.pf.c     = h->dpc32.as<float> ();
    a.pf.coef  = h->dpcoef.as<float> ();
    a.pnsplit  = std::max (1, std::min ({(1024 + bx - 1) / bx, 8, nsplit}));
  }

  return NCKDE_OK;
}

// FP64 eval (FP32 rung when f32 is set; ensure_f32 must have run); ws !=
// NULL: the split partials (two arrays of m x nsplit doubles, the second at
// an offset aligned to 256 bytes) come from it
static int
eval_fp64_ws (nckde *h, const double *dX, size_t m, double *dout, int linear, hipStream_t s, char *ws,
              bool f32 = false);

template <class Run>
static int eval_graph (nckde *h, const double *dX, size_t m, double *dout, int precision, int linear, hipStream_t s,
                       const Run &run);

static int
eval_device (nckde *h, const double *dX, size_t m, double *dout, int precision, int linear, hipStream_t s)
{
  if (m == 0)
    return NCKDE_OK;

  if (precision != NCKDE_PREC_FP64 && precision != NCKDE_PREC_MIXED && precision != NCKDE_PREC_FP32)
  {
    set_error ("nckde: unknown precision %d", precision);
    return NCKDE_EINVAL;
  }

  // the mixed and FP32 factors are built (or their weights refreshed) on the
  // handle's stream; an eval on another stream must not start before that
  // work ends.  Mixed factors whose probe exceeded the error bar evaluate in
  // FP64.
  if (precision == NCKDE_PREC_MIXED)
  {
    TRY (ensure_mixed (h));

    if (h->mixed_fallback)
      precision = NCKDE_PREC_FP64;
  }
  else if (precision == NCKDE_PREC_FP32)
  {
    TRY (ensure_f32 (h));
  }

  if (s != h->stream)
  {
    HIPCHK (hipEventRecord (h->ready_ev, h->stream));
    HIPCHK (hipStreamWaitEvent (s, h->ready_ev, 0));
  }

  auto run = [&] () -> int {
    if (precision == NCKDE_PREC_MIXED)
      return eval_mixed (h, dX, m, dout, linear, s);

    return eval_fp64_ws (h, dX, m, dout, linear, s, nullptr, precision == NCKDE_PREC_FP32);
  };

  if (!h->graphs || h->profiling || s == nullptr)
    return run ();

  return eval_graph (h, dX, m, dout, precision, linear, s, run);
}

// Replay of a captured eval: the first eval of a shape runs eagerly (it
// sizes every buffer and settles the stream bookkeeping), the second is
// captured into a graph (kernel launches only: the eval path has no memset
// or allocation once its buffers exist) and launched, later ones launch the
// graph.  Graphs of an older generation are destroyed.
template <class Run>
static int
eval_graph (nckde *h, const double *dX, size_t m, double *dout, int precision, int linear, hipStream_t s,
            const Run &run)
{
  auto &cache = h->graph_cache;

  for (size_t i = 0; i < cache.size ();)
  {
    if (cache[i].gen != h->graph_gen)
    {
      if (cache[i].exec)
        (void) hipGraphExecDestroy (cache[i].exec);

      cache.erase (cache.begin () + i);
    }
    else
    {
      ++i;
    }
  }

  nckde::EvalGraph *g = nullptr;

  for (auto &e : cache)
    if (e.dX == dX && e.dout == dout && e.m == m && e.precision == precision && e.linear == linear && e.s == s)
      g = &e;

  if (g != nullptr && g->exec != nullptr)
  {
    HIPCHK (hipGraphLaunch (g->exec, s));
    return NCKDE_OK;
  }

  if (g == nullptr)
  {
    if (cache.size () >= 16)
    {
      if (cache.front ().exec)
        (void) hipGraphExecDestroy (cache.front ().exec);

      cache.erase (cache.begin ());
    }

    cache.push_back ({dX, dout, m, precision, linear, s, h->graph_gen, 1, nullptr});
    return run ();
  }

  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  // nothing inside the capture may wait on another stream
  if (h->last_stream != nullptr && h->last_stream != s)
    HIPCHK (hipStreamSynchronize (h->last_stream));

  h->last_stream = s;

  HIPCHK (hipStreamBeginCapture (s, hipStreamCaptureModeThreadLocal));

  const int st = run ();
  const hipError_t ec = hipStreamEndCapture (s, &graph);

  if (st != NCKDE_OK)
  {
    if (graph)
      (void) hipGraphDestroy (graph);

    return st;
  }

  HIPCHK (ec);

  const hipError_t ei = hipGraphInstantiate (&exec, graph, nullptr, nullptr, 0);

  (void) hipGraphDestroy (graph);
  HIPCHK (ei);
  g->exec = exec;
  HIPCHK (hipGraphLaunch (exec, s));

  return NCKDE_OK;
}

static int
eval_fp64_ws (nckde *h, const double *dX, size_t m, double *dout, int linear, hipStream_t s, char *ws, bool f32)
{
  const int nsplit = fp64_nsplit (h, m, f32 ? NCKDE_PREC_FP32 : NCKDE_PREC_FP64);
  double *partM, *partS;

  if (ws == nullptr)
  {
    // the partial buffers belong to the handle: work queued on another
    // stream must finish before this call reuses them
    if (h->last_stream != nullptr && h->last_stream != s)
      HIPCHK (hipStreamSynchronize (h->last_stream));

    h->last_stream = s;

    HIPCHK (h->dpartM.ensure (sizeof (double) * m * nsplit));
    HIPCHK (h->dpartS.ensure (sizeof (double) * m * nsplit));
    partM = h->dpartM.as<double> ();
    partS = h->dpartS.as<double> ();
  }
  else
  {
    partM = (double *) ws;
    partS = (double *) (ws + align256 (sizeof (double) * m * nsplit));
  }

  nck::EvalArgs a;

  a.kind          = h->kind;
  a.kernel        = h->kernel;
  a.d             = h->d;
  a.D             = h->D;
  a.n             = (int) h->n_k;
  a.m             = (int) m;
  a.nu            = h->nu;
  a.kappa         = -0.5 * (h->nu + (double) (unsigned) h->d);
  a.inv_h2        = 1.0 / (h->href * h->href);
  a.lnh_d         = h->d * std::log (h->href);
  a.lnnorm_shared = (h->kind == NCKDE_KDE) ? h->kernel_lnnorm : 0.0;
  a.Th            = h->dTh.as<double> ();
  a.Wp            = h->dWp.as<double> ();
  a.b             = h->db.as<double> ();
  a.Ug            = h->dUg.as<double> ();
  a.X             = dX;
  a.partM         = partM;
  a.partS         = partS;
  a.nsplit        = nsplit;
  a.out           = dout;
  a.linear        = linear;

  if (f32)
  {
    a.Wf   = (h->kind == NCKDE_VKDE) ? h->dWf.as<float> () : nullptr;
    a.b2f  = h->db2f.as<float> ();
    a.bref = h->bref32;
  }

  if (h->profiling)
  {
    HIPCHK (hipEventRecord (h->pev[0], s));
    h->pev_recorded = true;
  }

  HIPCHK (nck::launch_eval_f64 (a, s));

  if (h->profiling)
    HIPCHK (hipEventRecord (h->pev[1], s));

  HIPCHK (nck::launch_eval_merge (a, s));

  return NCKDE_OK;
}

static int
eval_host (nckde *h, const double *X, size_t m, size_t ld, double *out, int precision, int linear)
{
  const int d = h->d;

  if (m == 0)
    return NCKDE_OK;

  if (X == nullptr || out == nullptr || ld < (size_t) d)
  {
    set_error ("nckde_eval: invalid arguments");
    return NCKDE_EINVAL;
  }

  HIPCHK (h->dXe.ensure (sizeof (double) * m * d));
  HIPCHK (h->dout.ensure (sizeof (double) * m));

  if (ld == (size_t) d)
    HIPCHK (hipMemcpyAsync (h->dXe.p, X, sizeof (double) * m * d, hipMemcpyHostToDevice, h->stream));
  else
    HIPCHK (hipMemcpy2DAsync (h->dXe.p, sizeof (double) * d, X, sizeof (double) * ld, sizeof (double) * d, m,
                              hipMemcpyHostToDevice, h->stream));

  TRY (eval_device (h, h->dXe.as<double> (), m, h->dout.as<double> (), precision, linear, h->stream));
  HIPCHK (hipMemcpyAsync (out, h->dout.p, sizeof (double) * m, hipMemcpyDeviceToHost, h->stream));
  HIPCHK (hipStreamSynchronize (h->stream));

  return NCKDE_OK;
}

static int
ensure_wcum (nckde *h)
{
  if (!h->wcum_ready)
  {
    HIPCHK (nck::launch_wcum (h->dw.as<double> (), (int) h->n_k, h->dwcum.as<double> (), h->stream));
    h->wcum_ready = true;
  }

  return NCKDE_OK;
}

// ---- cross-validated bandwidth (ncm_stats_dist.c:484-701, 1018-1073) ------
//
// The objectives run on the device (IM, NNLS, evaluation, Monte-Carlo pairs);
// the host keeps the scalar minimisers and the serial reductions whose order
// follows the reference's loops.

// over_smooth = exp(lnos) and href = factor * href(over_smooth), as every
// objective of the reference sets them before it evaluates
static void
cv_set_lnos (nckde *h, double lnos, double factor)
{
  h->over_smooth = std::exp (lnos);
  h->href        = factor * get_href (h);
  h->mixed_ready = false;
  h->graph_gen++;
}

// n(n-1) in the reference's unsigned arithmetic (guint products)
static double
cv_pairs (size_t n)
{
  return (double) (unsigned) ((unsigned) n * ((unsigned) n - 1u));
}

// IM row sums of every observation, without and with the diagonal term
static int
cv_im_sums (nckde *h, std::vector<double> &off, std::vector<double> &dg)
{
  const size_t no = h->n_obs;

  TRY (build_IM (h, nullptr, 0.0));
  HIPCHK (h->dcv2.ensure (sizeof (double) * 2 * no));
  HIPCHK (nck::launch_im_obs_sums (h->dAT.as<double> (), (int) h->n_k, (int) no, h->dcv2.as<double> (),
                                   h->dcv2.as<double> () + no, h->stream));
  off.resize (no);
  dg.resize (no);
  HIPCHK (hipMemcpyAsync (off.data (), h->dcv2.p, sizeof (double) * no, hipMemcpyDeviceToHost, h->stream));
  HIPCHK (hipMemcpyAsync (dg.data (), h->dcv2.as<double> () + no, sizeof (double) * no, hipMemcpyDeviceToHost,
                          h->stream));
  HIPCHK (hipStreamSynchronize (h->stream));

  return NCKDE_OK;
}

// m2lnp of rows [r0, r0 + m) of the sample, FP64
static int
cv_eval_rows (nckde *h, size_t r0, size_t m, std::vector<double> &v)
{
  v.resize (m);

  if (m == 0)
    return NCKDE_OK;

  HIPCHK (h->dcv1.ensure (sizeof (double) * m));
  TRY (eval_device (h, h->dX.as<double> () + r0 * h->d, m, h->dcv1.as<double> (), NCKDE_PREC_FP64, 0, h->stream));
  HIPCHK (hipMemcpyAsync (v.data (), h->dcv1.p, sizeof (double) * m, hipMemcpyDeviceToHost, h->stream));
  HIPCHK (hipStreamSynchronize (h->stream));

  return NCKDE_OK;
}

// _ncm_stats_dist_m2lnp (ncm_stats_dist.c:484-510): sum of m2lnp over the
// held-out observations, kernels at uniform weights
static int
cv_obj_m2lnp (nckde *h, double lnos, double &val)
{
  std::vector<double> v;

  cv_set_lnos (h, lnos, 1.0);
  TRY (cv_eval_rows (h, h->n_k, h->n_obs - h->n_k, v));

  double m2lnp = 0.0;

  for (double x : v)
    m2lnp += x;

  if (h->print_fit)
  {
    printf ("# over-smooth: % 22.15g, m2lnp = % 22.15g\n", h->over_smooth, m2lnp);
    fflush (stdout);
  }

  val = m2lnp;

  return NCKDE_OK;
}

// _ncm_stats_dist_amise_kde_gauss (ncm_stats_dist.c:512-559): the integral
// of the squared density is the IM sum at sqrt(2) href, the leave-one-out
// term the off-diagonal IM sum at href
static int
cv_obj_amise_kde_gauss (nckde *h, double lnos, double &val)
{
  const size_t n = h->n_k;
  const double n2 = (double) n * (double) n;
  std::vector<double> off, dg;
  double amise = 0.0;

  cv_set_lnos (h, lnos, std::sqrt (2.0));
  TRY (cv_im_sums (h, off, dg));

  for (size_t i = 0; i < n; ++i)
    amise += (off[i] + dg[i]) / n2;

  cv_set_lnos (h, lnos, 1.0);
  TRY (cv_im_sums (h, off, dg));

  for (size_t i = 0; i < n; ++i)
    amise -= 2.0 * off[i] / cv_pairs (n);

  if (h->print_fit)
  {
    printf ("# over-smooth: % 22.15g, amise = % 22.15g\n", h->over_smooth, amise);
    fflush (stdout);
  }

  val = amise;

  return NCKDE_OK;
}

// antithetic pairs per device batch of the Monte-Carlo integral below; pair t
// uses draws 2t and 2t+1 (mod the batch) of Philox stream (key 0, batch)
static constexpr int CV_MC_BATCH = 4096;

// _ncm_stats_dist_amise (ncm_stats_dist.c:563-660): leave-one-out term from
// the IM row sums; the integral of the squared density by Monte Carlo over
// antithetic kernel pairs (kernels ordered by their leave-one-out density),
// 100 warm-up pairs, then until the relative standard error of the mean
// density falls below 1e-2.  Every call restarts the generator from the same
// seed, as the reference does.
static int
cv_obj_amise (nckde *h, double lnos, double &val)
{
  const size_t n = h->n_k;
  const int d    = h->d;
  std::vector<double> off, dg;
  double amise = 0.0;

  cv_set_lnos (h, lnos, 1.0);
  TRY (cv_im_sums (h, off, dg));

  std::vector<double> lp (n);

  for (size_t i = 0; i < n; ++i)
  {
    amise -= 2.0 * off[i] / cv_pairs (n);
    lp[i] = (off[i] + dg[i]) / n;
  }

  std::vector<uint32_t> order (n);

  std::iota (order.begin (), order.end (), 0u);
  std::stable_sort (order.begin (), order.end (), [&] (uint32_t a, uint32_t b) { return lp[a] < lp[b]; });

  const int nu_int = (h->kernel == NCKDE_KERNEL_ST) ? (int) h->nu : 0;

  if (h->kernel == NCKDE_KERNEL_ST && (double) nu_int != h->nu)
  {
    set_error ("nckde: the leave-one-out Monte-Carlo integral needs an integer Student-t nu");
    return NCKDE_EUNSUPPORTED;
  }

  const int m2 = 2 * CV_MC_BATCH;

  HIPCHK (h->dorder.ensure (sizeof (uint32_t) * n));
  HIPCHK (hipMemcpyAsync (h->dorder.p, order.data (), sizeof (uint32_t) * n, hipMemcpyHostToDevice, h->stream));
  TRY (ensure_wcum (h));
  HIPCHK (h->du.ensure (sizeof (double) * m2));
  HIPCHK (h->dz.ensure (sizeof (double) * m2 * d));
  HIPCHK (h->dchi2.ensure (sizeof (double) * m2));
  HIPCHK (h->didx.ensure (sizeof (uint32_t) * m2));
  HIPCHK (h->dcv2.ensure (sizeof (double) * (size_t) m2 * d));
  HIPCHK (h->dcv1.ensure (sizeof (double) * m2));

  const uint64_t max_iter = 100000000;
  std::vector<double> p (m2);
  nck::PairStats st;
  uint64_t count = 0;
  double mean    = 0.0;
  bool done      = false;

  for (uint64_t b = 0; !done; ++b)
  {
    HIPCHK (nck::launch_draws (m2, d, nu_int, 0, b, h->du.as<double> (), h->dz.as<double> (),
                               nu_int > 0 ? h->dchi2.as<double> () : nullptr, h->stream));
    HIPCHK (nck::launch_pair_index (h->dwcum.as<double> (), (int) n, h->du.as<double> (), CV_MC_BATCH,
                                    h->dorder.as<uint32_t> (), h->didx.as<uint32_t> (), h->stream));
    HIPCHK (nck::launch_sample_at (h->didx.as<uint32_t> (), h->dz.as<double> (), h->dchi2.as<double> (), m2, d,
                                   h->kind, h->kernel, h->nu, h->href,
                                   (h->kind == NCKDE_VKDE) ? h->dU.as<double> () : h->dUg.as<double> (),
                                   h->dX.as<double> (), h->dcv2.as<double> (), h->stream));
    TRY (eval_device (h, h->dcv2.as<double> (), m2, h->dcv1.as<double> (), NCKDE_PREC_FP64, 1, h->stream));
    HIPCHK (hipMemcpyAsync (p.data (), h->dcv1.p, sizeof (double) * m2, hipMemcpyDeviceToHost, h->stream));
    HIPCHK (hipStreamSynchronize (h->stream));

    for (int t = 0; t < CV_MC_BATCH && !done; ++t)
    {
      nck::pair_stats_update (st, p[2 * t], p[2 * t + 1]);
      count++;

      if (count <= 100)
        continue;

      const uint64_t i = count - 101;

      mean = 0.5 * (st.mean[0] + st.mean[1]);

      const double var = 0.25 * (st.var[0] * st.bias + st.var[1] * st.bias + 2.0 * (st.cov01 * st.bias));
      const double msd = std::sqrt (var / (i + 101.0)) / mean;

      if (msd < 1.0e-2 || i + 1 >= max_iter)
        done = true;
    }
  }

  amise += mean;

  if (h->print_fit)
  {
    printf ("# over-smooth: % 22.15g, amise = % 22.15g\n", h->over_smooth, amise);
    fflush (stdout);
  }

  val = amise;

  return NCKDE_OK;
}

// _ncm_stats_dist_minimize_obj (ncm_stats_dist.c:664-701): Nelder-Mead over
// ln(over_smooth) from the current value with step 0.1, until the simplex
// size is below 1e-3 or 1000 iterations.  over_smooth and href keep the
// values of the last evaluation, as in the reference.
static int
cv_prepare_minimize (nckde *h)
{
  int (*obj) (nckde *, double, double &) = cv_obj_m2lnp;

  if (h->cv == NCKDE_CV_LOO)
    obj = (h->kind == NCKDE_KDE && h->kernel == NCKDE_KERNEL_GAUSS) ? cv_obj_amise_kde_gauss : cv_obj_amise;

  std::vector<double> x {std::log (h->over_smooth)}, step {0.1};
  nck::NMInfo info {};
  nck::ObjFn f = [&] (const double *v, double &val) {
    h->cv_nobj++;
    return obj (h, v[0], val);
  };
  const auto t0 = std::chrono::steady_clock::now ();

  h->cv_nobj = 0;
  TRY (nck::nm_minimize (f, x, step, 1.0e-3, 1000, &info));
  h->cv_seconds = std::chrono::duration<double> (std::chrono::steady_clock::now () - t0).count ();

  if (h->print_fit)
  {
    printf ("# iter: %d, over-smooth: % 22.15g, m2lnp = % 22.15g, gsl status (%d)\n", info.iter, h->over_smooth,
            info.fval, info.status == 0 ? 0 : (info.status == 1 ? -2 : 5));
    fflush (stdout);
  }

  return NCKDE_OK;
}

// IM scaled by 1/f at over_smooth = exp(lnos) and its NNLS fit (weights in
// dw, unnormalised): _ncm_stats_dist_compute_IM_full + NCM_NNLS_SOLVE
static int
cv_fit_at (nckde *h, double lnos, double min_m2lnp, double &rnorm)
{
  h->cv_nobj++;
  cv_set_lnos (h, lnos, 1.0);
  TRY (build_IM (h, h->dm2lnp.as<double> (), min_m2lnp));
  TRY (nnls_solve (h));
  rnorm = h->rnorm;

  return NCKDE_OK;
}

// CV_SPLIT branch of _ncm_stats_dist_prepare_interp (ncm_stats_dist.c:1018-1073):
// ten Gaussian restarts of ln(over_smooth) (sigma 0.5) keep the smallest NNLS
// residual; then levmar fits ln(over_smooth) to the residuals
// expm1(-(m2lnp_interp_i - (m2lnp_i - min))/2) over every observation; the
// final NNLS fit stays in the weights.  The restarts draw from a per-handle
// counter-based normal stream (key 0), where the reference uses its own
// mt19937 stream seeded with 0.
static int
cv_split_fit (nckde *h, const double *m2lnp, double min_m2lnp)
{
  const size_t no = h->n_obs;
  double ln_os    = std::log (h->over_smooth);
  double rnorm0, rnorm_try;
  const auto t0 = std::chrono::steady_clock::now ();

  h->cv_nobj = 0;
  TRY (cv_fit_at (h, ln_os, min_m2lnp, rnorm0));

  for (int i = 0; i < 10; ++i)
  {
    const double ln_os_try = ln_os + 0.5 * nck::host_normal (0, h->cv_draws++);

    TRY (cv_fit_at (h, ln_os_try, min_m2lnp, rnorm_try));

    if (rnorm_try < rnorm0)
    {
      ln_os  = ln_os_try;
      rnorm0 = rnorm_try;
    }
  }

  std::vector<double> v;
  nck::ResFn f = [&] (double lp, double *hx) -> int {
    double rn;

    TRY (cv_fit_at (h, lp, min_m2lnp, rn));
    TRY (refresh_bvec (h));
    TRY (cv_eval_rows (h, 0, no, v));

    double fn2 = 0.0;

    for (size_t i = 0; i < no; ++i)
    {
      hx[i] = std::expm1 (-0.5 * (v[i] - (m2lnp[i] - min_m2lnp)));
      fn2 += hx[i] * hx[i];
    }

    if (h->print_fit)
    {
      printf ("# over-smooth: % 22.15g, rnorm = % 22.15g, fnorm = % 22.15g\n", h->over_smooth, rn, std::sqrt (fn2));
      fflush (stdout);
    }

    return NCKDE_OK;
  };
  const double opts[5] = {1.0e-3, 1.0e-7, 1.0e-7, 1.0e-10, 1.0e-6};
  nck::LMInfo info {};

  TRY (nck::lm_dif1 (f, ln_os, (int) no, 10000, opts, &info));

  double rn;

  TRY (cv_fit_at (h, ln_os, min_m2lnp, rn));
  h->cv_seconds = std::chrono::duration<double> (std::chrono::steady_clock::now () - t0).count ();

  return NCKDE_OK;
}

// ========================================================================
// C-ABI
// ========================================================================
// serialisation helpers (byte writer / bounds-checked reader)
namespace
{

struct Writer
{
  unsigned char *buf;
  size_t cap, len = 0;

  void
  put (const void *src, size_t n)
  {
    if (buf != nullptr && len + n <= cap)
      std::memcpy (buf + len, src, n);

    len += n;
  }

  template <class T>
  void
  val (T v)
  {
    put (&v, sizeof (T));
  }
};

struct Reader
{
  const unsigned char *buf;
  size_t len, pos = 0;
  bool ok = true;

  void
  get (void *dst, size_t n)
  {
    if (!ok || pos + n > len)
    {
      ok = false;
      return;
    }

    std::memcpy (dst, buf + pos, n);
    pos += n;
  }

  template <class T>
  T
  val ()
  {
    T v {};

    get (&v, sizeof (T));

    return v;
  }

  size_t
  left () const
  {
    return ok ? len - pos : 0;
  }
};

const char SER_MAGIC[8] = {'N', 'C', 'K', 'D', 'E', '0', '2', '\0'};

} // namespace

extern "C" {

const char *
nckde_last_error (void)
{
  return nck::g_err;
}

int
nckde_create (nckde **out, int kind, int kernel, double nu, unsigned dim, int cv_type)
{
  if (out == nullptr)
  {
    set_error ("nckde_create: NULL output pointer");
    return NCKDE_EINVAL;
  }

  *out = nullptr;

  if ((kind != NCKDE_KDE && kind != NCKDE_VKDE) || (kernel != NCKDE_KERNEL_GAUSS && kernel != NCKDE_KERNEL_ST) ||
      dim == 0)
  {
    set_error ("nckde_create: invalid kind/kernel/dimension");
    return NCKDE_EINVAL;
  }

  if (kernel == NCKDE_KERNEL_ST && !(nu >= 1.0))
  {
    set_error ("nckde_create: Student-t nu must be >= 1");
    return NCKDE_EINVAL;
  }

  if (cv_type < NCKDE_CV_NONE || cv_type > NCKDE_CV_LOO)
  {
    set_error ("nckde_create: invalid cross-validation type %d", cv_type);
    return NCKDE_EINVAL;
  }

  int ndev = 0;

  if (hipGetDeviceCount (&ndev) != hipSuccess || ndev == 0)
  {
    set_error ("nckde_create: no HIP device available");
    return NCKDE_EHIP;
  }

  nckde *h = new nckde ();

  h->kind   = kind;
  h->kernel = kernel;
  h->cv     = cv_type;
  h->nu     = (kernel == NCKDE_KERNEL_ST) ? nu : 0.0;
  h->d      = (int) dim;

  if (hipGetDevice (&h->device) != hipSuccess)
  {
    delete h;
    set_error ("nckde_create: hipGetDevice failed");
    return NCKDE_EHIP;
  }

  if (hipStreamCreateWithFlags (&h->stream, hipStreamNonBlocking) != hipSuccess)
  {
    set_error ("nckde_create: hipStreamCreate failed");
    delete h;
    return NCKDE_EHIP;
  }

  if (hipEventCreateWithFlags (&h->ready_ev, hipEventDisableTiming) != hipSuccess)
  {
    set_error ("nckde_create: hipEventCreate failed");
    (void) hipStreamDestroy (h->stream);
    delete h;
    return NCKDE_EHIP;
  }

  *out = h;

  return NCKDE_OK;
}

int
nckde_destroy (nckde *h)
{
  if (h == nullptr)
    return NCKDE_OK;

  if (h->stream)
  {
    (void) hipStreamSynchronize (h->stream);
    (void) hipStreamDestroy (h->stream);
  }

  for (hipEvent_t &e : h->pev)
    if (e)
      (void) hipEventDestroy (e);

  if (h->ready_ev)
    (void) hipEventDestroy (h->ready_ev);

  for (auto &g : h->graph_cache)
    if (g.exec)
      (void) hipGraphExecDestroy (g.exec);

  delete h;

  return NCKDE_OK;
}

int
nckde_set_over_smooth (nckde *h, double v)
{
  TRY (check_handle (h));

  if (!(v >= 1.0e-5))
  {
    set_error ("nckde_set_over_smooth: value must be >= 1e-5");
    return NCKDE_EINVAL;
  }

  h->over_smooth = v;

  return NCKDE_OK;
}

int
nckde_set_shrink (nckde *h, double v)
{
  TRY (check_handle (h));

  if (!(v >= 0.0 && v <= 1.0))
  {
    set_error ("nckde_set_shrink: shrink must be in [0, 1]");
    return NCKDE_EINVAL;
  }

  h->shrink = v;

  return NCKDE_OK;
}

int
nckde_get_over_smooth (nckde *h, double *v)
{
  TRY (check_handle (h));

  if (v == nullptr)
  {
    set_error ("nckde_get_over_smooth: NULL output pointer");
    return NCKDE_EINVAL;
  }

  *v = h->over_smooth;

  return NCKDE_OK;
}

int
nckde_get_cv_stats (nckde *h, unsigned *n_objective, double *seconds)
{
  TRY (check_handle (h));

  if (n_objective)
    *n_objective = h->cv_nobj;

  if (seconds)
    *seconds = h->cv_seconds;

  return NCKDE_OK;
}

int
nckde_set_eval_split (nckde *h, unsigned nsplit)
{
  TRY (check_handle (h));
  std::lock_guard<std::mutex> lk (h->mtx);

  h->split_override = nsplit;
  h->graph_gen++;

  return NCKDE_OK;
}

int
nckde_set_graphs (nckde *h, int on)
{
  TRY (check_handle (h));
  std::lock_guard<std::mutex> lk (h->mtx);

  h->graphs = (on != 0);
  h->graph_gen++;

  return NCKDE_OK;
}

int
nckde_set_mixed_tolerance (nckde *h, double tol)
{
  TRY (check_handle (h));

  if (!(tol > 0.0))
  {
    set_error ("nckde_set_mixed_tolerance: the tolerance must be positive");
    return NCKDE_EINVAL;
  }

  std::lock_guard<std::mutex> lk (h->mtx);
  h->mixed_tol = tol;

  // a probe already made is judged again against the new tolerance
  if (h->mixed_ready)
    h->mixed_fallback = !(h->mixed_probe_err <= tol);

  return NCKDE_OK;
}

int
nckde_get_mixed_check (nckde *h, double *max_rel_err, int *fallback)
{
  TRY (check_handle (h));
  TRY (check_prepared (h));
  std::lock_guard<std::mutex> lk (h->mtx);

  TRY (ensure_mixed (h));

  if (max_rel_err)
    *max_rel_err = h->mixed_probe_err;

  if (fallback)
    *fallback = h->mixed_fallback ? 1 : 0;

  return NCKDE_OK;
}

int
nckde_set_nnls_umethod (nckde *h, int v)
{
  TRY (check_handle (h));

  if (v == 3 || v == 4)
  {
    set_error ("nckde_set_nnls_umethod: the DGELSD and GSL methods are not implemented");
    return NCKDE_EUNSUPPORTED;
  }

  if (v < NCKDE_NNLS_NORMAL || v > NCKDE_NNLS_QR)
  {
    set_error ("nckde_set_nnls_umethod: unknown method %d", v);
    return NCKDE_EINVAL;
  }

  std::lock_guard<std::mutex> lk (h->mtx);
  h->nnls_umethod = v;

  return NCKDE_OK;
}

int
nckde_get_nnls_umethod (nckde *h, int *v)
{
  TRY (check_handle (h));

  if (v == nullptr)
  {
    set_error ("nckde_get_nnls_umethod: NULL output");
    return NCKDE_EINVAL;
  }

  *v = h->nnls_umethod;

  return NCKDE_OK;
}

int
nckde_set_nnls_refine (nckde *h, unsigned v)
{
  TRY (check_handle (h));

  if (v > 16)
  {
    set_error ("nckde_set_nnls_refine: at most 16 steps");
    return NCKDE_EINVAL;
  }

  std::lock_guard<std::mutex> lk (h->mtx);
  h->nnls_refine = v;

  return NCKDE_OK;
}

int
nckde_get_nnls_refine (nckde *h, unsigned *v)
{
  TRY (check_handle (h));

  if (v == nullptr)
  {
    set_error ("nckde_get_nnls_refine: NULL output");
    return NCKDE_EINVAL;
  }

  *v = h->nnls_refine;

  return NCKDE_OK;
}

int
nckde_get_nnls_rungs (nckde *h, unsigned counts[4])
{
  TRY (check_handle (h));

  if (counts == nullptr)
  {
    set_error ("nckde_get_nnls_rungs: NULL output");
    return NCKDE_EINVAL;
  }

  for (int i = 0; i < 4; ++i)
    counts[i] = h->nnls_rungs[i];

  return NCKDE_OK;
}

int
nckde_set_split_frac (nckde *h, double v)
{
  TRY (check_handle (h));

  if (!(v >= 0.01 && v <= 1.0))
  {
    set_error ("nckde_set_split_frac: split_frac must be in [0.01, 1]");
    return NCKDE_EINVAL;
  }

  h->split_frac = v;

  return NCKDE_OK;
}

int
nckde_get_split_frac (nckde *h, double *v)
{
  TRY (check_handle (h));

  if (v == nullptr)
  {
    set_error ("nckde_get_split_frac: NULL output pointer");
    return NCKDE_EINVAL;
  }

  *v = h->split_frac;

  return NCKDE_OK;
}

int
nckde_set_print_fit (nckde *h, int v)
{
  TRY (check_handle (h));
  h->print_fit = v != 0;

  return NCKDE_OK;
}

int
nckde_get_print_fit (nckde *h, int *v)
{
  TRY (check_handle (h));

  if (v == nullptr)
  {
    set_error ("nckde_get_print_fit: NULL output pointer");
    return NCKDE_EINVAL;
  }

  *v = h->print_fit ? 1 : 0;

  return NCKDE_OK;
}

int
nckde_set_cv_type (nckde *h, int v)
{
  TRY (check_handle (h));

  if (v < NCKDE_CV_NONE || v > NCKDE_CV_LOO)
  {
    set_error ("nckde_set_cv_type: invalid cross-validation type %d", v);
    return NCKDE_EINVAL;
  }

  h->cv = v;

  return NCKDE_OK;
}

int
nckde_get_cv_type (nckde *h, int *v)
{
  TRY (check_handle (h));

  if (v == nullptr)
  {
    set_error ("nckde_get_cv_type: NULL output pointer");
    return NCKDE_EINVAL;
  }

  *v = h->cv;

  return NCKDE_OK;
}

int
nckde_set_local_frac (nckde *h, double v)
{
  TRY (check_handle (h));

  if (!(v >= 0.001 && v <= 1.0))
  {
    set_error ("nckde_set_local_frac: local_frac must be in [0.001, 1]");
    return NCKDE_EINVAL;
  }

  h->local_frac = v;

  return NCKDE_OK;
}

int
nckde_set_use_rot_href (nckde *h, int v)
{
  TRY (check_handle (h));
  h->use_rot_href = v ? 1 : 0;

  return NCKDE_OK;
}

int
nckde_set_nearPD_maxiter (nckde *h, unsigned v)
{
  TRY (check_handle (h));
  h->nearPD_maxiter = v;

  return NCKDE_OK;
}

int
nckde_set_use_threads (nckde *h, int v)
{
  TRY (check_handle (h));
  std::lock_guard<std::mutex> lk (h->mtx);

  h->use_threads = (v != 0);

  return NCKDE_OK;
}

int
nckde_get_use_threads (nckde *h, int *v)
{
  TRY (check_handle (h));

  if (v == nullptr)
  {
    set_error ("nckde_get_use_threads: NULL output");
    return NCKDE_EINVAL;
  }

  std::lock_guard<std::mutex> lk (h->mtx);

  *v = h->use_threads ? 1 : 0;

  return NCKDE_OK;
}

int
nckde_get_shrink (nckde *h, double *v)
{
  TRY (check_handle (h));

  if (v == nullptr)
  {
    set_error ("nckde_get_shrink: NULL output");
    return NCKDE_EINVAL;
  }

  std::lock_guard<std::mutex> lk (h->mtx);

  *v = h->shrink;

  return NCKDE_OK;
}

int
nckde_get_local_frac (nckde *h, double *v)
{
  TRY (check_handle (h));

  if (v == nullptr)
  {
    set_error ("nckde_get_local_frac: NULL output");
    return NCKDE_EINVAL;
  }

  std::lock_guard<std::mutex> lk (h->mtx);

  *v = h->local_frac;

  return NCKDE_OK;
}

int
nckde_get_use_rot_href (nckde *h, int *v)
{
  TRY (check_handle (h));

  if (v == nullptr)
  {
    set_error ("nckde_get_use_rot_href: NULL output");
    return NCKDE_EINVAL;
  }

  std::lock_guard<std::mutex> lk (h->mtx);

  *v = h->use_rot_href ? 1 : 0;

  return NCKDE_OK;
}

int
nckde_get_nearPD_maxiter (nckde *h, unsigned *v)
{
  TRY (check_handle (h));

  if (v == nullptr)
  {
    set_error ("nckde_get_nearPD_maxiter: NULL output");
    return NCKDE_EINVAL;
  }

  std::lock_guard<std::mutex> lk (h->mtx);

  *v = h->nearPD_maxiter;

  return NCKDE_OK;
}

int
nckde_get_cov_type (nckde *h, int *v)
{
  TRY (check_handle (h));

  if (v == nullptr)
  {
    set_error ("nckde_get_cov_type: NULL output");
    return NCKDE_EINVAL;
  }

  std::lock_guard<std::mutex> lk (h->mtx);

  *v = h->cov_type;

  return NCKDE_OK;
}

int
nckde_get_cov_fixed (nckde *h, double *cov)
{
  TRY (check_handle (h));
  std::lock_guard<std::mutex> lk (h->mtx);

  if (!h->has_cov_fixed)
  {
    set_error ("nckde_get_cov_fixed: no fixed covariance was set");
    return NCKDE_ESTATE;
  }

  if (cov == nullptr)
  {
    set_error ("nckde_get_cov_fixed: NULL output");
    return NCKDE_EINVAL;
  }

  std::copy (h->cov_fixed.begin (), h->cov_fixed.end (), cov);

  return NCKDE_OK;
}

// ncm_stats_dist_kde_set_cov_type / _set_cov_fixed (ncm_stats_dist_kde.c:
// 731-745, 770-791): a fixed covariance is factorised when it is set and
// must be positive definite.
static int
decompose_fixed (nckde *h)
{
  const int d = h->d;
  DBuf dc, du, dl, di;
  int info      = 0;
  hipStream_t s = h->stream;

  HIPCHK (dc.ensure (sizeof (double) * d * d));
  HIPCHK (du.ensure (sizeof (double) * d * d));
  HIPCHK (dl.ensure (sizeof (double)));
  HIPCHK (di.ensure (sizeof (int)));
  HIPCHK (hipMemcpyAsync (dc.p, h->cov_fixed.data (), sizeof (double) * d * d, hipMemcpyHostToDevice, s));
  HIPCHK (nck::launch_chol_batched (dc.as<double> (), du.as<double> (), di.as<int> (), dl.as<double> (), d, 1,
                                    h->kernel, (h->kernel == NCKDE_KERNEL_ST) ? st_lnnorm_const (h) : 0.0, s));
  h->cov_fixed_decomp.resize (d * d);
  HIPCHK (hipMemcpyAsync (&info, di.p, sizeof (int), hipMemcpyDeviceToHost, s));
  HIPCHK (hipMemcpyAsync (h->cov_fixed_decomp.data (), du.p, sizeof (double) * d * d, hipMemcpyDeviceToHost, s));
  HIPCHK (hipMemcpyAsync (&h->cov_fixed_lnnorm, dl.p, sizeof (double), hipMemcpyDeviceToHost, s));
  HIPCHK (hipStreamSynchronize (s));

  if (info != 0)
  {
    set_error ("ncm_stats_dist_kde_set_cov_fixed: matrix cov_fixed is not positive definite.");
    return NCKDE_ENUMERIC;
  }

  return NCKDE_OK;
}

int
nckde_set_cov_type (nckde *h, int cov_type)
{
  TRY (check_handle (h));

  if (cov_type < NCKDE_COV_SAMPLE || cov_type > NCKDE_COV_ROBUST)
  {
    set_error ("nckde_set_cov_type: invalid covariance type");
    return NCKDE_EINVAL;
  }

  std::lock_guard<std::mutex> lk (h->mtx);

  h->cov_type = cov_type;

  if (cov_type == NCKDE_COV_FIXED && h->has_cov_fixed)
    return decompose_fixed (h);

  return NCKDE_OK;
}

int
nckde_set_cov_fixed (nckde *h, const double *cov)
{
  TRY (check_handle (h));

  if (cov == nullptr)
  {
    set_error ("nckde_set_cov_fixed: NULL matrix");
    return NCKDE_EINVAL;
  }

  std::lock_guard<std::mutex> lk (h->mtx);

  h->cov_fixed.assign (cov, cov + h->d * h->d);
  h->has_cov_fixed = true;

  if (h->cov_type == NCKDE_COV_FIXED)
    return decompose_fixed (h);

  return NCKDE_OK;
}

int
nckde_add_obs (nckde *h, const double *x)
{
  return nckde_add_obs_batch (h, x, 1, h ? (size_t) h->d : 0);
}

int
nckde_add_obs_batch (nckde *h, const double *X, size_t n, size_t ld)
{
  TRY (check_handle (h));

  if ((X == nullptr && n > 0) || ld < (size_t) h->d)
  {
    set_error ("nckde_add_obs: invalid arguments");
    return NCKDE_EINVAL;
  }

  std::lock_guard<std::mutex> lk (h->mtx);

  for (size_t i = 0; i < n; ++i)
    h->X.insert (h->X.end (), X + i * ld, X + i * ld + h->d);

  h->n_samples += n;

  return NCKDE_OK;
}

int
nckde_reset (nckde *h)
{
  TRY (check_handle (h));
  std::lock_guard<std::mutex> lk (h->mtx);

  h->X.clear ();
  h->n_samples = 0;

  return NCKDE_OK;
}

int
nckde_prepare (nckde *h)
{
  TRY (check_handle (h));
  std::lock_guard<std::mutex> lk (h->mtx);

  return do_prepare (h);
}

int
nckde_prepare_interp (nckde *h, const double *m2lnp, size_t n)
{
  TRY (check_handle (h));

  if (m2lnp == nullptr)
  {
    set_error ("nckde_prepare_interp: NULL m2lnp");
    return NCKDE_EINVAL;
  }

  std::lock_guard<std::mutex> lk (h->mtx);

  for (unsigned &c : h->nnls_rungs)
    c = 0;

  return do_prepare_interp (h, m2lnp, n);
}

int
nckde_compute_IM (nckde *h, double *IM)
{
  TRY (check_prepared (h));
  std::lock_guard<std::mutex> lk (h->mtx);
  DBuf tmp;

  TRY (build_IM (h, nullptr, 0.0));
  HIPCHK (tmp.ensure (sizeof (double) * h->n_obs * h->n_k));
  HIPCHK (nck::launch_transpose (h->dAT.as<double> (), (int) h->n_k, (int) h->n_obs, tmp.as<double> (), h->stream));
  HIPCHK (hipMemcpyAsync (IM, tmp.p, sizeof (double) * h->n_obs * h->n_k, hipMemcpyDeviceToHost, h->stream));
  HIPCHK (hipStreamSynchronize (h->stream));

  return NCKDE_OK;
}

int
nckde_eval_m2lnp (nckde *h, const double *X, size_t m, size_t ld, double *out, int precision)
{
  TRY (check_prepared (h));
  std::lock_guard<std::mutex> lk (h->mtx);

  return eval_host (h, X, m, ld, out, precision, 0);
}

int
nckde_eval (nckde *h, const double *X, size_t m, size_t ld, double *out, int precision)
{
  TRY (check_prepared (h));
  std::lock_guard<std::mutex> lk (h->mtx);

  return eval_host (h, X, m, ld, out, precision, 1);
}

int
nckde_eval_m2lnp_device (nckde *h, const double *dX, size_t m, double *dout, int precision, void *stream)
{
  TRY (check_prepared (h));
  std::lock_guard<std::mutex> lk (h->mtx);

  return eval_device (h, dX, m, dout, precision, 0, stream ? (hipStream_t) stream : h->stream);
}

// ---- allocation-free eval (no reference counterpart) ---------------------
// For callers that cannot allocate device memory (code under graph capture,
// tight serving loops): the eval's scratch -- point work and split partials
// -- lives in this file-scope device array, of which each GPU holds one copy
// for the life of the process.  Its device address is taken once per device
// with hipGetSymbolAddress (the array's name is only a host-side stand-in)
// and every call on that device shares it.
constexpr size_t SCRATCH_BYTES = (size_t) 64 << 20;
constexpr int SCRATCH_MAX_DEVICES = 64;

static __device__ __align__ (256) unsigned char scratch[SCRATCH_BYTES];
static std::mutex scratch_mtx;
static void *scratch_ptr[SCRATCH_MAX_DEVICES];

static int
scratch_on_device (int dev, char **out)
{
  if (dev < 0 || dev >= SCRATCH_MAX_DEVICES)
  {
    set_error ("nckde_eval_m2lnp_device_noalloc: device %d out of range", dev);
    return NCKDE_EINVAL;
  }

  std::lock_guard<std::mutex> lk (scratch_mtx);

  if (scratch_ptr[dev] == nullptr)
  {
    void *p        = nullptr;
    const hipError_t e = hipGetSymbolAddress (&p, HIP_SYMBOL (scratch));

    if (e != hipSuccess || p == nullptr)
    {
      set_error ("nckde_eval_m2lnp_device_noalloc: hipGetSymbolAddress failed: %s", hipGetErrorString (e));
      return NCKDE_EHIP;
    }

    scratch_ptr[dev] = p;
  }

  *out = (char *) scratch_ptr[dev];

  return NCKDE_OK;
}

size_t
nckde_noalloc_scratch_bytes (void)
{
  return SCRATCH_BYTES;
}

int
nckde_eval_m2lnp_device_noalloc (nckde *h, const double *dX, size_t m, double *dout, int precision, void *stream)
{
  TRY (check_handle (h));
  TRY (check_prepared (h));

  if (m == 0)
    return NCKDE_OK;

  if (dX == nullptr || dout == nullptr)
  {
    set_error ("nckde_eval_m2lnp_device_noalloc: NULL points or output");
    return NCKDE_EINVAL;
  }

  if (precision != NCKDE_PREC_FP64 && precision != NCKDE_PREC_MIXED && precision != NCKDE_PREC_FP32)
  {
    set_error ("nckde: unknown precision %d", precision);
    return NCKDE_EINVAL;
  }

  std::lock_guard<std::mutex> lk (h->mtx);
  hipStream_t s = stream ? (hipStream_t) stream : h->stream;
  int dev       = -1;

  HIPCHK (hipGetDevice (&dev));

  if (dev != h->device)
  {
    set_error ("nckde_eval_m2lnp_device_noalloc: the handle lives on device %d, device %d is current", h->device,
               dev);
    return NCKDE_EINVAL;
  }

  if (precision == NCKDE_PREC_MIXED)
  {
    if (!h->mixed_ready || !h->mixed_b_ready)
    {
      set_error ("nckde_eval_m2lnp_device_noalloc: the mixed-precision factors are not built; an allocating "
                 "mixed eval or nckde_get_mixed_check after the prepare builds them");
      return NCKDE_ESTATE;
    }

    if (h->mixed_fallback)
      precision = NCKDE_PREC_FP64;
  }

  if (precision == NCKDE_PREC_FP32 && (!h->f32_ready || !h->f32_b_ready))
  {
    set_error ("nckde_eval_m2lnp_device_noalloc: the FP32 factors are not built; an allocating FP32 eval after "
               "the prepare builds them");
    return NCKDE_ESTATE;
  }

  size_t need;

  if (precision == NCKDE_PREC_MIXED)
  {
    size_t sz[4];

    mixed_work_bytes (h, m, sz);
    need = align256 (sz[0]) + align256 (sz[1]) + align256 (sz[2]) + sz[3];
  }
  else
  {
    const size_t part = sizeof (double) * m * fp64_nsplit (h, m, precision);

    need = align256 (part) + part;
  }

  if (need > SCRATCH_BYTES)
  {
    set_error ("nckde_eval_m2lnp_device_noalloc: %zu points need %zu bytes of scratch, the maximum is %zu; "
               "evaluate fewer points per call", m, need, SCRATCH_BYTES);
    return NCKDE_EINVAL;
  }

  char *ws = nullptr;

  TRY (scratch_on_device (dev, &ws));

  if (s != h->stream)
  {
    HIPCHK (hipEventRecord (h->ready_ev, h->stream));
    HIPCHK (hipStreamWaitEvent (s, h->ready_ev, 0));
  }

  if (precision == NCKDE_PREC_MIXED)
    return eval_mixed_ws (h, dX, m, dout, 0, s, ws);

  return eval_fp64_ws (h, dX, m, dout, 0, s, ws, precision == NCKDE_PREC_FP32);
}

// ---- several handles' evals in fused launches (no reference counterpart) -
// Mixed-precision evals of handles that share the MFMA dimension and kernel
// type run as one launch per pass over a segment table; every segment's
// blocks compute what the handle's own eval computes, so the results are
// bit-identical to separate calls.  Other precisions, and handles whose
// mixed factors fall back to FP64, run one by one.
int
nckde_eval_m2lnp_multi (nckde *const *hs, const double *const *X, const size_t *m, const size_t *ld,
                        double *const *out, size_t count, int precision)
{
  if (count == 0)
    return NCKDE_OK;

  if (hs == nullptr || X == nullptr || m == nullptr || out == nullptr)
  {
    set_error ("nckde_eval_m2lnp_multi: NULL argument");
    return NCKDE_EINVAL;
  }

  if (precision != NCKDE_PREC_FP64 && precision != NCKDE_PREC_MIXED && precision != NCKDE_PREC_FP32)
  {
    set_error ("nckde: unknown precision %d", precision);
    return NCKDE_EINVAL;
  }

  std::vector<nckde *> order (hs, hs + count);

  for (size_t i = 0; i < count; ++i)
  {
    TRY (check_handle (hs[i]));
    TRY (check_prepared (hs[i]));

    const size_t ldi = ld ? ld[i] : (size_t) hs[i]->d;

    if (m[i] > 0 && (X[i] == nullptr || out[i] == nullptr || ldi < (size_t) hs[i]->d))
    {
      set_error ("nckde_eval_m2lnp_multi: invalid buffers for evaluation %zu", i);
      return NCKDE_EINVAL;
    }
  }

  // the handles' locks in address order (no handle twice)
  std::sort (order.begin (), order.end ());

  if (std::adjacent_find (order.begin (), order.end ()) != order.end ())
  {
    set_error ("nckde_eval_m2lnp_multi: a handle appears twice");
    return NCKDE_EINVAL;
  }

  std::vector<std::unique_lock<std::mutex>> locks;

  for (nckde *h : order)
    locks.emplace_back (h->mtx);

  // groups of fusable evals by (MFMA dimension, kernel type)
  std::vector<std::vector<size_t>> groups;
  std::vector<int> keys;

  for (size_t i = 0; i < count; ++i)
  {
    nckde *h = hs[i];

    if (m[i] == 0)
      continue;

    bool fuse = (precision == NCKDE_PREC_MIXED);

    if (fuse)
    {
      TRY (ensure_mixed (h));
      fuse = !h->mixed_fallback && h->pilot_n == 0;
    }

    if (!fuse)
    {
      TRY (eval_host (h, X[i], m[i], ld ? ld[i] : (size_t) h->d, out[i], precision, 0));
      continue;
    }

    const int key = nck::mfma_dim (h->d) * 2 + h->kernel;
    size_t g      = 0;

    while (g < keys.size () && keys[g] != key)
      ++g;

    if (g == keys.size ())
    {
      keys.push_back (key);
      groups.emplace_back ();
    }

    groups[g].push_back (i);
  }

  for (size_t g = 0; g < groups.size (); ++g)
  {
    const std::vector<size_t> &mem = groups[g];
    nckde *h0                      = hs[mem[0]];
    hipStream_t S                  = h0->stream;
    const int nseg                 = (int) mem.size ();
    std::vector<nck::MixedArgs> args (nseg);
    std::vector<int> block0 (2 * (nseg + 1), 0);

    for (int k = 0; k < nseg; ++k)
    {
      nckde *h         = hs[mem[k]];
      const size_t mi  = m[mem[k]];
      const size_t ldi = ld ? ld[mem[k]] : (size_t) h->d;
      const int d      = h->d;

      // the handle's own work (factor builds) ends before S uses its
      // buffers, and so does an earlier eval on another stream
      if (h->stream != S)
      {
        HIPCHK (hipEventRecord (h->ready_ev, h->stream));
        HIPCHK (hipStreamWaitEvent (S, h->ready_ev, 0));
      }

      if (h->last_stream != nullptr && h->last_stream != S)
        HIPCHK (hipStreamSynchronize (h->last_stream));

      h->last_stream = S;

      HIPCHK (h->dXe.ensure (sizeof (double) * mi * d));
      HIPCHK (h->dout.ensure (sizeof (double) * mi));

      if (ldi == (size_t) d)
        HIPCHK (hipMemcpyAsync (h->dXe.p, X[mem[k]], sizeof (double) * mi * d, hipMemcpyHostToDevice, S));
      else
        HIPCHK (hipMemcpy2DAsync (h->dXe.p, sizeof (double) * d, X[mem[k]], sizeof (double) * ldi,
                                  sizeof (double) * d, mi, hipMemcpyHostToDevice, S));

      TRY (mixed_args (h, h->dXe.as<double> (), mi, h->dout.as<double> (), 0, S, nullptr, args[k]));
      HIPCHK (nck::launch_mixed_points (args[k], S));
      block0[k + 1]            = block0[k] + nck::mixed_eval_blocks (args[k]);
      block0[nseg + 1 + k + 1] = block0[nseg + 1 + k] + nck::mixed_merge_blocks (args[k]);
    }

    HIPCHK (h0->dmulti.ensure (sizeof (nck::MixedArgs) * nseg + sizeof (int) * block0.size ()));

    nck::MixedArgs *dargs = h0->dmulti.as<nck::MixedArgs> ();
    int *dblock0          = (int *) (dargs + nseg);

    HIPCHK (hipMemcpyAsync (dargs, args.data (), sizeof (nck::MixedArgs) * nseg, hipMemcpyHostToDevice, S));
    HIPCHK (hipMemcpyAsync (dblock0, block0.data (), sizeof (int) * block0.size (), hipMemcpyHostToDevice, S));
    HIPCHK (nck::launch_eval_mixed_multi (args[0].f.DP, h0->kernel, dargs, dblock0, nseg, block0[nseg],
                                          block0[2 * nseg + 1], S));

    for (int k = 0; k < nseg; ++k)
      HIPCHK (hipMemcpyAsync (out[mem[k]], hs[mem[k]]->dout.p, sizeof (double) * m[mem[k]], hipMemcpyDeviceToHost,
                              S));

    // the argument table and block offsets stay alive until S is done
    HIPCHK (hipStreamSynchronize (S));
  }

  return NCKDE_OK;
}

int
nckde_kernel_choose (nckde *h, const double *u, size_t m, uint32_t *idx)
{
  TRY (check_prepared (h));
  std::lock_guard<std::mutex> lk (h->mtx);

  if (m == 0)
    return NCKDE_OK;

  TRY (ensure_wcum (h));
  HIPCHK (h->du.ensure (sizeof (double) * m));
  HIPCHK (h->didx.ensure (sizeof (uint32_t) * m));
  HIPCHK (hipMemcpyAsync (h->du.p, u, sizeof (double) * m, hipMemcpyHostToDevice, h->stream));
  HIPCHK (nck::launch_choose (h->dwcum.as<double> (), (int) h->n_k, h->du.as<double> (), (int) m,
                              h->didx.as<uint32_t> (), h->stream));
  HIPCHK (hipMemcpyAsync (idx, h->didx.p, sizeof (uint32_t) * m, hipMemcpyDeviceToHost, h->stream));
  HIPCHK (hipStreamSynchronize (h->stream));

  return NCKDE_OK;
}

int
nckde_sample (nckde *h, const double *u, const double *z, const double *chi2, size_t m, double *out, uint32_t *idx)
{
  TRY (check_prepared (h));
  std::lock_guard<std::mutex> lk (h->mtx);
  const int d = h->d;

  if (m == 0)
    return NCKDE_OK;

  if (h->kernel == NCKDE_KERNEL_ST && chi2 == nullptr)
  {
    set_error ("nckde_sample: the Student-t kernel needs chi2 draws");
    return NCKDE_EINVAL;
  }

  TRY (ensure_wcum (h));
  HIPCHK (h->du.ensure (sizeof (double) * m));
  HIPCHK (h->dz.ensure (sizeof (double) * m * d));
  HIPCHK (h->dchi2.ensure (sizeof (double) * m));
  HIPCHK (h->dout.ensure (sizeof (double) * m * d));
  HIPCHK (h->didx.ensure (sizeof (uint32_t) * m));
  HIPCHK (hipMemcpyAsync (h->du.p, u, sizeof (double) * m, hipMemcpyHostToDevice, h->stream));
  HIPCHK (hipMemcpyAsync (h->dz.p, z, sizeof (double) * m * d, hipMemcpyHostToDevice, h->stream));

  if (chi2)
    HIPCHK (hipMemcpyAsync (h->dchi2.p, chi2, sizeof (double) * m, hipMemcpyHostToDevice, h->stream));

  HIPCHK (nck::launch_sample (h->dwcum.as<double> (), (int) h->n_k, h->du.as<double> (), h->dz.as<double> (),
                              h->dchi2.as<double> (), (int) m, d, h->kind, h->kernel, h->nu, h->href,
                              (h->kind == NCKDE_VKDE) ? h->dU.as<double> () : h->dUg.as<double> (),
                              h->dX.as<double> (), h->dout.as<double> (), h->didx.as<uint32_t> (), h->stream));
  HIPCHK (hipMemcpyAsync (out, h->dout.p, sizeof (double) * m * d, hipMemcpyDeviceToHost, h->stream));

  if (idx)
    HIPCHK (hipMemcpyAsync (idx, h->didx.p, sizeof (uint32_t) * m, hipMemcpyDeviceToHost, h->stream));

  HIPCHK (hipStreamSynchronize (h->stream));

  return NCKDE_OK;
}

int
nckde_get_dim (nckde *h, unsigned *dim)
{
  TRY (check_handle (h));
  *dim = (unsigned) h->d;

  return NCKDE_OK;
}

int
nckde_get_sample_size (nckde *h, size_t *n)
{
  TRY (check_handle (h));
  *n = h->n_obs;

  return NCKDE_OK;
}

int
nckde_get_n_kernels (nckde *h, size_t *n)
{
  TRY (check_handle (h));
  *n = h->n_k;

  return NCKDE_OK;
}

int
nckde_get_n_samples (nckde *h, size_t *n)
{
  TRY (check_handle (h));
  *n = h->n_samples;

  return NCKDE_OK;
}

int
nckde_get_href (nckde *h, double *href)
{
  TRY (check_prepared (h));
  *href = h->href;

  return NCKDE_OK;
}

int
nckde_get_rnorm (nckde *h, double *rnorm)
{
  TRY (check_prepared (h));
  *rnorm = h->rnorm * h->rnorm;

  return NCKDE_OK;
}

int
nckde_get_lnnorm (nckde *h, size_t i, double *lnnorm)
{
  TRY (check_prepared (h));
  std::lock_guard<std::mutex> lk (h->mtx);

  if (i >= h->n_k)
  {
    set_error ("nckde_get_lnnorm: index out of range");
    return NCKDE_EINVAL;
  }

  double l = h->kernel_lnnorm;

  if (h->kind == NCKDE_VKDE)
    HIPCHK (hipMemcpy (&l, h->dlnnorm.as<double> () + i, sizeof (double), hipMemcpyDeviceToHost));

  *lnnorm = l + h->d * std::log (h->href);

  return NCKDE_OK;
}

int
nckde_get_weights (nckde *h, double *w)
{
  TRY (check_prepared (h));
  std::lock_guard<std::mutex> lk (h->mtx);
  HIPCHK (hipMemcpy (w, h->dw.p, sizeof (double) * h->n_k, hipMemcpyDeviceToHost));

  return NCKDE_OK;
}

int
nckde_get_cov_decomp (nckde *h, size_t i, double *U)
{
  TRY (check_prepared (h));
  std::lock_guard<std::mutex> lk (h->mtx);
  const size_t dd = (size_t) h->d * h->d;

  if (i >= h->n_k)
  {
    set_error ("nckde_get_cov_decomp: index out of range");
    return NCKDE_EINVAL;
  }

  if (h->kind == NCKDE_VKDE)
    HIPCHK (hipMemcpy (U, h->dU.as<double> () + i * dd, sizeof (double) * dd, hipMemcpyDeviceToHost));
  else
    HIPCHK (hipMemcpy (U, h->dUg.p, sizeof (double) * dd, hipMemcpyDeviceToHost));

  return NCKDE_OK;
}

int
nckde_get_nnls_state (nckde *h, double *x, double *g, double *r)
{
  TRY (check_prepared (h));
  std::lock_guard<std::mutex> lk (h->mtx);

  if (h->nnls_x.empty ())
  {
    set_error ("nckde_get_nnls_state: the last prepare ran no NNLS fit");
    return NCKDE_ESTATE;
  }

  if (x)
    std::copy (h->nnls_x.begin (), h->nnls_x.end (), x);

  if (g)
    std::copy (h->nnls_g.begin (), h->nnls_g.end (), g);

  if (r)
    std::copy (h->nnls_r.begin (), h->nnls_r.end (), r);

  return NCKDE_OK;
}

int
nckde_get_Ki (nckde *h, size_t i, double *y_i, double *cov_i, double *n_i, double *w_i)
{
  TRY (check_prepared (h));

  if (y_i == nullptr || cov_i == nullptr || n_i == nullptr || w_i == nullptr)
  {
    set_error ("nckde_get_Ki: NULL output");
    return NCKDE_EINVAL;
  }

  const size_t d = (size_t) h->d;
  std::vector<double> U (d * d);
  double lnnorm = 0.0, href = 0.0;

  TRY (nckde_get_cov_decomp (h, i, U.data ()));
  TRY (nckde_get_lnnorm (h, i, &lnnorm));
  TRY (nckde_get_href (h, &href));

  {
    std::lock_guard<std::mutex> lk (h->mtx);

    std::copy (h->X.begin () + i * d, h->X.begin () + (i + 1) * d, y_i);
    HIPCHK (hipMemcpy (w_i, h->dw.as<double> () + i, sizeof (double), hipMemcpyDeviceToHost));
  }

  // cov_i = href^2 U^T U over the upper triangle of the factor (the
  // reference's triang_to_sym also zeroes the stored factor's lower triangle;
  // the stored factor is left as it is here)
  for (size_t a = 0; a < d; ++a)
    for (size_t b = 0; b < d; ++b)
    {
      double acc = 0.0;

      for (size_t k = 0; k <= std::min (a, b); ++k)
        acc += U[k * d + a] * U[k * d + b];

      cov_i[a * d + b] = acc * href * href;
    }

  *n_i = std::exp (lnnorm);

  return NCKDE_OK;
}

int
nckde_get_full_cov (nckde *h, double *cov)
{
  TRY (check_prepared (h));
  std::lock_guard<std::mutex> lk (h->mtx);
  HIPCHK (hipMemcpy (cov, h->dcovg.p, sizeof (double) * h->d * h->d, hipMemcpyDeviceToHost));

  return NCKDE_OK;
}

int
nckde_get_full_cov_decomp (nckde *h, double *U)
{
  TRY (check_prepared (h));
  std::lock_guard<std::mutex> lk (h->mtx);
  HIPCHK (hipMemcpy (U, h->dUg.p, sizeof (double) * h->d * h->d, hipMemcpyDeviceToHost));

  return NCKDE_OK;
}

int
nckde_get_sample (nckde *h, double *X)
{
  TRY (check_handle (h));
  std::lock_guard<std::mutex> lk (h->mtx);
  std::memcpy (X, h->X.data (), sizeof (double) * h->n_samples * h->d);

  return NCKDE_OK;
}

int
nckde_get_invUsample (nckde *h, double *Y)
{
  TRY (check_prepared (h));
  std::lock_guard<std::mutex> lk (h->mtx);
  HIPCHK (hipMemcpy (Y, h->dY.p, sizeof (double) * h->n_obs * h->d, hipMemcpyDeviceToHost));

  return NCKDE_OK;
}

int
nckde_get_knn (nckde *h, size_t *k, int32_t *idx)
{
  TRY (check_prepared (h));
  std::lock_guard<std::mutex> lk (h->mtx);

  if (h->kind != NCKDE_VKDE)
  {
    *k = 0;
    return NCKDE_OK;
  }

  *k = h->k;

  if (idx)
    HIPCHK (hipMemcpy (idx, h->dknn.p, sizeof (int32_t) * h->n_k * h->k, hipMemcpyDeviceToHost));

  return NCKDE_OK;
}

} // extern "C"

namespace nck
{

// Mixture draws whose variates come from the counter-based device generator
// instead of the caller (used by the APES walker; the transform is the one
// of nckde_sample).
int
sample_counter_rng (nckde *h, size_t m, uint64_t key, uint64_t ctr, double *out)
{
  TRY (check_prepared (h));
  std::lock_guard<std::mutex> lk (h->mtx);
  const int d = h->d;

  if (m == 0)
    return NCKDE_OK;

  const int nu_int = (h->kernel == NCKDE_KERNEL_ST) ? (int) h->nu : 0;

  if (h->kernel == NCKDE_KERNEL_ST && (double) nu_int != h->nu)
  {
    set_error ("nckde: device draws need an integer Student-t nu");
    return NCKDE_EUNSUPPORTED;
  }

  TRY (ensure_wcum (h));
  HIPCHK (h->du.ensure (sizeof (double) * m));
  HIPCHK (h->dz.ensure (sizeof (double) * m * d));
  HIPCHK (h->dchi2.ensure (sizeof (double) * m));
  HIPCHK (h->dout.ensure (sizeof (double) * m * d));
  HIPCHK (h->didx.ensure (sizeof (uint32_t) * m));
  HIPCHK (nck::launch_draws ((int) m, d, nu_int, key, ctr, h->du.as<double> (), h->dz.as<double> (),
                             nu_int > 0 ? h->dchi2.as<double> () : nullptr, h->stream));
  HIPCHK (nck::launch_sample (h->dwcum.as<double> (), (int) h->n_k, h->du.as<double> (), h->dz.as<double> (),
                              h->dchi2.as<double> (), (int) m, d, h->kind, h->kernel, h->nu, h->href,
                              (h->kind == NCKDE_VKDE) ? h->dU.as<double> () : h->dUg.as<double> (),
                              h->dX.as<double> (), h->dout.as<double> (), h->didx.as<uint32_t> (), h->stream));
  HIPCHK (hipMemcpyAsync (out, h->dout.p, sizeof (double) * m * d, hipMemcpyDeviceToHost, h->stream));
  HIPCHK (hipStreamSynchronize (h->stream));

  return NCKDE_OK;
}

} // namespace nck

extern "C" {

// ---- serialisation --------------------------------------------------------
//
// Layout (native byte order): "NCKDE02" + NUL; int32 kind, kernel, d, cv,
// use_rot_href, cov_type, nearPD_maxiter, has_cov_fixed, prepared,
// print_fit; double nu, over_smooth, shrink, local_frac, split_frac; uint64
// cv_draws; [d*d cov_fixed]; uint64 n_samples and
// the sample rows; if prepared: uint64 n_obs, n_k, k; double href,
// kernel_lnnorm, rnorm; the whitened sample (n_obs x d), the global
// covariance and its factor (d x d each), the weights (n_k); VKDE only: the
// local covariances and factors (n_k x d x d each), their lnnorms (n_k) and
// the neighbour lists (n_k x k int32).  Evaluation factors are rebuilt on
// load.


int
nckde_serialize (nckde *h, int with_state, void *buf, size_t cap, size_t *len)
{
  TRY (check_handle (h));

  if (len == nullptr)
  {
    set_error ("nckde_serialize: NULL length pointer");
    return NCKDE_EINVAL;
  }

  std::lock_guard<std::mutex> lk (h->mtx);
  const int d          = h->d;
  const bool state     = with_state && h->prepared;
  const size_t dd      = (size_t) d * d;
  Writer w {(unsigned char *) buf, cap};

  if (state)
    HIPCHK (hipStreamSynchronize (h->stream));

  w.put (SER_MAGIC, sizeof (SER_MAGIC));

  for (int v : {h->kind, h->kernel, d, h->cv, h->use_rot_href, h->cov_type, (int) h->nearPD_maxiter,
                (int) h->has_cov_fixed, (int) state, (int) h->print_fit})
    w.val<int32_t> (v);

  for (double v : {h->nu, h->over_smooth, h->shrink, h->local_frac, h->split_frac})
    w.val<double> (v);

  w.val<uint64_t> (h->cv_draws);

  if (h->has_cov_fixed)
    w.put (h->cov_fixed.data (), sizeof (double) * dd);

  w.val<uint64_t> (h->n_samples);
  w.put (h->X.data (), sizeof (double) * h->n_samples * d);

  if (state)
  {
    w.val<uint64_t> (h->n_obs);
    w.val<uint64_t> (h->n_k);
    w.val<uint64_t> (h->k);
    w.val<double> (h->href);
    w.val<double> (h->kernel_lnnorm);
    w.val<double> (h->rnorm);

    auto dev = [&] (const DBuf &b, size_t bytes) -> int {
      std::vector<unsigned char> tmp (bytes);

      if (bytes > 0)
        HIPCHK (hipMemcpy (tmp.data (), b.p, bytes, hipMemcpyDeviceToHost));

      w.put (tmp.data (), bytes);

      return NCKDE_OK;
    };

    TRY (dev (h->dY, sizeof (double) * h->n_obs * d));
    TRY (dev (h->dcovg, sizeof (double) * dd));
    TRY (dev (h->dUg, sizeof (double) * dd));
    TRY (dev (h->dw, sizeof (double) * h->n_k));

    if (h->kind == NCKDE_VKDE)
    {
      TRY (dev (h->dcov, sizeof (double) * h->n_k * dd));
      TRY (dev (h->dU, sizeof (double) * h->n_k * dd));
      TRY (dev (h->dlnnorm, sizeof (double) * h->n_k));
      TRY (dev (h->dknn, sizeof (int32_t) * h->n_k * h->k));
    }
  }

  *len = w.len;

  if (buf != nullptr && w.len > cap)
  {
    set_error ("nckde_serialize: buffer of %zu bytes, %zu needed", cap, w.len);
    return NCKDE_EINVAL;
  }

  return NCKDE_OK;
}

} // extern "C"

// bytes left to read must cover every array before it is sized from the
// header's counts (a corrupt or truncated buffer then fails cleanly instead
// of asking for an allocation its length cannot back)
static int
deserialize_impl (nckde **out, const void *buf, size_t len)
{
  Reader r {(const unsigned char *) buf, len};
  char magic[8];

  r.get (magic, sizeof (magic));

  if (!r.ok || std::memcmp (magic, SER_MAGIC, sizeof (magic)) != 0)
  {
    set_error ("nckde_deserialize: not an nckde serialisation");
    return NCKDE_EINVAL;
  }

  const int kind = r.val<int32_t> (), kernel = r.val<int32_t> (), d = r.val<int32_t> (), cv = r.val<int32_t> ();
  const int rot = r.val<int32_t> (), cov_type = r.val<int32_t> (), maxit = r.val<int32_t> ();
  const int has_fixed = r.val<int32_t> (), state = r.val<int32_t> (), pfit = r.val<int32_t> ();
  const double nu = r.val<double> (), os = r.val<double> (), shrink = r.val<double> (), lf = r.val<double> ();
  const double sf = r.val<double> ();
  const uint64_t cv_draws = r.val<uint64_t> ();

  if (!r.ok || d <= 0 || d > 4096)
  {
    set_error ("nckde_deserialize: truncated or invalid header");
    return NCKDE_EINVAL;
  }

  nckde *h = nullptr;

  TRY (nckde_create (&h, kind, kernel, nu, (unsigned) d, cv));

  std::unique_ptr<nckde, int (*) (nckde *)> guard (h, nckde_destroy);
  const size_t dd = (size_t) d * d;

  h->use_rot_href   = rot;
  h->over_smooth    = os;
  h->shrink         = shrink;
  h->local_frac     = lf;
  h->split_frac     = sf;
  h->print_fit      = pfit != 0;
  h->cv_draws       = cv_draws;
  h->nearPD_maxiter = (unsigned) maxit;

  if (has_fixed)
  {
    if (sizeof (double) * dd > r.left ())
    {
      set_error ("nckde_deserialize: truncated fixed covariance");
      return NCKDE_EINVAL;
    }

    std::vector<double> cf (dd);

    r.get (cf.data (), sizeof (double) * dd);

    if (!r.ok)
    {
      set_error ("nckde_deserialize: truncated fixed covariance");
      return NCKDE_EINVAL;
    }

    TRY (nckde_set_cov_fixed (h, cf.data ()));
  }

  TRY (nckde_set_cov_type (h, cov_type));

  const uint64_t ns = r.val<uint64_t> ();

  if (!r.ok || ns > (r.left () / sizeof (double)) / (size_t) d)
  {
    set_error ("nckde_deserialize: truncated sample");
    return NCKDE_EINVAL;
  }

  h->X.resize (ns * d);
  r.get (h->X.data (), sizeof (double) * ns * d);
  h->n_samples = ns;

  if (state)
  {
    const uint64_t n_obs = r.val<uint64_t> (), n_k = r.val<uint64_t> (), k = r.val<uint64_t> ();

    h->href          = r.val<double> ();
    h->kernel_lnnorm = r.val<double> ();
    h->rnorm         = r.val<double> ();

    if (!r.ok || n_obs > ns || n_k > n_obs || (kind == NCKDE_VKDE && (k == 0 || k > n_obs)))
    {
      set_error ("nckde_deserialize: invalid prepared state");
      return NCKDE_EINVAL;
    }

    // every count is bounded by ns <= len / (8 d) here, so none of these
    // products overflows; together they must fit in what is left
    size_t need = sizeof (double) * (n_obs * d + 2 * dd + n_k);

    if (kind == NCKDE_VKDE)
    {
      if (k > r.left () / sizeof (int32_t) / std::max<uint64_t> (n_k, 1))
        need = SIZE_MAX;
      else
        need += sizeof (double) * (2 * n_k * dd + n_k) + sizeof (int32_t) * n_k * k;
    }

    if (need > r.left ())
    {
      set_error ("nckde_deserialize: truncated prepared state");
      return NCKDE_EINVAL;
    }

    h->n_obs = n_obs;
    h->n_k   = n_k;
    h->k     = k;
    h->D     = nck::padded_dim (d);

    auto dev = [&] (DBuf &b, size_t bytes) -> int {
      if (bytes > r.left ())
      {
        set_error ("nckde_deserialize: truncated prepared state");
        return NCKDE_EINVAL;
      }

      std::vector<unsigned char> tmp (bytes);

      r.get (tmp.data (), bytes);

      if (!r.ok)
      {
        set_error ("nckde_deserialize: truncated prepared state");
        return NCKDE_EINVAL;
      }

      HIPCHK (b.ensure (bytes));

      if (bytes > 0)
        HIPCHK (hipMemcpy (b.p, tmp.data (), bytes, hipMemcpyHostToDevice));

      return NCKDE_OK;
    };

    HIPCHK (h->dX.ensure (sizeof (double) * n_obs * d));
    HIPCHK (hipMemcpy (h->dX.p, h->X.data (), sizeof (double) * n_obs * d, hipMemcpyHostToDevice));
    TRY (dev (h->dY, sizeof (double) * n_obs * d));
    TRY (dev (h->dcovg, sizeof (double) * dd));
    TRY (dev (h->dUg, sizeof (double) * dd));
    TRY (dev (h->dw, sizeof (double) * n_k));

    if (kind == NCKDE_VKDE)
    {
      TRY (dev (h->dcov, sizeof (double) * n_k * dd));
      TRY (dev (h->dU, sizeof (double) * n_k * dd));
      TRY (dev (h->dlnnorm, sizeof (double) * n_k));
      TRY (dev (h->dknn, sizeof (int32_t) * n_k * k));
    }

    HIPCHK (h->dTh.ensure (sizeof (double) * n_k * h->D));

    if (kind == NCKDE_VKDE)
    {
      HIPCHK (h->dWp.ensure (sizeof (double) * n_k * (h->D * (h->D + 1) / 2)));
      HIPCHK (nck::launch_vkde_factors (h->dU.as<double> (), h->dX.as<double> (), (int) n_k, d, h->D,
                                        h->dWp.as<double> (), h->dTh.as<double> (), h->stream));
    }
    else
    {
      HIPCHK (nck::launch_kde_factors (h->dY.as<double> (), (int) n_k, d, h->D, h->dTh.as<double> (), h->stream));
    }

    TRY (refresh_bvec (h));
    HIPCHK (h->dwcum.ensure (sizeof (double) * (n_k + 1)));
    HIPCHK (hipStreamSynchronize (h->stream));
    h->prepared = true;
  }

  if (!r.ok)
  {
    set_error ("nckde_deserialize: truncated input");
    return NCKDE_EINVAL;
  }

  *out = guard.release ();

  return NCKDE_OK;
}

extern "C" {

int
nckde_deserialize (nckde **out, const void *buf, size_t len)
{
  if (out == nullptr || buf == nullptr)
  {
    set_error ("nckde_deserialize: NULL argument");
    return NCKDE_EINVAL;
  }

  *out = nullptr;

  try
  {
    return deserialize_impl (out, buf, len);
  }
  catch (const std::bad_alloc &)
  {
    set_error ("nckde_deserialize: host allocation failed");
    return NCKDE_ENOMEM;
  }
  catch (const std::exception &e)
  {
    set_error ("nckde_deserialize: %s", e.what ());
    return NCKDE_EINVAL;
  }
}

int
nckde_set_profiling (nckde *h, int on)
{
  TRY (check_handle (h));
  std::lock_guard<std::mutex> lk (h->mtx);

  if (on && h->pev[0] == nullptr)
  {
    HIPCHK (hipEventCreate (&h->pev[0]));
    HIPCHK (hipEventCreate (&h->pev[1]));
  }

  h->profiling    = (on != 0);
  h->pev_recorded = false;

  return NCKDE_OK;
}

int
nckde_get_kernel_ms (nckde *h, float *ms)
{
  TRY (check_handle (h));
  std::lock_guard<std::mutex> lk (h->mtx);

  if (ms == nullptr || !h->pev_recorded)
  {
    set_error ("nckde_get_kernel_ms: no profiled eval on this handle");
    return NCKDE_ESTATE;
  }

  HIPCHK (hipEventSynchronize (h->pev[1]));
  HIPCHK (hipEventElapsedTime (ms, h->pev[0], h->pev[1]));

  return NCKDE_OK;
}

int
nckde_request_bytes (const nckde_request *req, size_t *bytes)
{
  if (req == nullptr || bytes == nullptr)
  {
    set_error ("nckde_request_bytes: NULL argument");
    return NCKDE_EINVAL;
  }

  TRY (check_handle (req->h));

  const nckde *h  = req->h;
  const size_t d  = (size_t) h->d;
  const size_t m  = req->m;
  const size_t D  = (size_t) nck::padded_dim (h->d);
  const size_t DP = (size_t) nck::mfma_dim (h->d);
  const size_t nk = std::max<size_t> (h->n_k, 1);

  switch (req->op)
  {
    case NCKDE_REQ_EVAL_M2LNP:
    {
      // points, output and the split partial sums (at most 2048 blocks of
      // partials per point block, see eval_mixed / eval_device)
      const size_t parts = std::min<size_t> ((nk + 63) / 64, 64);
      size_t b           = m * d * 8 + m * 8 + m * parts * 16;

      if (req->precision == NCKDE_PREC_MIXED)
        b += m * (4 * DP + 16);

      *bytes = b;
      break;
    }
    case NCKDE_REQ_SAMPLE:
      *bytes = m * (8 + 8 * d + 8 + 8 * d + 4);
      break;
    case NCKDE_REQ_PREPARE_INTERP:
    {
      const size_t n  = m;
      const size_t kk = std::max<size_t> ((size_t) (h->local_frac * n), 2);
      const size_t nq = std::min<size_t> (n, std::max<size_t> (1, ((size_t) 512 << 20) / (8 * std::max<size_t> (n, 1))));
      size_t b        = n * d * 8 * 2 + n * d * d * 8 * 2 + n * (D * (D + 1) / 2 + D + 4) * 8;

      if (h->kind == NCKDE_VKDE)
        b += n * kk * 4 + n * nq * 8;

      b += 3 * n * n * 8 + 12 * n * 8;  // IM^T, M = A^T A, the active-set matrix, vectors
      *bytes = b;
      break;
    }
    default:
      set_error ("nckde_request_bytes: unknown op %d", req->op);
      return NCKDE_EINVAL;
  }

  return NCKDE_OK;
}

} // extern "C"

namespace nck
{

static std::array<DBuf *, N_WORK_BUFS>
work_bufs (nckde *h)
{
  return {&h->dkeys, &h->dAT, &h->dM, &h->dS, &h->dlu, &h->dscratch, &h->drw, &h->drobust, &h->dpd};
}

// Frees the transient buffers of prepare / prepare_interp (kNN keys, IM^T,
// normal matrices); the prepared factors and weights stay.
int
trim_work (nckde *h)
{
  TRY (check_handle (h));
  std::lock_guard<std::mutex> lk (h->mtx);

  HIPCHK (hipStreamSynchronize (h->stream));

  for (DBuf *b : work_bufs (h))
    b->release ();

  return NCKDE_OK;
}

int
mixed_fuse_key (nckde *h)
{
  if (check_handle (h) != NCKDE_OK)
    return -1;

  std::lock_guard<std::mutex> lk (h->mtx);

  if (!h->prepared || !h->mixed_ready || !h->mixed_b_ready || h->mixed_fallback || h->pilot_n > 0)
    return -1;

  return nck::mfma_dim (h->d) * 2 + h->kernel;
}

int
swap_work (nckde *h, WorkSet &w)
{
  TRY (check_handle (h));
  std::lock_guard<std::mutex> lk (h->mtx);

  HIPCHK (hipStreamSynchronize (h->stream));

  const auto bufs = work_bufs (h);

  for (int i = 0; i < N_WORK_BUFS; ++i)
  {
    std::swap (bufs[i]->p, w.b[i].p);
    std::swap (bufs[i]->bytes, w.b[i].bytes);
  }

  return NCKDE_OK;
}

} // namespace nck
