// nckde_internal.h -- shared declarations of the gfx950 kernel-density engine.
//
// Device layout of one prepared handle (all row-major):
//   X      n_obs x d        FP64  sample theta_j
//   Y      n_obs x d        FP64  whitened sample y_j (U_g^T y_j = theta_j)
//   U      n_k x d x d      FP64  per-kernel Cholesky factors (cov_i = U_i^T U_i)
//   Wp     n_k x P(D)       FP64  packed lower W_i = U_i^{-T}, D = padded dim,
//                                 P(D) = D(D+1)/2 (VKDE) / none (KDE)
//   Th     n_k x D          FP64  kernel centres padded to D
//   lnnorm n_k              FP64  kernel log-normalisation (without d ln h)
//   w      n_k              FP64  weights;  b_i = ln w_i - lnnorm_i
//   Mixed-precision factors (VKDE): A_i = s_i W_i U_g^T / h split in FP16
//   hi/lo (n_k x 2 x DP x DP), c_i = A_i yc_i in FP32 (n_k x DP) and per-kernel
//   FP32 coefficients; yc = centred whitened coordinates.
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <vector>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/nckde.h"

namespace nck
{

constexpr double LN2PI = 1.8378770664093454835606594728112353;
constexpr double LNPI  = 1.1447298858494001741434273513530587;
constexpr double LOG_DBL_EPSILON = -3.6043653389117154e+01;

void set_error (const char *fmt, ...);

// Device buffer with grow-only capacity.
struct DBuf
{
  void *p      = nullptr;
  size_t bytes = 0;

  DBuf () = default;
  DBuf (const DBuf &) = delete;
  DBuf &operator= (const DBuf &) = delete;
  ~DBuf () { release (); }

  hipError_t ensure (size_t nbytes);
  void release ();

  template <class T> T *as () const { return static_cast<T *> (p); }
};

// Pinned host buffer (grows, never shrinks) for the read-backs and uploads
// of host-driven loops: copies from and to it are DMA transfers, without the
// staging copy of pageable memory.
struct HBuf
{
  void *p      = nullptr;
  size_t bytes = 0;

  HBuf () = default;
  HBuf (const HBuf &) = delete;
  HBuf &operator= (const HBuf &) = delete;
  ~HBuf () { release (); }

  hipError_t
  ensure (size_t nbytes)
  {
    if (nbytes <= bytes && p != nullptr)
      return hipSuccess;

    release ();

    const hipError_t e = hipHostMalloc (&p, nbytes ? nbytes : 8, hipHostMallocDefault);

    if (e != hipSuccess)
    {
      p = nullptr;
      return e;
    }

    bytes = nbytes;

    return hipSuccess;
  }

  void
  release ()
  {
    if (p)
      (void) hipHostFree (p);

    p     = nullptr;
    bytes = 0;
  }

  template <class T> T *as () const { return static_cast<T *> (p); }
};

// Frees a handle's transient prepare buffers (defined in api.cpp).
int trim_work (nckde *h);

// The transient buffers of prepare / prepare_interp (kNN keys, IM^T, normal
// matrices, factor and solve scratch).  A request server keeps one set per
// worker and lends it to the session it refits, so that refits reuse device
// memory instead of allocating and freeing it per request.
constexpr int N_WORK_BUFS = 9;
struct WorkSet
{
  DBuf b[N_WORK_BUFS];

  size_t
  bytes () const
  {
    size_t t = 0;

    for (const DBuf &x : b)
      t += x.bytes;

    return t;
  }

  void
  release ()
  {
    for (DBuf &x : b)
      x.release ();
  }
};

// Exchanges the handle's transient buffers with w (waits for the handle's
// stream first; defined in api.cpp).
int swap_work (nckde *h, WorkSet &w);

// Key under which the handle's mixed evals can share a fused launch (MFMA
// dimension and kernel type), or -1 (factors not built yet, FP64 fallback,
// or a pilot pass); defined in api.cpp.
int mixed_fuse_key (nckde *h);

// Padded dimension used by the templated per-pair kernels.
int padded_dim (int d);
// Padded dimension of the MFMA tile (multiples of 4 up to 64, see k_eval_mixed.hip).
int mfma_dim (int d);

// ---- prepare kernels (k_prepare.hip) ------------------------------------
hipError_t launch_welford_cov (const double *X, int d, int nrows, const int *rows, int rows_stride,
                               int nmats, double *cov, hipStream_t s);
hipError_t launch_chol_batched (const double *cov, double *U, int *info, double *lnnorm, int d,
                                int nmats, int kernel, double lnnorm_const, hipStream_t s);
// nearPD / diagonal fallback of the nfix matrices listed in which (d <= 64),
// one wave each; scratch holds 3 d^2 doubles per matrix
hipError_t launch_nearpd_fix (const int *which, int nfix, const double *cov, double *U, double *lnnorm,
                              double *scratch, int d, int maxiter, int kernel, double lnnorm_const,
                              int *status, hipStream_t s);
hipError_t launch_whiten (const double *X, const double *Ug, double *Y, int n, int d, hipStream_t s);
// true when the kNN selection keeps a query's keys in LDS (no key buffer)
bool knn_keys_in_lds (int n, int k);
// Yt: the whitened sample transposed (d x n, coordinate l of point j at l n + j)
hipError_t launch_knn (const double *Yt, int n, int d, int q0, int nq, int k, uint64_t *keys,
                       int *out, int *status, hipStream_t s);
size_t knn_max_k ();
hipError_t launch_vkde_factors (const double *U, const double *X, int nk, int d, int D, double *Wp,
                                double *Th, hipStream_t s);
hipError_t launch_kde_factors (const double *Y, int nk, int d, int D, double *Th, hipStream_t s);
hipError_t launch_bvec (const double *w, const double *lnnorm, int nk, double *b, hipStream_t s);

// ---- evaluation (k_eval.hip) --------------------------------------------
struct EvalArgs
{
  int kind, kernel, d, D, n, m;
  double nu, kappa, inv_h2, lnh_d, lnnorm_shared;
  const double *Th, *Wp, *b;       // kernel factors
  const double *Ug;                // KDE: global factor (d x d)
  const double *X;                 // points, m x d
  double *partM, *partS;           // nsplit x m
  int nsplit;
  double *out;                     // m
  int linear;                      // 0: m2lnp, 1: density
  // FP32 rung (b2f != NULL): FP32 factors (VKDE), b_i - bref in FP32
  const float *Wf = nullptr, *b2f = nullptr;
  double bref     = 0.0;
};
// FP64 eval, or the FP32 rung when a.b2f is set
hipError_t launch_eval_f64 (const EvalArgs &a, hipStream_t s);
hipError_t launch_f32_factors (const double *Wp, size_t count, float *Wf, hipStream_t s);
hipError_t launch_f32_b2 (const double *b, double bref, int n, float *b2f, hipStream_t s);
hipError_t launch_eval_merge (const EvalArgs &a, hipStream_t s);

// ---- mixed-precision evaluation (k_eval_mixed.hip) ------------------------
struct MixedFactors
{
  int DP;                 // padded MFMA dimension (4, 8, 16, 32, 64)
  const void *A;          // n_pad x 2 x DP x DP fp16 (hi, lo), row-major per kernel
  const float *c;         // n_pad x DP
  const float *coef;      // n_pad x 4: (q_i or r_i, b2_i, 1/s_i^2, 0)
};
struct MixedArgs
{
  int kernel, d, n, m, nsplit;
  double nu, kappa, lnh_d;
  double lnnorm_shared = 0.0;  // KDE: the kernel's shared lnnorm (VKDE: 0, lnnorm_i sits in b_i)
  MixedFactors f;
  const double *X;        // points m x d (FP64, device)
  const double *mu;       // centring (d)
  const double *G;        // U_g (d x d, upper): y' = U_g^{-T} (x - mu)
  double ysc;             // power-of-two scale of the whitened coordinates
  void *Yh;               // workspace: Yh, Yl (m x DP fp16 each), then m float2
  float *partM;           // nsplit x m: largest base-2 term
  double *partS;          // nsplit x m: sum of 2^(lnt2 - M0)
  int *flag;              // m: 1 where the terms underflowed (second pass)
  float *M0;              // m: reference of the second pass
  double bref;            // max_i b_i
  double *out;
  int linear;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;  // when set: recorded around the main evaluation kernel
  // pilot (pn > 0): the factors of pn kernels spread over the sample, whose
  // largest term per point is the main pass's per-point reference
  MixedFactors pf;
  int pn = 0, pnsplit = 0;
};
// what a merge of the split partials does
enum MixedMerge
{
  MERGE_PLAIN   = 0,  // terms against b_ref: results, flags of the underflowed points, their maxima
  MERGE_FLAGGED = 1,  // flagged points, terms against their maxima: results
  MERGE_PILOT   = 2,  // pilot pass: every point's largest term becomes its reference, every point flagged
  MERGE_CHECK   = 3,  // terms against the pilot reference: results; points whose largest term is too far
                      // above it (or whose sum is not finite) flagged, with their maxima
};
int mixed_points_per_block (int DP);
hipError_t launch_mixed_points (const MixedArgs &a, hipStream_t s);
hipError_t launch_eval_mixed (const MixedArgs &a, hipStream_t s);
hipError_t launch_mixed_merge (const MixedArgs &a, double bref, int mode, hipStream_t s);
// several mixed evals (one DP and kernel type) in one launch per pass: dsegs
// holds their arguments on the device, dblock0 the first eval block of each
// (nseg + 1 entries) then the first merge block of each (nseg + 1 entries);
// the points kernels run before, one per segment
int mixed_eval_blocks (const MixedArgs &a);
int mixed_merge_blocks (const MixedArgs &a);
hipError_t launch_eval_mixed_multi (int DP, int kernel, const MixedArgs *dsegs, const int *dblock0, int nseg,
                                    int eval_blocks, int merge_blocks, hipStream_t s);
hipError_t mixed_build (const double *X, int n, int d, int D, int DP, const double *Wp, const double *Ug,
                        const double *b, double bref, double inv_h, int kernel, double nu, double *mu, double *Yc,
                        unsigned long long *ymax_bits, double *ysc_out, _Float16 *A, float *c, float *coef,
                        float *risk,
                        hipStream_t s);
hipError_t mixed_refresh_b (const double *b, double bref, int n, float *coef, hipStream_t s);

// ---- robust covariances (k_robust.hip) -------------------------------------
size_t robust_work_doubles (int type, int d, int nrows, int nmats);
hipError_t launch_robust_cov (int type, const double *X, int d, int nrows, const int *rows, int rows_stride,
                              int nmats, double *cov, double *work, hipStream_t s);

// ---- interpolation matrix & linear algebra (k_linalg.hip) ----------------
struct IMArgs
{
  int kind, kernel, d, D, n_obs, n_k;
  double nu, inv_h2, h2, lnnorm_href, kde_scale;
  const double *X, *Y, *Th, *Wp, *lnnorm;
  const double *m2lnp;     // NULL: no row scaling
  double min_m2lnp;
  double *AT;              // n_k x n_obs (column-major A)
};
hipError_t launch_build_IM (const IMArgs &a, hipStream_t s);
hipError_t launch_transpose (const double *in, int rows, int cols, double *out, hipStream_t s);
hipError_t launch_gram_nt (const double *AT, int n, int K, double *M, hipStream_t s);
hipError_t launch_row_dot (const double *AT, int n, int K, const double *v, double *out, hipStream_t s);
size_t residual_work_doubles (int nk, int nobs);
hipError_t launch_residual (const double *AT, int nk, int nobs, const double *x, const double *f,
                            double *r, double *partial, double *rnorm2, double *work, hipStream_t s);
hipError_t launch_gather_sym (const double *M, int n, const int *P, int np, double *S,
                              const double *b, double *bs, hipStream_t s);
size_t potrf_diag_doubles (int np);
hipError_t launch_potrf (double *S, int np, int *info, double *diagbuf, hipStream_t s);
hipError_t launch_trsv_pair (const double *S, int np, double *b, double *tmp, hipStream_t s);
// the same solve as two flag-chained launches (np / 64 panels at most
// trsv_chain_max_panels ()); flags holds 1 + 2 ceil(np / 64) words, zero or
// from earlier epochs; flags[0] == epoch afterwards means the chain gave up
// and b is not the solution
int trsv_chain_max_panels ();
hipError_t launch_trsv_chain (const double *S, int np, double *b, double *tmp, unsigned *flags, unsigned epoch,
                              hipStream_t s);
hipError_t launch_lu_solve (double *S, int np, double *b, int *info, hipStream_t s);
hipError_t launch_trsv_backward (const double *S, int np, double *y, double *x, hipStream_t s);
hipError_t launch_mirror_lower (double *S, int np, hipStream_t s);
// k_nnls.hip: the symmetric-solve and QR rungs, refinement helpers
hipError_t launch_lu_factor (double *S, int np, double *b, double *colv, int *info, hipStream_t s);
size_t qr_work_doubles (int nobs, int np);
hipError_t launch_qr_lstsq (const double *AT, int nobs, const int *P, int np, const double *f, double *work,
                            int *iwork, double *x, hipStream_t s);
hipError_t launch_row_dot_idx (const double *AT, const int *P, int np, int K, const double *v, double *out,
                               hipStream_t s);
hipError_t launch_add_update (double *x, const double *dx, int np, double *stat, hipStream_t s);
hipError_t launch_diag_range (const double *S, int np, double *out, hipStream_t s);
// k_probe.hip: the np highest-risk kernels (sel) and two probe variates each
// (kernel idx, z, chi2); the largest relative difference of two results
hipError_t launch_probe_points (const float *risk, int n, int np, int d, double c, double chi2v, uint32_t *sel,
                                uint32_t *idx, double *z, double *chi2, hipStream_t s);
hipError_t launch_probe_err (const double *a, const double *b, int m, double *out, hipStream_t s);
hipError_t launch_scatter (const double *sol, const int *P, int np, int n, double *x, hipStream_t s);
hipError_t launch_shrink (double *w, int n, double shrink, int *status, hipStream_t s);

// ---- sampling (k_sample.hip) ----------------------------------------------
hipError_t launch_wcum (const double *w, int n, double *wcum, hipStream_t s);
hipError_t launch_choose (const double *wcum, int n, const double *u, int m, uint32_t *idx, hipStream_t s);
hipError_t launch_draws (int m, int d, int nu_int, uint64_t key, uint64_t ctr, double *u, double *z, double *chi2,
                        hipStream_t s);
// m mixture draws with device-generated variates (Philox stream (key, ctr));
// defined in api.cpp; nu must be an integer for the Student-t kernel
int sample_counter_rng (nckde *h, size_t m, uint64_t key, uint64_t ctr, double *out);
hipError_t launch_sample (const double *wcum, int n, const double *u, const double *z, const double *chi2,
                          int m, int d, int kind, int kernel, double nu, double href, const double *U,
                          const double *X, double *out, uint32_t *idx, hipStream_t s);
// as launch_sample, with the kernel of draw p given by idx_in[p]
hipError_t launch_sample_at (const uint32_t *idx_in, const double *z, const double *chi2, int m, int d, int kind,
                             int kernel, double nu, double href, const double *U, const double *X, double *out,
                             hipStream_t s);
// antithetic pairs: draw 2t picks i = choose(u[2t]); idx[2t] = order[i],
// idx[2t+1] = order[n-1-i]
hipError_t launch_pair_index (const double *wcum, int n, const double *u, int npairs, const uint32_t *order,
                              uint32_t *idx, hipStream_t s);
// per observation j of IM^T (n_k x n_obs): off[j] = sum over kernels i != j
// in ascending i, diag[j] = IM^T[j][j] (0 for j >= n_k)
hipError_t launch_im_obs_sums (const double *AT, int nk, int nobs, double *off, double *diag, hipStream_t s);

// ---- cross-validation minimisers (cv_min.cpp) -----------------------------
#define TRY_OK(call)                                                                 \
  do {                                                                               \
    int r_ok_ = (call);                                                              \
    if (r_ok_ != NCKDE_OK)                                                           \
      return r_ok_;                                                                  \
  } while (0)

using ObjFn = std::function<int (const double *x, double &val)>;
using ResFn = std::function<int (double p, double *hx)>;

struct NMInfo
{
  int iter, status;    // status: 0 size below tolerance, 1 iteration cap, -1 failed contraction
  double fval, size;
};
// Nelder-Mead (GSL nmsimplex2) from x with initial steps step, until the
// simplex size falls below epsabs or maxiter iterations
int nm_minimize (const ObjFn &f, std::vector<double> &x, const std::vector<double> &step, double epsabs,
                 int maxiter, NMInfo *info);

struct LMInfo
{
  double init_eL2, eL2;
  int iter, stop, nfev, njac;
};
// levmar's forward-difference Levenberg-Marquardt for one parameter p and
// n residuals hx (zero measurement vector); opts = {mu, eps1, eps2, eps3, delta}
int lm_dif1 (const ResFn &func, double &p, int n, int itmax, const double opts[5], LMInfo *info);

struct PairStats
{
  double W = 0.0, W2 = 0.0, bias = 0.0;
  double mean[2] = {0.0, 0.0}, var[2] = {0.0, 0.0}, cov01 = 0.0;
};
void pair_stats_update (PairStats &st, double x0, double x1);

// third counter word of the host normal stream (separates it from the
// sampling streams that share a key)
constexpr uint32_t HOST_NORMAL_TAG = 0x43565350u;
// standard normal number ctr of the stream with this key (Box-Muller, cosine branch)
double host_normal (uint64_t key, uint64_t ctr);

} // namespace nck
