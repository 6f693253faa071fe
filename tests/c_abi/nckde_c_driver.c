/*
 * nckde_c_driver.c -- a plain C99 caller of the nckde C-ABI (include/nckde.h),
 * with no GLib and no Python: the path a NumCosmo build would take through
 * the boundary.  It drives create -> add_obs -> prepare_interp -> eval_m2lnp
 * (FP64 and mixed) -> kernel_choose -> sample -> getters -> serialise, and
 * checks every result against the FP64 CPU oracle (oracle/nckde_oracle.h),
 * the test checker.
 *
 * Exit status 0 and a last line "nckde C driver: OK" on success; 1 with the
 * failing check on stderr otherwise; 2 when no device is available.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "nckde.h"
#include "nckde_oracle.h"

#define D 4
#define N 600
#define M 3000
#define NU 3.0

static int failures = 0;

#define CHECK(cond, ...)                                          \
  do {                                                            \
    if (!(cond))                                                  \
    {                                                             \
      fprintf (stderr, "FAIL %s:%d: ", __FILE__, __LINE__);       \
      fprintf (stderr, __VA_ARGS__);                              \
      fprintf (stderr, "\n");                                     \
      failures++;                                                 \
    }                                                             \
  } while (0)

#define CALL(expr)                                                                     \
  do {                                                                                 \
    int st_ = (expr);                                                                  \
    if (st_ != NCKDE_OK)                                                               \
    {                                                                                  \
      fprintf (stderr, "FAIL %s:%d: %s -> %d (%s)\n", __FILE__, __LINE__, #expr, st_, \
               nckde_last_error ());                                                   \
      return 1;                                                                        \
    }                                                                                  \
  } while (0)

/* xorshift64* and Box-Muller: a self-contained seeded stream */
static uint64_t rng_state = 0x9E3779B97F4A7C15ull;

static double
uniform01 (void)
{
  rng_state ^= rng_state >> 12;
  rng_state ^= rng_state << 25;
  rng_state ^= rng_state >> 27;
  return (double) ((rng_state * 2685821657736338717ull) >> 11) * (1.0 / 9007199254740992.0);
}

static double
normal (void)
{
  double u1 = uniform01 (), u2 = uniform01 ();

  if (u1 < 1e-300)
    u1 = 1e-300;

  return sqrt (-2.0 * log (u1)) * cos (6.283185307179586 * u2);
}

/* banana target: x2 += 0.5 (x1^2 - 1); m2lnp exact (Jacobian 1) */
static void
banana (double *x, double *m2lnp)
{
  double z[D], s = 0.0;
  int a;

  for (a = 0; a < D; ++a)
    z[a] = normal ();

  for (a = 0; a < D; ++a)
  {
    x[a] = z[a];
    s += z[a] * z[a];
  }

  x[1] += 0.5 * (x[0] * x[0] - 1.0);

  if (m2lnp)
    *m2lnp = s + D * 1.8378770664093453;
}

static double
rel (double a, double b)
{
  return fabs (a - b) / fmax (fabs (b), 1.0);
}

int
main (void)
{
  static double X[N * D], m2[N], Xe[M * D], got[M], ref[M], w[N], wo[N], U[D * D];
  static double u[64], z[64 * D], chi2[64], xs[64 * D], xo[D];
  static uint32_t idx[64];
  nckde *h = NULL, *h2 = NULL;
  ok_sd *o;
  int i, a;
  double e, l1 = 0.0, wsum = 0.0;

  for (i = 0; i < N; ++i)
    banana (X + i * D, m2 + i);

  for (i = 0; i < M; ++i)
    banana (Xe + i * D, NULL);

  if (nckde_create (&h, NCKDE_VKDE, NCKDE_KERNEL_ST, NU, D, NCKDE_CV_NONE) != NCKDE_OK)
  {
    fprintf (stderr, "no device: %s\n", nckde_last_error ());
    return 2;
  }

  /* eval before prepare: the reference's state error */
  CHECK (nckde_eval_m2lnp (h, Xe, 1, D, got, NCKDE_PREC_FP64) == NCKDE_ESTATE, "eval before prepare");

  for (i = 0; i < N; ++i)
    CALL (nckde_add_obs (h, X + i * D));

  CALL (nckde_prepare_interp (h, m2, N));

  o = ok_new (OK_VKDE, OK_ST, NU, D);

  for (i = 0; i < N; ++i)
    ok_add_obs (o, X + i * D);

  CHECK (ok_prepare_interp (o, m2, N) == 0, "oracle prepare_interp: %s", ok_last_error ());

  /* NNLS weights: relative L1 distance */
  CALL (nckde_get_weights (h, w));
  ok_get_weights (o, wo);

  for (i = 0; i < N; ++i)
  {
    l1 += fabs (w[i] - wo[i]);
    wsum += fabs (wo[i]);
  }

  CHECK (l1 / wsum <= 1e-7, "weights rel L1 %.3e", l1 / wsum);

  /* local factors are bit-exact */
  {
    int bad = 0;

    for (i = 0; i < N; i += 37)
    {
      double Uo[D * D];

      CALL (nckde_get_cov_decomp (h, (size_t) i, U));
      ok_get_cov_decomp (o, i, Uo);

      for (a = 0; a < D * D; ++a)
        bad += (U[a] != Uo[a]);
    }

    CHECK (bad == 0, "%d cov_decomp entries differ", bad);
  }

  /* evaluation against the oracle carrying the device's weights */
  ok_set_weights (o, w);
  ok_eval_m2lnp_batch (o, Xe, M, ref);

  CALL (nckde_eval_m2lnp (h, Xe, M, D, got, NCKDE_PREC_FP64));
  e = 0.0;

  for (i = 0; i < M; ++i)
    e = fmax (e, rel (got[i], ref[i]));

  CHECK (e <= 1e-12, "FP64 eval_m2lnp rel err %.3e", e);
  printf ("fp64 eval_m2lnp: max rel err %.2e over %d points\n", e, M);

  CALL (nckde_eval_m2lnp (h, Xe, M, D, got, NCKDE_PREC_MIXED));
  e = 0.0;

  for (i = 0; i < M; ++i)
    e = fmax (e, rel (got[i], ref[i]));

  CHECK (e <= 1e-6, "mixed eval_m2lnp rel err %.3e", e);
  printf ("mixed eval_m2lnp: max rel err %.2e over %d points\n", e, M);

  /* linear eval */
  CALL (nckde_eval (h, Xe, 200, D, got, NCKDE_PREC_FP64));
  ok_eval_batch (o, Xe, 200, ref);
  e = 0.0;

  for (i = 0; i < 200; ++i)
    e = fmax (e, fabs (got[i] - ref[i]) / fmax (fabs (ref[i]), 1e-300));

  CHECK (e <= 1e-11, "eval rel err %.3e", e);

  /* kernel_choose and sample: exact given the uniforms and variates */
  for (i = 0; i < 64; ++i)
  {
    u[i]    = (i == 0) ? 0.0 : uniform01 ();
    chi2[i] = 0.0;

    for (a = 0; a < D; ++a)
      z[i * D + a] = normal ();

    for (a = 0; a < 3; ++a)
    {
      double g = normal ();
      chi2[i] += g * g;
    }
  }

  CALL (nckde_kernel_choose (h, u, 64, idx));

  for (i = 0; i < 64; ++i)
    CHECK ((int) idx[i] == ok_kernel_choose (o, u[i]), "kernel_choose u=%.17g: %u vs %d", u[i], idx[i],
           ok_kernel_choose (o, u[i]));

  CALL (nckde_sample (h, u, z, chi2, 64, xs, idx));
  e = 0.0;

  for (i = 0; i < 64; ++i)
  {
    ok_sample (o, u[i], z + i * D, chi2[i], xo);

    for (a = 0; a < D; ++a)
      e = fmax (e, rel (xs[i * D + a], xo[a]));
  }

  CHECK (e <= 1e-14, "sample rel err %.3e", e);

  /* getters */
  {
    double v, yi[D], ci[D * D], ni, wi, href, lnn;
    int iv;
    unsigned uv;

    CALL (nckde_get_shrink (h, &v));
    CHECK (v == 0.01, "shrink %g", v);
    CALL (nckde_get_local_frac (h, &v));
    CHECK (v == 0.05, "local_frac %g", v);
    CALL (nckde_get_use_rot_href (h, &iv));
    CHECK (iv == 0, "use_rot_href %d", iv);
    CALL (nckde_get_nearPD_maxiter (h, &uv));
    CHECK (uv == 200, "nearPD_maxiter %u", uv);
    CALL (nckde_get_cov_type (h, &iv));
    CHECK (iv == NCKDE_COV_SAMPLE, "cov_type %d", iv);
    CHECK (nckde_get_cov_fixed (h, ci) == NCKDE_ESTATE, "cov_fixed without one set");
    CALL (nckde_set_use_threads (h, 1));
    CALL (nckde_get_use_threads (h, &iv));
    CHECK (iv == 1, "use_threads %d", iv);

    CALL (nckde_get_Ki (h, 5, yi, ci, &ni, &wi));
    CALL (nckde_get_cov_decomp (h, 5, U));
    CALL (nckde_get_href (h, &href));
    CALL (nckde_get_lnnorm (h, 5, &lnn));

    for (a = 0; a < D; ++a)
      CHECK (yi[a] == X[5 * D + a], "Ki centre");

    for (a = 0; a < D; ++a)
    {
      int b;

      for (b = 0; b < D; ++b)
      {
        double acc = 0.0;
        int k;

        for (k = 0; k <= (a < b ? a : b); ++k)
          acc += U[k * D + a] * U[k * D + b];

        CHECK (rel (ci[a * D + b], acc * href * href) <= 1e-15, "Ki covariance (%d,%d)", a, b);
      }
    }

    CHECK (rel (ni, exp (lnn)) <= 1e-15 && wi == w[5], "Ki norm / weight");
  }

  /* serialise with the prepared state; the restored object evaluates the
   * same numbers */
  {
    size_t len = 0;
    void *buf;
    static double got2[M];

    CALL (nckde_serialize (h, 1, NULL, 0, &len));
    buf = malloc (len);
    CALL (nckde_serialize (h, 1, buf, len, &len));
    CALL (nckde_deserialize (&h2, buf, len));
    {
      nckde *h3 = NULL;

      CHECK (nckde_deserialize (&h3, buf, len / 2) == NCKDE_EINVAL && h3 == NULL,
             "truncated serialisation accepted");
    }
    CALL (nckde_eval_m2lnp (h2, Xe, M, D, got2, NCKDE_PREC_FP64));
    CALL (nckde_eval_m2lnp (h, Xe, M, D, got, NCKDE_PREC_FP64));
    CHECK (memcmp (got, got2, sizeof (double) * M) == 0, "restored object evaluates differently");
    free (buf);
    nckde_destroy (h2);
  }

  ok_free (o);
  nckde_destroy (h);

  if (failures)
  {
    fprintf (stderr, "nckde C driver: %d check(s) failed\n", failures);
    return 1;
  }

  printf ("nckde C driver: OK\n");

  return 0;
}
