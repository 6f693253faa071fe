"""Parity of the HIP engine (through the C-ABI) with the FP64 CPU oracle.

Bars:
* integer / index results (neighbour lists, chosen kernels) .... bit-exact
* prepare factors (covariances, Cholesky factors, whitening) ... bit-exact
  (same operation order, IEEE division / sqrt, no contraction), lnnorm to
  1e-14 relative (device log vs glibc log)
* FP64 eval_m2lnp ........................ |rel err| <= 1e-12
* prepare_interp weights ................. rel L1 <= 1e-7, KKT of the fit
"""
import sys

import numpy as np
import pytest

from numcosmo_amd import synth

pytestmark = pytest.mark.gpu

GAUSS, ST = 0, 1


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gpu():
    import numcosmo_amd as na
    return na


def _pair(gpu, oracle, kind, kernel, d, X, nu=3.0, **kw):
    sdk = gpu.StatsDistKernelGauss.new(d) if kernel == GAUSS else gpu.StatsDistKernelST.new(d, nu)
    sd = (gpu.StatsDistVKDE if kind == 1 else gpu.StatsDistKDE).new(sdk, gpu.StatsDistCV.NONE)
    od = oracle.StatsDist(kind, kernel, d, nu)
    for k, v in kw.items():
        getattr(sd, "set_" + k)(v)
        getattr(od, "set_" + k)(v)
    sd.add_obs_array(X)
    od.add_obs_array(X)
    return sd, od


CASES = [
    # kind, kernel, d, n, nu
    (1, GAUSS, 2, 300, 0.0),
    (1, ST, 3, 600, 3.0),
    (1, ST, 8, 1200, 3.0),
    (1, GAUSS, 16, 1000, 0.0),
    (0, GAUSS, 4, 500, 0.0),
    (0, ST, 5, 400, 4.0),
    # above d=16: padded dimensions 24, 32 and 64 of the FP64 eval and IM kernels
    (1, ST, 24, 1200, 3.0),
    (1, GAUSS, 32, 1500, 0.0),
    (1, ST, 64, 1800, 5.0),
    (0, GAUSS, 40, 900, 0.0),
]


@pytest.mark.parametrize("kind,kernel,d,n,nu", CASES)
def test_prepare_factors_bit_exact(gpu, oracle, kind, kernel, d, n, nu):
    _say(f"[prepare] kind={kind} kernel={kernel} d={d} n={n}")
    g = np.random.default_rng(1000 + d + n)
    X = synth.Banana(d).draw(g, n) if d >= 2 else g.normal(size=(n, d))
    sd, od = _pair(gpu, oracle, kind, kernel, d, X, nu)
    sd.prepare()
    od.prepare()
    assert sd.get_sample_size() == od.get_sample_size() == n
    assert sd.get_href() == od.get_href()
    np.testing.assert_array_equal(sd.peek_full_cov(), od.get_full_cov())
    np.testing.assert_array_equal(np.triu(sd.peek_full_cov_decomp()), np.triu(od.get_full_cov_decomp()))
    np.testing.assert_array_equal(sd.peek_invUsample(), od.get_invUsample())
    if kind == 1:
        np.testing.assert_array_equal(sd.peek_knn(), od.get_knn())
        for i in range(0, n, max(1, n // 50)):
            np.testing.assert_array_equal(sd.peek_cov_decomp(i), od.get_cov_decomp(i))
    for i in range(0, n, max(1, n // 20)):
        assert sd.get_lnnorm(i) == pytest.approx(od.get_lnnorm(i), rel=1e-14, abs=1e-14)
    np.testing.assert_array_equal(sd.peek_weights(), od.get_weights())


@pytest.mark.parametrize("kind,kernel,d,n,nu", CASES)
def test_eval_m2lnp_fp64(gpu, oracle, kind, kernel, d, n, nu):
    _say(f"[eval fp64] kind={kind} kernel={kernel} d={d} n={n}")
    g = np.random.default_rng(2000 + d + n)
    tgt = synth.Banana(d) if d >= 2 else None
    X = tgt.draw(g, n)
    sd, od = _pair(gpu, oracle, kind, kernel, d, X, nu, over_smooth=0.8)
    sd.prepare()
    od.prepare()
    Xe = tgt.draw(g, 700)
    Xe[:5] = X[:5]             # points on kernel centres
    Xe[5] = 40.0               # far outlier: every kernel deep in its tail
    ref = od.eval_m2lnp_batch(Xe)
    got = sd.eval_m2lnp_batch(Xe)
    err = np.abs(got - ref) / np.maximum(np.abs(ref), 1.0)
    _say(f"   max rel err {err.max():.3e}")
    assert err.max() <= 1e-12
    lin = sd.eval_batch(Xe[:50])
    lref = od.eval_batch(Xe[:50])
    assert np.allclose(lin, lref, rtol=1e-11, atol=0)
    # the single-point entry routes to the same engine
    assert sd.eval_m2lnp(Xe[7]) == got[7]


def _kkt_rel(A, x):
    g = A.T @ (1.0 - A @ x)
    scale = np.abs(A.T @ np.ones(A.shape[0])).max()
    act = x > 0
    return max(np.abs(g[act]).max(initial=0.0), g[~act].max(initial=0.0)) / scale


@pytest.mark.parametrize("kind,kernel,d,n,nu", [
    (1, ST, 2, 200, 3.0), (1, GAUSS, 3, 400, 0.0), (1, ST, 8, 800, 3.0), (0, GAUSS, 4, 300, 0.0),
    (1, ST, 24, 600, 3.0), (1, GAUSS, 32, 800, 0.0), (1, ST, 64, 1400, 5.0)])
def test_prepare_interp_weights(gpu, oracle, kind, kernel, d, n, nu):
    _say(f"[prepare_interp] kind={kind} kernel={kernel} d={d} n={n}")
    g = np.random.default_rng(3000 + d + n)
    tgt = synth.Gaussian(d, g)
    X = tgt.draw(g, n)
    m2lnp = tgt.m2lnp(X)
    sd, od = _pair(gpu, oracle, kind, kernel, d, X, nu)
    sd.prepare_interp(m2lnp)
    od.prepare_interp(m2lnp)
    w, wo = sd.peek_weights(), od.get_weights()
    rel_l1 = np.abs(w - wo).sum() / np.abs(wo).sum()
    _say(f"   weights rel L1 {rel_l1:.3e}, rnorm {sd.get_rnorm():.6e} vs {od.get_rnorm():.6e}")
    assert rel_l1 <= 1e-7
    assert sd.get_rnorm() == pytest.approx(od.get_rnorm(), rel=1e-7, abs=1e-14)
    # the device interpolation matrix equals the oracle's: entries K e^{-lnnorm}
    # to 1e-12 in their logarithm relative to its size (a kernel value deep
    # in the tail, K = e^{-chi2/2} with chi2 ~ 1e3, carries the rounding of
    # chi2 times chi2/2 in its relative error)
    IM, IMo = sd.compute_IM(), od.compute_IM()
    live = IMo > 1e-290
    assert np.array_equal(IM > 1e-290, live)
    lg, lgo = np.log(IM[live]), np.log(IMo[live])
    assert np.all(np.abs(lg - lgo) <= 1e-12 * np.maximum(1.0, np.abs(lgo)))
    big = IMo > 1e-3 * IMo.max()
    assert np.allclose(IM[big], IMo[big], rtol=1e-12, atol=0)
    # KKT of the device solution, held to the oracle's own (the active-set
    # loop stops when a single added index no longer lowers the residual by
    # a relative 2^-52, not at an exact KKT point)
    fvec = np.exp(-0.5 * (m2lnp - m2lnp.min()))
    A = IMo / fvec[:, None]

    def unshrunk(wv):
        v = wv - 0.01 / n
        Av = A @ v
        return v * (Av.sum() / (Av @ Av))

    kd, ko = _kkt_rel(A, unshrunk(w)), _kkt_rel(A, unshrunk(wo))
    _say(f"   KKT residual device {kd:.2e}, oracle {ko:.2e}")
    assert kd < max(1e-8, 10.0 * ko)
    # evaluation after the fit
    Xe = tgt.draw(g, 300)
    ref = od.eval_m2lnp_batch(Xe)
    got = sd.eval_m2lnp_batch(Xe)
    assert np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1.0)) < 1e-7


def test_prepare_interp_cut_and_fallback(gpu, oracle):
    _say("[prepare_interp cut/fallback]")
    g = np.random.default_rng(3100)
    d, n = 2, 120
    X = g.normal(size=(n, d))
    # fallback: more than half beyond the window
    m2 = np.zeros(n)
    m2[:70] = 400.0 + g.uniform(size=70)
    sd, od = _pair(gpu, oracle, 1, GAUSS, d, X, local_frac=0.1)
    sd.prepare_interp(m2)
    od.prepare_interp(m2)
    np.testing.assert_array_equal(sd.peek_weights(), od.get_weights())
    # cut: a tenth beyond the window, recursion on the rest
    m2 = np.sum(X * X, 1)
    m2[::10] += 1000.0
    sd, od = _pair(gpu, oracle, 1, GAUSS, d, X, local_frac=0.1)
    sd.prepare_interp(m2)
    od.prepare_interp(m2)
    assert sd.get_n_samples() == od.get_n_samples() == 108
    np.testing.assert_array_equal(sd.peek_sample_array(), od.get_sample())
    assert np.abs(sd.peek_weights() - od.get_weights()).sum() < 1e-7


@pytest.mark.parametrize("kind,kernel,nu", [(1, GAUSS, 0.0), (1, ST, 3.0), (0, ST, 5.0)])
def test_kernel_choose_and_sample_exact(gpu, oracle, kind, kernel, nu):
    _say(f"[sample] kind={kind} kernel={kernel}")
    g = np.random.default_rng(4000 + kernel)
    d, n, m = 3, 500, 20000
    X = synth.Banana(d).draw(g, n)
    sd, od = _pair(gpu, oracle, kind, kernel, d, X, nu)
    sd.prepare()
    od.prepare()
    u = g.uniform(size=m)
    u[:3] = [0.0, np.nextafter(1.0, 0.0), 0.5]
    idx = sd.kernel_choose_batch(u)
    ref = np.array([od.kernel_choose(p) for p in u])
    np.testing.assert_array_equal(idx, ref)
    z = g.normal(size=(200, d))
    c2 = g.chisquare(nu if kernel == ST else 1.0, size=200)
    out, ii = sd.sample_batch(u[:200], z, c2 if kernel == ST else None)
    for p in range(200):
        i, x = od.sample(u[p], z[p], c2[p])
        assert ii[p] == i
        np.testing.assert_allclose(out[p], x, rtol=1e-15, atol=1e-15)


def test_kernel_choose_after_fit(gpu, oracle):
    """Selection from the fitted weights is the reference's bisection on the
    serial prefix sum of those weights."""
    g = np.random.default_rng(4100)
    d, n = 2, 300
    tgt = synth.Gaussian(d, g)
    X = tgt.draw(g, n)
    sd, _ = _pair(gpu, oracle, 1, ST, d, X, 3.0)
    sd.prepare_interp(tgt.m2lnp(X))
    w = sd.peek_weights()
    wcum = np.concatenate([[0.0], np.cumsum(w)])
    wcum = wcum * (1.0 / wcum[-1])
    u = g.uniform(size=50000)
    ref = np.maximum(np.searchsorted(wcum[:n], u, side="right") - 1, 0)
    np.testing.assert_array_equal(sd.kernel_choose_batch(u), ref)


def test_errors_match_reference_text(gpu):
    sdk = gpu.StatsDistKernelGauss.new(3)
    sd = gpu.StatsDistKDE.new(sdk, gpu.StatsDistCV.NONE)
    sd.add_obs_array(np.eye(3))
    with pytest.raises(gpu.NckdeError, match="the sample is too small"):
        sd.prepare()
    sd = gpu.StatsDistVKDE.new(sdk, gpu.StatsDistCV.NONE)
    sd.add_obs_array(np.random.default_rng(0).normal(size=(20, 3)))
    with pytest.raises(gpu.NckdeError, match="Too few observations"):
        sd.prepare()
    with pytest.raises(gpu.NckdeError, match="not prepared"):
        gpu.StatsDistVKDE.new(sdk).eval_m2lnp(np.zeros(3))


def test_kde_fixed_covariance(gpu, oracle):
    g = np.random.default_rng(4200)
    d, n = 3, 200
    cov = synth.spd(g, d)
    X = g.normal(size=(n, d)) @ np.linalg.cholesky(cov).T
    sd, od = _pair(gpu, oracle, 0, GAUSS, d, X)
    for s in (sd, od):
        s.set_cov_fixed(cov)
        s.set_cov_type(1)
    sd.prepare()
    od.prepare()
    Xe = g.normal(size=(100, d))
    assert np.allclose(sd.eval_m2lnp_batch(Xe), od.eval_m2lnp_batch(Xe), rtol=1e-12)
    bad = -np.eye(d)
    sd2, _ = _pair(gpu, oracle, 0, GAUSS, d, X)
    sd2.set_cov_type(1)
    with pytest.raises(gpu.NckdeError, match="not positive definite"):
        sd2.set_cov_fixed(bad)


def test_nearpd_path_small_k(gpu, oracle):
    """k = 2 neighbours give rank-1 local covariances: the nearPD / diagonal
    fallback runs on the device for every kernel."""
    g = np.random.default_rng(4300)
    d, n = 3, 40
    X = g.normal(size=(n, d))
    sd, od = _pair(gpu, oracle, 1, GAUSS, d, X, local_frac=0.05)
    sd.prepare()
    od.prepare()
    for i in range(n):
        Ug, Uo = np.triu(sd.peek_cov_decomp(i)), np.triu(od.get_cov_decomp(i))
        assert np.allclose(Ug.T @ Ug, Uo.T @ Uo, rtol=1e-9, atol=1e-12)
    Xe = g.normal(size=(50, d))
    assert np.allclose(sd.eval_m2lnp_batch(Xe), od.eval_m2lnp_batch(Xe), rtol=1e-9)


@pytest.mark.parametrize("d,n,local_frac,maxiter", [(12, 200, 0.05, 200), (33, 80, 0.3, 10)])
def test_nearpd_path_rank_deficient(gpu, oracle, d, n, local_frac, maxiter):
    """Fewer neighbours than dimensions (k = 10 < 12, k = 24 < 33): every
    local covariance is singular and runs the nearPD iteration to its end
    (one wave per matrix, 16- and 64-wide LDS forms).  The factors equal the
    oracle's serial algorithm entry for entry."""
    _say(f"[nearPD] d={d} n={n} local_frac={local_frac} maxiter={maxiter}")
    g = np.random.default_rng(4400 + d)
    X = synth.Gaussian(d, g).draw(g, n)
    sd, od = _pair(gpu, oracle, 1, ST, d, X, 3.0, local_frac=local_frac, nearPD_maxiter=maxiter)
    sd.prepare()
    od.prepare()
    worst = 0.0
    for i in range(n):
        Ug, Uo = np.triu(sd.peek_cov_decomp(i)), np.triu(od.get_cov_decomp(i))
        worst = max(worst, float(np.max(np.abs(Ug - Uo) / np.maximum(np.abs(Uo), 1e-300))))
    _say(f"   factor max rel diff {worst:.3e}")
    assert worst <= 1e-12
    Xe = synth.Gaussian(d, g).draw(g, 50)
    assert np.allclose(sd.eval_m2lnp_batch(Xe), od.eval_m2lnp_batch(Xe), rtol=1e-9)


MIXED_CASES = [
    # kernel, d (padded MFMA dim 4, 8, 16, 32, 64), n, nu
    (GAUSS, 3, 700, 0.0),
    (ST, 8, 1500, 3.0),
    (ST, 12, 900, 1.0),
    (GAUSS, 16, 2000, 0.0),
    (ST, 24, 1200, 5.0),
    (ST, 40, 1300, 5.0),
    # d > 16: from 1024 kernels on, a pilot pass over 256 of them sets the
    # main pass's per-point reference; below, the underflowed points run twice
    (ST, 33, 800, 2.0),
    (GAUSS, 60, 1100, 0.0),
]


@pytest.mark.parametrize("kind", [1, 0])
@pytest.mark.parametrize("kernel,d,n,nu", MIXED_CASES)
def test_eval_m2lnp_mixed_precision(gpu, oracle, kind, kernel, d, n, nu):
    """FP16-split MFMA distances with FP32 kernel sums: relative m2lnp error
    |dm2lnp| / max(|m2lnp|, 1) <= 1e-6 against the FP64 oracle (VKDE, and
    KDE, whose kernels share the global factor)."""
    _say(f"[eval mixed] kind={kind} kernel={kernel} d={d} n={n}")
    g = np.random.default_rng(5000 + d)
    tgt = synth.Gaussian(d, g)
    X = tgt.draw(g, n)
    sd, od = _pair(gpu, oracle, kind, kernel, d, X, nu)
    sd.prepare_interp(tgt.m2lnp(X)) if d <= 16 else sd.prepare()
    od.prepare()
    od.set_weights(sd.peek_weights())
    Xe = tgt.draw(g, 1000)
    Xe[:5] = X[:5]                 # on kernel centres
    Xe[5] = 50.0 * Xe[5]           # far outlier: per-point rescaled path
    Xe[6] = 1.0e4 * Xe[6]
    ref = od.eval_m2lnp_batch(Xe)
    got = sd.eval_m2lnp_batch(Xe, gpu.Precision.MIXED)
    err = np.abs(got - ref) / np.maximum(np.abs(ref), 1.0)
    _say(f"   mixed max rel err {err.max():.3e} p99 {np.quantile(err, 0.99):.3e}")
    assert np.all(np.isfinite(got))
    assert err.max() <= 1e-6
    # batch composition does not change a point's result beyond the tolerance
    alone = sd.eval_m2lnp_batch(Xe[10:11], gpu.Precision.MIXED)
    assert abs(alone[0] - got[10]) <= 1e-6 * max(abs(ref[10]), 1.0)


FP32_CASES = [(GAUSS, 3, 700, 0.0), (ST, 8, 1500, 3.0), (ST, 16, 2000, 1.0), (ST, 40, 1300, 5.0), (ST, 64, 1500, 5.0)]


@pytest.mark.parametrize("kind", [1, 0])
@pytest.mark.parametrize("kernel,d,n,nu", FP32_CASES)
def test_eval_m2lnp_fp32_rung(gpu, oracle, kind, kernel, d, n, nu):
    """The FP32 rung of the precision ladder (single-precision distances and
    kernel sums, FP64 differences x - theta_i): within 1e-6 of the FP64
    oracle, including points on kernel centres and far outliers; it follows
    a refit's new weights."""
    _say(f"[eval fp32] kind={kind} kernel={kernel} d={d} n={n}")
    g = np.random.default_rng(5200 + d)
    tgt = synth.Gaussian(d, g)
    X = tgt.draw(g, n)
    sd, od = _pair(gpu, oracle, kind, kernel, d, X, nu)
    sd.prepare_interp(tgt.m2lnp(X)) if d <= 16 else sd.prepare()
    od.prepare()
    od.set_weights(sd.peek_weights())
    Xe = tgt.draw(g, 1000)
    Xe[:5] = X[:5]
    Xe[5] = 50.0 * Xe[5]
    ref = od.eval_m2lnp_batch(Xe)
    got = sd.eval_m2lnp_batch(Xe, gpu.Precision.FP32)
    err = np.abs(got - ref) / np.maximum(np.abs(ref), 1.0)
    _say(f"   fp32 max rel err {err.max():.3e} p99 {np.quantile(err, 0.99):.3e}")
    assert np.all(np.isfinite(got))
    assert err.max() <= 1e-6
    # the linear density carries the same error in its logarithm:
    # |d ln p| = |d m2lnp| / 2
    lin, lref = sd.eval_batch(Xe[:50], gpu.Precision.FP32), od.eval_batch(Xe[:50])
    live = lref > 1e-300
    assert np.all(np.abs(np.log(lin[live]) - np.log(lref[live])) <= 0.5e-6 * np.maximum(np.abs(ref[:50][live]), 1.0))
    # new weights after a refit reach the FP32 factors
    if d <= 16:
        X2 = tgt.draw(g, n)
        sd.reset()
        sd.add_obs_array(X2)
        sd.prepare_interp(tgt.m2lnp(X2))
        od2 = oracle.StatsDist(od.kind, od.kernel, d, nu)
        od2.add_obs_array(X2)
        od2.prepare()
        od2.set_weights(sd.peek_weights())
        ref2 = od2.eval_m2lnp_batch(Xe[:200])
        got2 = sd.eval_m2lnp_batch(Xe[:200], gpu.Precision.FP32)
        assert np.max(np.abs(got2 - ref2) / np.maximum(np.abs(ref2), 1.0)) <= 1e-6


def test_mixed_after_refit_uses_new_weights(gpu, oracle):
    g = np.random.default_rng(5100)
    d, n = 8, 800
    tgt = synth.Banana(d)
    X = tgt.draw(g, n)
    sd, od = _pair(gpu, oracle, 1, ST, d, X, 3.0)
    sd.prepare()
    Xe = tgt.draw(g, 300)
    a = sd.eval_m2lnp_batch(Xe, gpu.Precision.MIXED)
    sd.prepare_interp(tgt.m2lnp(X))
    b = sd.eval_m2lnp_batch(Xe, gpu.Precision.MIXED)
    od.prepare()
    od.set_weights(sd.peek_weights())
    ref = od.eval_m2lnp_batch(Xe)
    assert np.max(np.abs(b - ref) / np.maximum(np.abs(ref), 1.0)) <= 1e-6
    assert np.max(np.abs(a - b)) > 1e-3
