"""Pins of the oracle's cross-validated bandwidth restatement (CPU only).

What fixes each part:
- Philox4x32-10: the known-answer vectors of Random123 (Salmon et al., SC'11).
- Nelder-Mead (GSL nmsimplex2 in one dimension, the minimiser of
  ncm_stats_dist.c:664-701):
  - a trajectory derived by hand for f(x) = (x - 0.33)^2 from x0 = 0, step 0.1;
  - agreement with Brent's method to the simplex size tolerance.
- levmar's dlevmar_dif for one parameter (ncm_stats_dist.c:1063-1065):
  - the first two damped steps on a linear residual, which have a closed form
    (initial mu = 1e-3 max diag JtJ, then mu/3 after an exact-model step);
  - agreement with MINPACK's Levenberg-Marquardt (scipy) on a nonlinear fit.
- NcmStatsVec two-variable statistics (ncm_stats_vec.c:510-565, 2254-2297): numpy.
- Leave-one-out objectives:
  - KDE Gauss AMISE (ncm_stats_dist.c:512-559) against Gaussian-mixture closed
    forms evaluated with scipy;
  - the Monte-Carlo AMISE (ncm_stats_dist.c:563-660) against the VKDE Gauss
    closed form, within its own stopping error.
- Split fits:
  - SPLIT_NOFIT's result against a grid scan of the held-out objective;
  - every CV type against the reference's density-interpolation check
    (test_ncm_stats_dist.c:452-491, 677-838): Jensen-Shannon estimate below 0.5.
"""
import math

import numpy as np
import pytest
import scipy.optimize
import scipy.stats


def _gauss_target(rng, d, n):
    A = rng.standard_normal((d, d))
    S = A @ A.T + d * np.eye(d)
    L = np.linalg.cholesky(S)
    X = rng.standard_normal((n, d)) @ L.T
    Si = np.linalg.inv(S)
    c = math.log(np.linalg.det(2 * math.pi * S))
    return S, L, X, (lambda Y: np.einsum("ij,jk,ik->i", Y, Si, Y) + c)


# ---- Philox ------------------------------------------------------------------
def test_philox_known_answers(oracle):
    kat = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, out in kat:
        assert oracle.philox4x32_10(ctr, key) == list(out)


def test_draw_stream_layout_and_moments(oracle):
    # the uniform of draw j is the top 53 bits of block (j, 0, ctr_lo, ctr_hi)
    for key, ctr, j in [(0, 0, 0), (7, 3, 11), (2**40 + 5, 2**33 + 1, 4095)]:
        u, z, c = oracle.draw(key, ctr, j, 3, 2)
        r = oracle.philox4x32_10((j, 0, ctr & 0xFFFFFFFF, ctr >> 32), (key & 0xFFFFFFFF, key >> 32))
        assert u == ((r[0] >> 5) * 67108864.0 + (r[1] >> 6)) / 9007199254740992.0
    zs, cs, us = [], [], []
    for j in range(6000):
        u, z, c = oracle.draw(0, 9, j, 3, 3)
        us.append(u)
        zs.append(z)
        cs.append(c)
    zs = np.array(zs).ravel()
    assert 0.0 <= min(us) and max(us) < 1.0
    assert abs(np.mean(us) - 0.5) < 0.02
    assert abs(zs.mean()) < 0.03 and abs(zs.var() - 1.0) < 0.04
    assert abs(np.mean(cs) - 3.0) < 0.1
    assert scipy.stats.kstest(zs, "norm").pvalue > 1e-3


# ---- Nelder-Mead ---------------------------------------------------------------
def test_nm_trajectory_by_hand(oracle):
    seen = []

    def f(x):
        seen.append(x)
        return (x - 0.33) ** 2

    oracle.nm_minimize1(f, 0.0, 0.1, 1e-3, 4)
    # start corners, reflection + expansion, reflection accepted above the
    # second point then contraction to the centre, contraction, contraction
    expect = [0.0, 0.1, 0.2, 0.3, 0.5, 0.4, 0.2, 0.35, 0.4, 0.325]
    np.testing.assert_allclose(seen, expect, atol=1e-12)


@pytest.mark.parametrize("fn,x0", [
    (lambda x: x ** 4 - 3 * x + math.cos(x), 0.0),
    (lambda x: math.exp(x) - 2.5 * x, 1.2),
    (lambda x: (x + 1.7) ** 2 * (1 + 0.1 * math.sin(5 * x)), -0.5),
])
def test_nm_agrees_with_brent(oracle, fn, x0):
    x, fv, it, st = oracle.nm_minimize1(fn, x0, 0.1, 1e-3, 1000)
    xb = scipy.optimize.minimize_scalar(fn, bracket=(x0 - 3, x0 + 3), tol=1e-12).x
    assert st == 0
    assert abs(x - xb) < 3e-3
    assert fv == fn(x)


# ---- levmar ----------------------------------------------------------------------
def test_lm_first_steps_closed_form(oracle):
    rng = np.random.default_rng(3)
    a = rng.uniform(0.5, 2.0, 40)
    b = a * 1.7 + rng.normal(0, 0.05, 40)
    pstar = (a @ b) / (a @ a)
    res = lambda p: a * p - b
    p0 = 0.2
    # one step: Dp = J^T e / (J^T J + mu), mu = 1e-3 J^T J
    p1, stop, it = oracle.lm_dif1(res, p0, 40, itmax=1)
    assert it == 1 and stop == 3
    assert p1 == pytest.approx(p0 + (pstar - p0) / 1.001, rel=1e-9)
    # an exact-model step scales mu by 1/3 before the second step
    p2, stop, it = oracle.lm_dif1(res, p0, 40, itmax=2)
    assert p2 == pytest.approx(p1 + (pstar - p1) / (1.0 + 1e-3 / 3.0), rel=1e-9)


def test_lm_agrees_with_minpack(oracle):
    rng = np.random.default_rng(8)
    t = np.linspace(0.0, 4.0, 60)
    y = np.exp(-0.83 * t) + rng.normal(0.0, 0.01, t.size)
    res = lambda p: np.exp(-p * t) - y
    p, stop, it = oracle.lm_dif1(res, 0.2, t.size)
    ref = scipy.optimize.least_squares(lambda q: res(q[0]), [0.2], method="lm", xtol=1e-15, ftol=1e-15,
                                       gtol=1e-15).x[0]
    assert stop in (1, 2)
    assert p == pytest.approx(ref, rel=1e-6)


def test_pair_stats_match_numpy(oracle):
    rng = np.random.default_rng(2)
    x0 = rng.lognormal(size=500)
    x1 = 0.3 * x0 + rng.lognormal(size=500)
    m0, m1, v0, v1, c01 = oracle.pair_stats(x0, x1)
    C = np.cov(np.vstack([x0, x1]))
    assert m0 == pytest.approx(x0.mean(), rel=1e-13)
    assert m1 == pytest.approx(x1.mean(), rel=1e-13)
    assert v0 == pytest.approx(C[0, 0], rel=1e-12)
    assert v1 == pytest.approx(C[1, 1], rel=1e-12)
    assert c01 == pytest.approx(C[0, 1], rel=1e-12)


# ---- objectives ------------------------------------------------------------------
def _prepared(oracle, kind, kernel, d, X, nu=3.0, cv=0, local_frac=None):
    sd = oracle.StatsDist(kind, kernel, d, nu)
    if local_frac is not None:
        sd.set_local_frac(local_frac)
    sd.add_obs_array(X)
    sd.prepare()
    sd.set_cv_type(cv)
    return sd


@pytest.mark.parametrize("d", [1, 2, 3])
def test_amise_kde_gauss_closed_form(oracle, d):
    rng = np.random.default_rng(10 + d)
    S, L, X, _ = _gauss_target(rng, d, 70)
    sd = _prepared(oracle, oracle.KDE, oracle.GAUSS, d, X, cv=oracle.CV_LOO)
    Sg = np.cov(X.T, ddof=1).reshape(d, d)
    n = X.shape[0]
    D = X[:, None, :] - X[None, :, :]
    for lnos in (-0.4, 0.0, 0.3):
        val = sd.cv_objective(lnos)
        h = sd.get_href()   # the objective leaves href at over_smooth * rot(n)
        int_p2 = scipy.stats.multivariate_normal(np.zeros(d), 2 * h * h * Sg).pdf(D.reshape(-1, d)).sum() / n ** 2
        K = scipy.stats.multivariate_normal(np.zeros(d), h * h * Sg).pdf(D.reshape(-1, d)).reshape(n, n)
        loo = (K.sum() - np.trace(K)) * 2.0 / (n * (n - 1))
        assert val == pytest.approx(int_p2 - loo, rel=1e-11)


def test_amise_monte_carlo_closed_form(oracle):
    d = 2
    rng = np.random.default_rng(21)
    S, L, X, _ = _gauss_target(rng, d, 90)
    sd = _prepared(oracle, oracle.VKDE, oracle.GAUSS, d, X, cv=oracle.CV_LOO, local_frac=0.2)
    n = X.shape[0]
    for lnos in (-0.3, 0.2):
        val = sd.cv_objective(lnos)
        h = sd.get_href()
        covs = []
        for i in range(n):
            U = np.triu(sd.get_cov_decomp(i))   # the factor is stored in place (upper triangle)
            covs.append(h * h * (U.T @ U))
        int_p2 = 0.0
        loo = 0.0
        for i in range(n):
            for j in range(n):
                int_p2 += scipy.stats.multivariate_normal(X[j], covs[i] + covs[j]).pdf(X[i]) / n ** 2
                if i != j:
                    loo += scipy.stats.multivariate_normal(X[j], covs[j]).pdf(X[i])
        exact = int_p2 - 2.0 * loo / (n * (n - 1))
        # the integral stops at a 1% relative standard error
        assert abs(val - exact) < 0.04 * int_p2


def test_split_nofit_minimises_held_out_objective(oracle):
    d = 2
    rng = np.random.default_rng(4)
    S, L, X, m2f = _gauss_target(rng, d, 240)
    sd = oracle.StatsDist(oracle.KDE, oracle.GAUSS, d, 3.0, oracle.CV_SPLIT_NOFIT)
    sd.add_obs_array(X)
    sd.prepare()
    assert sd.get_n_kernels() == 120
    lnos = math.log(sd.get_over_smooth())
    grid = np.linspace(lnos - 1.0, lnos + 1.0, 201)
    vals = [sd.cv_objective(g) for g in grid]
    best = grid[int(np.argmin(vals))]
    assert abs(lnos - best) < 0.02
    assert sd.cv_objective(lnos) <= min(vals) + 1e-3 * abs(min(vals))


def _js(oracle, sd, d, L, m2f, rng, ntests=500, nu=3.0, kernel=0):
    Y = rng.standard_normal((ntests, d)) @ L.T
    a0 = 0.25 * (sd.eval_m2lnp_batch(Y) - m2f(Y))
    Ys = []
    for _ in range(ntests):
        z = rng.standard_normal(d)
        c = rng.chisquare(nu) if kernel == oracle.ST else 1.0
        Ys.append(sd.sample(rng.uniform(), z, c)[1])
    Ys = np.array(Ys)
    a1 = 0.25 * (m2f(Ys) - sd.eval_m2lnp_batch(Ys))
    return 0.5 * (np.mean(np.log1p(np.tanh(a0))) + np.mean(np.log1p(np.tanh(a1))))


@pytest.mark.parametrize("kind,kernel,cv", [
    (0, 0, 1), (0, 0, 2), (0, 0, 3), (1, 0, 1), (1, 1, 2), (1, 1, 3),
])
def test_dens_interp_cv_types(oracle, kind, kernel, cv):
    d = 2
    rng = np.random.default_rng(100 + 10 * kind + 3 * kernel + cv)
    S, L, X, m2f = _gauss_target(rng, d, 400)
    sd = oracle.StatsDist(kind, kernel, d, 3.0, cv)
    if kind == oracle.VKDE:
        sd.set_local_frac(0.1)
    sd.add_obs_array(X)
    sd.prepare_interp(m2f(X))
    w = sd.get_weights()
    assert np.all(w > 0) and w.sum() == pytest.approx(1.0, rel=1e-12)
    assert sd.get_n_kernels() == (200 if cv in (1, 2) else 400)
    assert 0.1 < sd.get_over_smooth() < 10.0
    assert _js(oracle, sd, d, L, m2f, rng, kernel=kernel) < 0.5
