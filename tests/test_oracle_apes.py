"""The oracle's APES transition density against scipy (no GPU).

ncm_fit_esmcmc_walker_apes.c:819-859 mixes the kernel density with a
truncated Gaussian random walk; ncm_util.c:646-690 gives the truncation
integral.  Both are pinned here against scipy's normal and truncated-normal
distributions, which compute the same quantities by other routes: a
truncation centred at the wrong end, a dropped normalisation or a mixture
with swapped weights fails one of these checks.
"""
import math

import numpy as np
import pytest
from scipy import special, stats


def _log_interval_scipy(zl, zu):
    lo, hi = min(zl, zu), max(zl, zu)
    if lo > 0:
        return math.log(special.ndtr(-lo) - special.ndtr(-hi))
    return math.log(special.ndtr(hi) - special.ndtr(lo))


@pytest.mark.parametrize("zl,zu", [(-1.0, 2.0), (4.5, 6.0), (-7.0, -5.0), (-5.0, 5.0), (0.5, 10.0),
                                   (-10.0, -4.3), (-0.2, 0.1), (3.0, 40.0), (-4.0, 4.5), (6.0, 4.5)])
def test_log_normal_gaussian_integral(oracle, zl, zu):
    got = oracle.log_normal_gaussian_integral(zl, zu)
    assert got == pytest.approx(_log_interval_scipy(zl, zu), rel=1e-12, abs=1e-13)


def test_log_normal_gaussian_integral_empty(oracle):
    assert oracle.log_normal_gaussian_integral(1.5, 1.5) == -math.inf


def _sd(oracle):
    rng = np.random.default_rng(3)
    X = rng.standard_normal((80, 2)) @ np.array([[1.0, 0.0], [0.6, 0.5]]).T
    sd = oracle.StatsDist(oracle.VKDE, oracle.ST, 2, 3.0)
    sd.set_over_smooth(0.2)
    sd.set_local_frac(0.2)
    sd.add_obs_array(X)
    sd.prepare()
    return sd, X


def test_rw_std_is_a_quarter_of_the_sample_sigma(oracle):
    sd, X = _sd(oracle)
    np.testing.assert_allclose(oracle.apes_rw_std(sd.get_full_cov()),
                               0.25 * np.sqrt(np.diag(np.cov(X, rowvar=False, ddof=1))), rtol=1e-12)


def test_transition_without_random_walk_is_the_density(oracle):
    sd, _ = _sd(oracle)
    x, y = np.array([0.1, -0.3]), np.array([0.5, 0.2])
    assert oracle.apes_transition_m2lnp(sd, x, y, [-1, -1], [1, 1], [0.2, 0.2], 0.0) == sd.eval_m2lnp(y)


@pytest.mark.parametrize("lb,ub", [([-1.0, -0.8], [1.2, 0.9]), ([-math.inf, -0.5], [0.7, math.inf]),
                                   ([-math.inf, -math.inf], [math.inf, math.inf])])
def test_transition_matches_scipy_mixture(oracle, lb, ub):
    sd, _ = _sd(oracle)
    std = oracle.apes_rw_std(sd.get_full_cov())
    p = 0.3
    rng = np.random.default_rng(8)
    for _ in range(20):
        lo = np.maximum(np.array(lb), -2.0)
        hi = np.minimum(np.array(ub), 2.0)
        x, y = rng.uniform(lo, hi), rng.uniform(lo, hi)
        for a, b in ((x, y), (y, x)):
            # the walk from a to b: a Gaussian centred at a, truncated to [lb, ub]
            rw = np.prod([stats.truncnorm.pdf(b[i], (lb[i] - a[i]) / std[i], (ub[i] - a[i]) / std[i],
                                              loc=a[i], scale=std[i]) for i in range(2)])
            want = -2.0 * math.log((1.0 - p) * math.exp(-0.5 * sd.eval_m2lnp(b)) + p * rw)
            got = oracle.apes_transition_m2lnp(sd, a, b, lb, ub, std, p)
            assert got == pytest.approx(want, rel=1e-10, abs=1e-10)
