"""Robust covariance types on the device (ncm_stats_vec.c:1834-2064 via
ncm_stats_dist_kde.c:423-441 and ncm_stats_dist_vkde.c:467-472).

The device Qn, pair scales, Jacobi rotation sequence and OGK assembly follow
the oracle operation for operation, so the prepared factors are compared
bit for bit, and FP64 eval to 1e-12.
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


def _pair(oracle, kind, kernel, d, nu, cov_type, X):
    import numcosmo_amd as na
    sdk = na.StatsDistKernelGauss.new(d) if kernel == GAUSS else na.StatsDistKernelST.new(d, nu)
    sd = (na.StatsDistVKDE if kind == 1 else na.StatsDistKDE).new(sdk)
    sd.set_cov_type(cov_type)
    od = oracle.StatsDist(kind, kernel, d, nu)
    od.set_cov_type(cov_type)
    sd.add_obs_array(X)
    od.add_obs_array(X)
    return sd, od


@pytest.mark.parametrize("kind,kernel,d,n,nu,cov_type", [
    (0, GAUSS, 3, 300, 0.0, 2), (0, GAUSS, 3, 300, 0.0, 3), (0, ST, 5, 257, 3.0, 3),
    (1, ST, 4, 600, 3.0, 2), (1, ST, 4, 600, 3.0, 3), (1, GAUSS, 8, 900, 0.0, 3),
    (0, GAUSS, 3, 3000, 0.0, 3),   # 3000 values per set: the block path
])
def test_robust_prepare_bit_exact(oracle, kind, kernel, d, n, nu, cov_type):
    _say(f"[robust] kind={kind} kernel={kernel} d={d} n={n} cov_type={cov_type}")
    g = np.random.default_rng(5000 + d + n + cov_type)
    X = synth.Banana(d).draw(g, n)
    X[: n // 10] += 8.0          # a cluster of outliers
    sd, od = _pair(oracle, kind, kernel, d, nu, cov_type, X)
    sd.prepare()
    od.prepare()
    np.testing.assert_array_equal(sd.peek_full_cov(), od.get_full_cov())
    np.testing.assert_array_equal(np.triu(sd.peek_full_cov_decomp()), np.triu(od.get_full_cov_decomp()))
    np.testing.assert_array_equal(sd.peek_invUsample(), od.get_invUsample())
    if kind == 1:
        for i in range(0, n, max(1, n // 40)):
            np.testing.assert_array_equal(sd.peek_cov_decomp(i), od.get_cov_decomp(i))
    Xe = synth.Banana(d).draw(g, 500)
    got, ref = sd.eval_m2lnp_batch(Xe), od.eval_m2lnp_batch(Xe)
    assert np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1.0)) <= 1e-12


def test_robust_too_few_neighbours(oracle):
    import numcosmo_amd as na
    sd = na.StatsDistVKDE.new(na.StatsDistKernelST.new(2, 3.0))
    sd.set_cov_type(3)
    sd.add_obs_array(np.random.default_rng(1).normal(size=(60, 2)))   # k = 3 neighbours
    with pytest.raises(na.NckdeError, match="too few points to estimate the covariance"):
        sd.prepare()
