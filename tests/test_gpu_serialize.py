"""Serialisation of configuration, sample and prepared state.

Follows the reference's serialisation test (test_ncm_stats_dist.c:932-990):
an object restored from its serialisation and fed the same sample gives the
same fit; and, beyond the reference, a restored prepared state evaluates and
samples bit for bit like the original without a new prepare.
"""
import sys

import numpy as np
import pytest

from numcosmo_amd import synth

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _vkde(d=4, nu=3.0):
    import numcosmo_amd as na
    sd = na.StatsDistVKDE.new(na.StatsDistKernelST.new(d, nu))
    sd.set_over_smooth(0.7)
    sd.set_shrink(0.02)
    sd.set_local_frac(0.1)
    return sd


def test_config_round_trip_gives_the_same_fit():
    from numcosmo_amd.ncm_stats_dist import deserialize
    _say("[serialize] configuration only")
    g = np.random.default_rng(60)
    tgt = synth.Gaussian(4, g)
    X = tgt.draw(g, 500)
    m2 = tgt.m2lnp(X)
    sd = _vkde()
    dup = deserialize(sd.serialize(with_state=False))
    assert dup.get_over_smooth() == 0.7 and dup.get_local_frac() == 0.1
    sd.add_obs_array(X)
    dup.add_obs_array(X)
    sd.prepare_interp(m2)
    dup.prepare_interp(m2)
    np.testing.assert_array_equal(sd.peek_weights(), dup.peek_weights())
    Xe = tgt.draw(g, 300)
    np.testing.assert_array_equal(sd.eval_m2lnp_batch(Xe), dup.eval_m2lnp_batch(Xe))


@pytest.mark.parametrize("kind", ["vkde", "kde_fixed"])
def test_prepared_state_round_trip(kind):
    import numcosmo_amd as na
    from numcosmo_amd.ncm_stats_dist import deserialize
    _say(f"[serialize] prepared state, {kind}")
    g = np.random.default_rng(61)
    tgt = synth.Gaussian(5, g)
    X = tgt.draw(g, 800)
    if kind == "vkde":
        sd = _vkde(5)
    else:
        sd = na.StatsDistKDE.new(na.StatsDistKernelGauss.new(5))
        sd.set_cov_type(na.StatsDistKDECovType.FIXED)
        sd.set_cov_fixed(tgt.cov)
    sd.add_obs_array(X)
    sd.prepare_interp(tgt.m2lnp(X))
    blob = sd.serialize(with_state=True)
    dup = deserialize(blob)
    np.testing.assert_array_equal(sd.peek_weights(), dup.peek_weights())
    assert dup.get_href() == sd.get_href() and dup.get_rnorm() == sd.get_rnorm()
    Xe = tgt.draw(g, 1000)
    np.testing.assert_array_equal(sd.eval_m2lnp_batch(Xe), dup.eval_m2lnp_batch(Xe))
    if kind == "vkde":
        np.testing.assert_array_equal(sd.eval_m2lnp_batch(Xe, precision=na.Precision.MIXED),
                                      dup.eval_m2lnp_batch(Xe, precision=na.Precision.MIXED))
        np.testing.assert_array_equal(sd.peek_knn(), dup.peek_knn())
    u = g.random(400)
    z = g.standard_normal((400, 5))
    np.testing.assert_array_equal(sd.kernel_choose_batch(u), dup.kernel_choose_batch(u))
    a, ia = sd.sample_batch(u, z, g.chisquare(3.0, 400) if kind == "vkde" else None)
    # the restored object serialises to the same bytes
    assert dup.serialize(with_state=True) == blob


def test_truncated_blob_is_rejected():
    import numcosmo_amd as na
    from numcosmo_amd.ncm_stats_dist import deserialize
    sd = _vkde()
    sd.add_obs_array(np.random.default_rng(62).normal(size=(200, 4)))
    sd.prepare()
    blob = sd.serialize()
    with pytest.raises(na.NckdeError, match="truncated|invalid"):
        deserialize(blob[: len(blob) - 100])
