"""APES ensemble sampler on the device engine.

Statistical checks in the manner of the reference's
tests/c/ncm/fit/test_ncm_fit_esmcmc.c (nwalkers = 100 dim; posterior
var(m2lnL) ~ 2 dim at reltol 0.6, :613,652,703), plus the target's moments
and the parameter bounds.
"""
import sys

import numpy as np
import pytest

from numcosmo_amd import synth

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _run(dim, ktype, steps, precision=0, bounds=None, seed=11, rw=0.02, start="target"):
    from numcosmo_amd.apes import APES, InterpolationKernel
    g = np.random.default_rng(seed)
    tgt = synth.Gaussian(dim, g)
    nw = 100 * dim
    sampler = APES(nwalkers=nw, ndim=dim, log_prob_batch=lambda th: -0.5 * tgt.m2lnp(th),
                   interpolation_kernel=InterpolationKernel(ktype), over_smooth=0.2,
                   precision=precision, bounds=bounds, seed=seed, random_walk_prob=rw)
    if start == "target":
        x0 = tgt.draw(g, nw)
    else:
        x0 = g.normal(size=(nw, dim)) * 0.1 + 0.5
    sampler.run_mcmc(x0, steps)
    return tgt, sampler


def _check_moments(tgt, th, m2, dim, mean_tol=0.2):
    sd = np.sqrt(np.diag(tgt.cov))
    _say(f"   mean offset / sd {np.abs(th.mean(0) - tgt.mean) / sd}")
    assert np.all(np.abs(th.mean(0) - tgt.mean) < mean_tol * sd)
    assert np.allclose(np.cov(th, rowvar=False), tgt.cov, rtol=0.3, atol=0.1 * sd.max() ** 2)
    # var(m2lnL) = 2 dim for a Gaussian posterior
    assert np.var(m2) == pytest.approx(2.0 * dim, rel=0.6)


@pytest.mark.parametrize("dim,ktype", [(2, 0), (3, 1), (2, 2)])
def test_apes_stationary(dim, ktype):
    """Walkers started from the target must stay distributed as the target:
    a wrong transition density or acceptance rule makes the ensemble drift."""
    _say(f"[apes] stationary dim={dim} ktype={ktype}")
    tgt, s = _run(dim, ktype, 60)
    chain = s.chain[10:]
    st = s.get_stats()
    _say(f"   acceptance {st['acceptance']:.3f} seconds {st['seconds']}")
    assert 0.05 < st["acceptance"] <= 1.0
    _check_moments(tgt, chain[..., 1:].reshape(-1, dim), chain[..., 0].ravel(), dim)


def test_apes_converges_from_a_ball():
    _say("[apes] burn-in from a small ball, dim=2 Cauchy kernel")
    tgt, s = _run(2, 0, 300, start="ball")
    chain = s.chain[200:]
    _check_moments(tgt, chain[..., 1:].reshape(-1, 2), chain[..., 0].ravel(), 2, mean_tol=0.25)


def test_apes_bounds_respected():
    lb = np.array([0.0, -1.0])
    ub = np.array([2.0, 1.5])
    _, s = _run(2, 0, 40, bounds=(lb, ub), rw=0.2, start="ball")
    th = s.chain[..., 1:].reshape(-1, 2)
    assert np.all(th >= lb) and np.all(th <= ub)


def test_apes_mixed_precision_stationary():
    _say("[apes] stationary dim=3 ST3 kernel, mixed-precision transition densities")
    tgt, s = _run(3, 1, 60, precision=1, seed=5)
    chain = s.chain[10:]
    _check_moments(tgt, chain[..., 1:].reshape(-1, 3), chain[..., 0].ravel(), 3)


def test_apes_is_reproducible_for_a_seed():
    _, s1 = _run(2, 1, 8, seed=21)
    _, s2 = _run(2, 1, 8, seed=21)
    np.testing.assert_array_equal(s1.chain, s2.chain)


def test_apes_chain_as_catalog(tmp_path):
    # the run's chain in NcmMSetCatalog's FITS layout: rows ordered by step
    # then walker, NcmFit:m2lnL first
    from numcosmo_amd.catalog import MSetCatalog
    _, s = _run(2, 0, 6, seed=31)
    p = str(tmp_path / "apes.fits")
    s.write_catalog(p, ["NcmModelMVND:mu_0", "NcmModelMVND:mu_1"])
    cat = MSetCatalog.read(p)
    nw = s.nwalkers
    assert cat.nchains == nw and len(cat) == 7 * nw   # the initial sample, then 6 steps
    assert cat.param_names == ["NcmModelMVND:mu_0", "NcmModelMVND:mu_1"]
    np.testing.assert_array_equal(cat.as_array()[nw:], s.chain.reshape(-1, 3))
    np.testing.assert_array_equal(cat.as_array(), s.get_catalog().as_array())
    th, m2 = s.get_state()
    np.testing.assert_array_equal(cat.as_array()[-nw:, 0], m2)
    np.testing.assert_array_equal(cat.as_array()[-nw:, 1:], th)
    cat.trim(1, 2)
    assert len(cat) == 3 * nw and cat.first_id == nw
    np.testing.assert_array_equal(cat.as_array(), s.chain[0::2].reshape(-1, 3))
