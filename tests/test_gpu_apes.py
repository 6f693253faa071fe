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


@pytest.mark.parametrize("ktype,precision,rtol", [(0, 0, 1e-7), (2, 0, 1e-7), (1, 1, 3e-6)])
def test_apes_moves_match_oracle(oracle, ktype, precision, rtol):
    """Every move of a short bounded run, replayed on the oracle: the state
    each half-step starts from (rebuilt from the initial sample and the
    recorded decisions), the two transition densities
    (ncm_fit_esmcmc_walker_apes.c:819-881), the exploration countdown
    (:815-816) and the acceptance rule (:884-919)."""
    import math
    from numcosmo_amd.apes import APES, InterpolationKernel
    dim, nw, steps, expl, rw, os_, lf = 2, 120, 4, 3, 0.25, 0.2, 0.2
    _say(f"[apes] trace vs oracle ktype={ktype} precision={precision} nw={nw} steps={steps}")
    g = np.random.default_rng(17 + ktype)
    tgt = synth.Gaussian(dim, g)
    sig = np.sqrt(np.diag(tgt.cov))
    lb, ub = tgt.mean - 1.2 * sig, tgt.mean + 1.8 * sig
    s = APES(nwalkers=nw, ndim=dim, log_prob_batch=lambda th: -0.5 * tgt.m2lnp(th),
             interpolation_kernel=InterpolationKernel(ktype), over_smooth=os_, local_fraction=lf,
             random_walk_prob=rw, bounds=(lb, ub), precision=precision, seed=5 + ktype)
    s.set_exploration(expl)
    x0 = tgt.draw(g, 20 * nw)
    x0 = x0[np.all((x0 >= lb) & (x0 <= ub), axis=1)][:nw]
    assert x0.shape == (nw, dim)
    s.run_mcmc(x0, steps, trace=True)
    T = s.trace
    assert T.shape == (steps * nw, 2 * dim + 9)

    kern, nu = {0: (oracle.ST, 1.0), 1: (oracle.ST, 3.0), 2: (oracle.GAUSS, 3.0)}[ktype]
    n2 = nw // 2
    theta, m2 = x0.copy(), s.m2lnL(x0)
    worst, decided = 0.0, 0
    for h in range(2 * steps):
        which = h % 2
        k0, other0 = (0, n2) if which == 0 else (n2, 0)
        R = T[h * n2:(h + 1) * n2]
        th, ts = R[:, 9:9 + dim], R[:, 9 + dim:]
        np.testing.assert_array_equal(R[:, 0], which)
        np.testing.assert_array_equal(R[:, 1], np.arange(k0, k0 + n2))
        np.testing.assert_array_equal(th, theta[k0:k0 + n2])
        np.testing.assert_array_equal(R[:, 3], m2[k0:k0 + n2])
        # exploration = 3 covers the first two half-steps
        exploring = h + 1 < expl
        np.testing.assert_array_equal(R[:, 2], 1.0 if exploring else 0.0)
        assert np.all((ts >= lb) & (ts <= ub))
        np.testing.assert_allclose(R[:, 4], s.m2lnL(ts), rtol=1e-14)

        od = oracle.StatsDist(oracle.VKDE, kern, dim, nu)
        od.set_over_smooth(os_)
        od.set_local_frac(lf)
        od.set_shrink(0.01)
        od.add_obs_array(theta[other0:other0 + n2])
        od.prepare_interp(m2[other0:other0 + n2])
        std = oracle.apes_rw_std(od.get_full_cov())
        for j in range(n2):
            ps = oracle.apes_transition_m2lnp(od, th[j], ts[j], lb, ub, std, rw)
            pc = oracle.apes_transition_m2lnp(od, ts[j], th[j], lb, ub, std, rw)
            e = max(abs(R[j, 5] - ps) / max(abs(ps), 1.0), abs(R[j, 6] - pc) / max(abs(pc), 1.0))
            worst = max(worst, e)
            assert e <= rtol, (h, j, R[j, 5], ps, R[j, 6], pc)
            m2lnq = 0.0 if exploring else pc - ps
            prob = min(math.exp(-0.5 * (R[j, 4] - R[j, 3] + m2lnq)), 1.0)
            if abs(R[j, 7] - prob) > 1e-4:
                assert R[j, 8] == float(R[j, 7] < prob), (h, j, R[j, 7], prob)
                decided += 1
        acc = R[:, 8] == 1.0
        theta[k0:k0 + n2][acc] = ts[acc]
        m2[k0:k0 + n2][acc] = R[acc, 4]
        if which == 1:
            np.testing.assert_array_equal(s.chain[h // 2, :, 1:], theta)
            np.testing.assert_array_equal(s.chain[h // 2, :, 0], m2)
    _say(f"   worst transition rel err {worst:.2e}, {decided} of {steps * nw} decisions checked")
    assert decided >= 0.9 * steps * nw
    assert 0 < T[:, 8].sum() < steps * nw
