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


def test_apes_resume_continues_the_run(tmp_path):
    """A run stopped after 3 steps, written to FITS and resumed to 7
    ensembles is the uninterrupted 6-step run, row for row: the walkers start
    from the catalog's last ensemble and the sampler's streams go on."""
    from numcosmo_amd.catalog import MSetCatalog
    _say("[apes] resume from a FITS catalog vs the uninterrupted run, dim=2 ST3")
    _, a = _run(2, 1, 6, seed=41)
    _, b = _run(2, 1, 3, seed=41)
    nw = b.nwalkers
    p = str(tmp_path / "b.fits")
    b.write_catalog(p)
    b.resume(p, 7)
    cat = MSetCatalog.read(p)
    assert len(cat) == 7 * nw
    np.testing.assert_array_equal(cat.as_array(), a.get_catalog().as_array())
    np.testing.assert_array_equal(b.chain, a.chain[3:])
    # ncm_fit_esmcmc_run: nothing to do once the catalog holds n_total ensembles
    b.resume(p, 5)
    assert b.chain is None and len(MSetCatalog.read(p)) == 7 * nw


def _apes_small(ktype, seed, lb, ub):
    from numcosmo_amd.apes import APES, InterpolationKernel
    g = np.random.default_rng(17)
    tgt = synth.Gaussian(2, g)
    s = APES(nwalkers=120, ndim=2, log_prob_batch=lambda th: -0.5 * tgt.m2lnp(th),
             interpolation_kernel=InterpolationKernel(ktype), over_smooth=0.2, local_fraction=0.2,
             random_walk_prob=0.25, bounds=(lb, ub), seed=seed)
    return tgt, s


@pytest.mark.parametrize("ki", [25, 73])
def test_apes_resume_inside_an_ensemble_matches_oracle(oracle, ki):
    """A catalog that ends ki walkers into an ensemble: the first resumed
    step moves walkers [ki, nwalkers) only, from each walker's latest row
    (ncm_fit_esmcmc.c:1855-1880, 2244-2268).  Every move of the two resumed
    steps is replayed on the oracle as in test_apes_moves_match_oracle."""
    import math
    from numcosmo_amd.catalog import MSetCatalog
    g = np.random.default_rng(17)
    tgt = synth.Gaussian(2, g)
    sig = np.sqrt(np.diag(tgt.cov))
    lb, ub = tgt.mean - 1.2 * sig, tgt.mean + 1.8 * sig
    _say(f"[apes] resume inside an ensemble, ki={ki}, nw=120, replayed on the oracle")
    tgt, s1 = _apes_small(0, 5, lb, ub)
    x0 = tgt.draw(g, 2400)
    x0 = x0[np.all((x0 >= lb) & (x0 <= ub), axis=1)][:120]
    s1.run_mcmc(x0, 2)
    nw, n2, dim = 120, 60, 2
    full = s1.get_catalog().as_array()
    cat = MSetCatalog([f"APES:theta_{i}" for i in range(dim)], nchains=nw)
    cat.add_rows(full[:2 * nw + ki])
    # each walker's latest row: ensemble 2 below ki, ensemble 1 from ki on
    theta = np.vstack([full[2 * nw:2 * nw + ki, 1:], full[nw + ki:2 * nw, 1:]])
    m2 = np.concatenate([full[2 * nw:2 * nw + ki, 0], full[nw + ki:2 * nw, 0]])

    _, s2 = _apes_small(0, 9, lb, ub)
    s2.resume(cat, 4, trace=True)
    assert len(cat) == 4 * nw and s2.chain.shape == (2, nw, 1 + dim)
    np.testing.assert_array_equal(cat.as_array()[:2 * nw + ki], full[:2 * nw + ki])
    np.testing.assert_array_equal(cat.as_array()[2 * nw + ki:], s2.chain.reshape(-1, 1 + dim)[ki:])

    halves = ([(0, ki, n2)] if ki < n2 else []) + [(1, max(ki, n2), nw), (0, 0, n2), (1, n2, nw)]
    T = s2.trace
    assert T.shape[0] == sum(k1 - k0 for _, k0, k1 in halves)
    pos, worst, decided = 0, 0.0, 0
    for h, (which, k0, k1) in enumerate(halves):
        other0 = n2 if which == 0 else 0
        R = T[pos:pos + k1 - k0]
        pos += k1 - k0
        th, ts = R[:, 9:9 + dim], R[:, 9 + dim:]
        np.testing.assert_array_equal(R[:, 0], which)
        np.testing.assert_array_equal(R[:, 1], np.arange(k0, k1))
        np.testing.assert_array_equal(th, theta[k0:k1])
        np.testing.assert_array_equal(R[:, 3], m2[k0:k1])
        np.testing.assert_array_equal(R[:, 2], 0.0)
        od = oracle.StatsDist(oracle.VKDE, oracle.ST, dim, 1.0)
        od.set_over_smooth(0.2)
        od.set_local_frac(0.2)
        od.set_shrink(0.01)
        od.add_obs_array(theta[other0:other0 + n2])
        od.prepare_interp(m2[other0:other0 + n2])
        std = oracle.apes_rw_std(od.get_full_cov())
        for j in range(k1 - k0):
            ps = oracle.apes_transition_m2lnp(od, th[j], ts[j], lb, ub, std, 0.25)
            pc = oracle.apes_transition_m2lnp(od, ts[j], th[j], lb, ub, std, 0.25)
            e = max(abs(R[j, 5] - ps) / max(abs(ps), 1.0), abs(R[j, 6] - pc) / max(abs(pc), 1.0))
            worst = max(worst, e)
            assert e <= 1e-7, (h, j, R[j, 5], ps, R[j, 6], pc)
            prob = min(math.exp(-0.5 * (R[j, 4] - R[j, 3] + pc - ps)), 1.0)
            if abs(R[j, 7] - prob) > 1e-4:
                assert R[j, 8] == float(R[j, 7] < prob), (h, j, R[j, 7], prob)
                decided += 1
        acc = R[:, 8] == 1.0
        theta[k0:k1][acc] = ts[acc]
        m2[k0:k1][acc] = R[acc, 4]
        if k1 == nw:
            step = 0 if h < len(halves) - 2 else 1
            np.testing.assert_array_equal(s2.chain[step, :, 1:], theta)
            np.testing.assert_array_equal(s2.chain[step, :, 0], m2)
    _say(f"   worst transition rel err {worst:.2e}, {decided} of {T.shape[0]} decisions checked")
    assert decided >= 0.9 * T.shape[0]


def test_apes_resume_drops_a_corrupt_last_ensemble(tmp_path):
    """ncm_fit_esmcmc.c:1883-1909: a last ensemble whose -2 ln L no longer
    matches the likelihood (relative difference > 1e-3) is removed, the file
    backed up, and the run resumes from the ensemble before it;
    skip_check keeps it."""
    import shutil
    from numcosmo_amd.catalog import MSetCatalog
    _say("[apes] resume drops a corrupt last ensemble")
    _, s = _run(2, 0, 3, seed=51)
    nw = s.nwalkers
    p = str(tmp_path / "c.fits")
    s.write_catalog(p)
    cat = MSetCatalog.read(p)
    rows = cat.as_array().copy()
    rows[-5, 0] *= 1.01
    bad = MSetCatalog(cat.param_names, nchains=nw)
    bad.add_rows(rows)
    bad.write(p)
    q = str(tmp_path / "skip.fits")
    shutil.copy(p, q)

    s.resume(p, 6, trace=True)
    np.testing.assert_array_equal(MSetCatalog.read(p + ".0.bak").as_array(), rows)
    out = MSetCatalog.read(p).as_array()
    assert out.shape[0] == 6 * nw
    np.testing.assert_array_equal(out[:3 * nw], rows[:3 * nw])
    # the first resumed half-step starts from ensemble 2
    np.testing.assert_array_equal(s.trace[:nw // 2, 9:11], rows[2 * nw:2 * nw + nw // 2, 1:])
    assert s.chain.shape[0] == 3

    s.resume(q, 5, skip_check=True)
    out = MSetCatalog.read(q).as_array()
    assert out.shape[0] == 5 * nw
    np.testing.assert_array_equal(out[:4 * nw], rows)
    assert not (tmp_path / "skip.fits.0.bak").exists()


def test_apes_resume_completes_the_first_ensemble():
    """A catalog with fewer rows than walkers: the missing walkers come from
    initial_sample, their rows are added, then the run goes on."""
    from numcosmo_amd.catalog import MSetCatalog
    _say("[apes] resume from an incomplete first ensemble")
    g = np.random.default_rng(61)
    tgt = synth.Gaussian(2, g)
    from numcosmo_amd.apes import APES
    s = APES(nwalkers=200, ndim=2, log_prob_batch=lambda th: -0.5 * tgt.m2lnp(th), seed=61)
    x0 = tgt.draw(g, 200)
    cat = MSetCatalog(["a", "b"], nchains=200)
    cat.add_rows(np.hstack([tgt.m2lnp(x0[:70])[:, None], x0[:70]]))
    with pytest.raises(ValueError, match="initial_sample"):
        s.resume(cat, 3)
    s.resume(cat, 3, initial_sample=x0)
    out = cat.as_array()
    assert out.shape[0] == 3 * 200
    np.testing.assert_array_equal(out[:200, 1:], x0)
    np.testing.assert_allclose(out[:200, 0], tgt.m2lnp(x0), rtol=1e-14)
    np.testing.assert_array_equal(out[200:], s.chain.reshape(-1, 3))


def test_apes_step_at_full_size(oracle):
    """One ensemble step at the benchmark's APES size: d=16, 2 x 3200
    walkers on the seeded 4-component mixture, Cauchy kernel, over_smooth 1,
    mixed-precision transition densities.  The oracle's active-set fit at
    n=3200 takes CPU minutes, so each half-step is pinned as the n=3200 fit
    in test_gpu_configs is:
    - the other half's VKDE, refitted through the Python API with the
      walker's settings: its weights satisfy the KKT conditions on the
      oracle's interpolation matrix;
    - the oracle, holding those weights and its own prepared factors,
      recomputes both transition densities (the VKDE mixed with the 2%
      random walk, ncm_fit_esmcmc_walker_apes.c:819-859) of 128 sampled
      walkers to the mixed-precision bar;
    - every one of the 6400 decisions follows the acceptance rule from the
      recorded values, and the chain is the state those decisions build."""
    import math
    import numcosmo_amd as na
    from numcosmo_amd.apes import APES, InterpolationKernel
    d, nw, n2, rw = 16, 6400, 3200, 0.02
    _say(f"[apes] one step at d={d}, {nw} walkers, mixture target, mixed precision")
    g = synth.rng(3)
    tgt = synth.Mixture(d, g)
    x0 = tgt.draw(g, nw)
    s = APES(nwalkers=nw, ndim=d, log_prob_batch=lambda th: -0.5 * tgt.m2lnp(th),
             interpolation_kernel=InterpolationKernel.CAUCHY, over_smooth=1.0, precision=1, seed=7)
    s.run_mcmc(x0, 1, trace=True)
    T = s.trace
    assert T.shape == (nw, 2 * d + 9)
    lb, ub = np.full(d, -np.inf), np.full(d, np.inf)
    theta, m2 = x0.copy(), s.m2lnL(x0)
    pick = np.sort(np.random.default_rng(3).choice(n2, size=128, replace=False))
    worst = 0.0
    for which in (0, 1):
        k0, other0 = (0, n2) if which == 0 else (n2, 0)
        R = T[which * n2:(which + 1) * n2]
        th, ts = R[:, 9:9 + d], R[:, 9 + d:]
        np.testing.assert_array_equal(R[:, 0], which)
        np.testing.assert_array_equal(R[:, 1], np.arange(k0, k0 + n2))
        np.testing.assert_array_equal(th, theta[k0:k0 + n2])
        np.testing.assert_array_equal(R[:, 3], m2[k0:k0 + n2])
        np.testing.assert_allclose(R[:, 4], s.m2lnL(ts), rtol=1e-14)

        Xo, mo = theta[other0:other0 + n2], m2[other0:other0 + n2]
        sd = na.StatsDistVKDE.new(na.StatsDistKernelST.new(d, 1.0))
        sd.set_over_smooth(1.0)
        sd.set_shrink(0.01)
        sd.add_obs_array(Xo)
        sd.prepare_interp(mo)
        w = sd.peek_weights()
        od = oracle.StatsDist(oracle.VKDE, oracle.ST, d, 1.0)
        od.set_over_smooth(1.0)
        od.set_shrink(0.01)
        od.add_obs_array(Xo)
        od.prepare()
        assert sd.get_n_kernels() == n2
        A = od.compute_IM() / np.exp(-0.5 * (mo - mo.min()))[:, None]
        v = w - 0.01 / n2
        Av = A @ v
        x = v * (Av.sum() / (Av @ Av))
        grad = A.T @ (1.0 - A @ x)
        scale = np.abs(A.T @ np.ones(n2)).max()
        act = x > 0
        kkt = max(np.abs(grad[act]).max(initial=0.0), grad[~act].max(initial=0.0)) / scale
        od.set_weights(w)
        std = oracle.apes_rw_std(od.get_full_cov())
        for j in pick:
            ps = oracle.apes_transition_m2lnp(od, th[j], ts[j], lb, ub, std, rw)
            pc = oracle.apes_transition_m2lnp(od, ts[j], th[j], lb, ub, std, rw)
            e = max(abs(R[j, 5] - ps) / max(abs(ps), 1.0), abs(R[j, 6] - pc) / max(abs(pc), 1.0))
            worst = max(worst, e)
            assert e <= 3e-6, (which, j, R[j, 5], ps, R[j, 6], pc)
        _say(f"   half {which}: KKT residual {kkt:.2e}, worst transition rel err so far {worst:.2e}, "
             f"accepted {int(R[:, 8].sum())} of {n2}")
        assert kkt < 1e-8
        prob = np.minimum(np.exp(-0.5 * ((R[:, 4] - R[:, 3]) + (R[:, 6] - R[:, 5]))), 1.0)
        clear = np.abs(R[:, 7] - prob) > 1e-12
        np.testing.assert_array_equal(R[clear, 8], (R[clear, 7] < prob[clear]).astype(float))
        acc = R[:, 8] == 1.0
        theta[k0:k0 + n2][acc] = ts[acc]
        m2[k0:k0 + n2][acc] = R[acc, 4]
    np.testing.assert_array_equal(s.chain[0, :, 1:], theta)
    np.testing.assert_array_equal(s.chain[0, :, 0], m2)
    assert 0 < T[:, 8].sum() < nw
