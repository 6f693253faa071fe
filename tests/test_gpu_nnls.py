"""The NNLS weight fit's unconstrained-solve rungs on the device against the
oracle (ncm_nnls.c:574-640, 655-706, 767-871).

- Each method of ncm_nnls_set_umethod -- NORMAL (Cholesky), NORMAL_LU
  (symmetric solve) and QR (least squares on A_P) -- on the same
  prepare_interp as the oracle with that method: weights to rel L1 1e-7 and
  the residual norm to 1e-9, with every solve ending on the rung the method
  names.
- Duplicated sample points: identical columns (and rows) in the
  interpolation matrix, so the normal matrix is exactly singular.  The fit
  must finish (no numerical error), reach the oracle's residual norm and
  density, and give each group of duplicates the oracle's total weight
  (the split inside a group is not unique).
- Refinement (off by default, as in the reference): on an ill-conditioned
  fit, the steps asked for are taken and the refined fit's residual norm is
  no larger than the unrefined one's.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _target(seed, d, n, m=400):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((d, d))
    S = A @ A.T + d * np.eye(d)
    L = np.linalg.cholesky(S)
    X = rng.standard_normal((n, d)) @ L.T
    Si = np.linalg.inv(S)
    c = np.log(np.linalg.det(2 * np.pi * S))
    m2f = lambda Y: np.einsum("ij,jk,ik->i", Y, Si, Y) + c
    Xe = rng.standard_normal((m, d)) @ L.T
    return X, m2f, Xe


def _pair(oracle, kind, kernel, d, nu, umethod, local_frac=0.1):
    import numcosmo_amd as na
    sdk = na.StatsDistKernelGauss.new(d) if kernel == "GAUSS" else na.StatsDistKernelST.new(d, nu)
    sd = (na.StatsDistKDE if kind == "KDE" else na.StatsDistVKDE).new(sdk)
    od = oracle.StatsDist(getattr(oracle, kind), getattr(oracle, kernel), d, nu)
    if kind == "VKDE":
        sd.set_local_frac(local_frac)
        od.set_local_frac(local_frac)
    sd.set_nnls_umethod(umethod)
    od.set_nnls_umethod(umethod)
    return sd, od


def _rel(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1.0)


CASES = [("VKDE", "ST", 3.0, 3, 300, 11), ("KDE", "GAUSS", 0.0, 3, 300, 12)]


@pytest.mark.parametrize("umethod", [0, 1, 2], ids=["NORMAL", "NORMAL_LU", "QR"])
@pytest.mark.parametrize("kind,kernel,nu,d,n,seed", CASES)
def test_umethod_matches_oracle(oracle, umethod, kind, kernel, nu, d, n, seed):
    _say(f"[nnls] {kind} {kernel} d={d} n={n} umethod={umethod}")
    X, m2f, Xe = _target(seed, d, n)
    sd, od = _pair(oracle, kind, kernel, d, nu, umethod)
    sd.add_obs_array(X)
    od.add_obs_array(X)
    oracle.nnls_rung_counts(reset=True)
    sd.prepare_interp(m2f(X))
    od.prepare_interp(m2f(X))
    ro = oracle.nnls_rung_counts(reset=True)
    rd = sd.get_nnls_rungs()
    w, wo = sd.peek_weights(), od.get_weights()
    rel_l1 = np.abs(w - wo).sum() / np.abs(wo).sum()
    _say(f"   rungs device {rd} oracle {ro}; weights rel L1 {rel_l1:.2e}; "
         f"rnorm^2 {sd.get_rnorm():.12e} vs {od.get_rnorm():.12e}")
    assert rd[umethod] > 0 and sum(rd[:3]) == rd[umethod]
    assert ro[umethod] > 0 and sum(ro) == ro[umethod]
    assert rel_l1 <= 1e-7
    assert sd.get_rnorm() == pytest.approx(od.get_rnorm(), rel=1e-9)
    assert np.max(_rel(sd.eval_m2lnp_batch(Xe), od.eval_m2lnp_batch(Xe))) <= 1e-9


@pytest.mark.parametrize("umethod", [0, 1, 2], ids=["NORMAL", "NORMAL_LU", "QR"])
@pytest.mark.parametrize("kind,kernel,nu", [("VKDE", "ST", 3.0), ("KDE", "GAUSS", 0.0)])
def test_duplicate_points(oracle, umethod, kind, kernel, nu):
    d, n, ndup = 3, 240, 40
    _say(f"[nnls] {kind} {kernel} duplicated points: {ndup} of {n} repeated, umethod={umethod}")
    X, m2f, Xe = _target(21, d, n)
    Xd = np.concatenate([X, X[:ndup]])
    sd, od = _pair(oracle, kind, kernel, d, nu, umethod)
    sd.add_obs_array(Xd)
    od.add_obs_array(Xd)
    oracle.nnls_rung_counts(reset=True)
    sd.prepare_interp(m2f(Xd))
    od.prepare_interp(m2f(Xd))
    ro = oracle.nnls_rung_counts(reset=True)
    rd = sd.get_nnls_rungs()
    w, wo = sd.peek_weights(), od.get_weights()
    # total weight of each group of identical points
    gw = np.concatenate([w[:ndup] + w[n:], w[ndup:n]])
    gwo = np.concatenate([wo[:ndup] + wo[n:], wo[ndup:n]])
    rel_l1 = np.abs(gw - gwo).sum() / np.abs(gwo).sum()
    e = np.max(_rel(sd.eval_m2lnp_batch(Xe), od.eval_m2lnp_batch(Xe)))
    _say(f"   rungs device {rd} oracle {ro}; group weights rel L1 {rel_l1:.2e}; eval {e:.2e}; "
         f"rnorm^2 {sd.get_rnorm():.12e} vs {od.get_rnorm():.12e}")
    assert sd.get_rnorm() == pytest.approx(od.get_rnorm(), rel=1e-7)
    assert rel_l1 <= 1e-6
    assert e <= 1e-7


def test_refinement_on_an_ill_conditioned_fit(oracle):
    """A smooth d=2 Gaussian target with a wide Gauss KDE: the interpolation
    matrix is nearly singular, the Cholesky pivots spread by far more than
    1e10, and the refinement steps asked for are taken; they never raise the
    residual.  Without a request there is none (the reference's solver)."""
    import numcosmo_amd as na
    X, m2f, Xe = _target(5, 2, 300)
    assert na.StatsDistKDE.new(na.StatsDistKernelGauss.new(2)).get_nnls_refine() == 0
    res = {}
    for steps in (0, 2):
        sd = na.StatsDistKDE.new(na.StatsDistKernelGauss.new(2))
        sd.set_over_smooth(3.0)
        sd.set_nnls_refine(steps)
        assert sd.get_nnls_refine() == steps
        sd.add_obs_array(X)
        sd.prepare_interp(m2f(X))
        res[steps] = (sd.get_rnorm(), sd.get_nnls_rungs())
    od = oracle.StatsDist(oracle.KDE, oracle.GAUSS, 2, 3.0)
    od.set_over_smooth(3.0)
    od.add_obs_array(X)
    od.prepare_interp(m2f(X))
    _say(f"   rnorm^2 unrefined {res[0][0]:.12e} {res[0][1]}, refined {res[2][0]:.12e} {res[2][1]}, "
         f"oracle {od.get_rnorm():.12e}")
    assert res[0][1][3] == 0
    assert res[2][1][3] > 0
    assert res[2][0] <= res[0][0] * (1 + 1e-12)
