"""Seeded synthetic inputs for the kernel-density workloads.

Every generator draws from ``numpy.random.Generator(Philox(key))`` with
``key = 20261202 + config_id`` unless a seed is given, so the GPU engine and
the CPU oracle always see identical inputs.

Workloads (sizes as used by the benchmark and the parity tests):

* ``gaussian``   N(0, Sigma), Sigma = Q diag(lambda) Q^T with Q Haar and
                 lambda_i = 10**U(-1, 1); m2lnp exact.  (KDE Gauss d=4, n=1000)
* ``banana``     x ~ N(0, I_d), x_2 += 0.5 (x_1**2 - 1); m2lnp exact (unit
                 Jacobian).  (VKDE Student-t nu=3, d=8, n=4000, m=1e5)
* ``mixture``    equal-weight 4-component Gaussian mixture, means U[-5,5]^d,
                 covariances as ``gaussian``; m2lnp by FP64 log-sum-exp.
                 (APES d=16, 2 x 3200 walkers)
* ``cond``       Gaussian with eigenvalues log-spaced on [1, cond] and a Haar
                 rotation.  (VKDE Student-t nu=5, d=64, n=50000)
"""
from __future__ import annotations

import numpy as np

BASE_KEY = 20261202
LN2PI = 1.8378770664093454835606594728112353


def rng(config_id: int = 0, seed: int | None = None) -> np.random.Generator:
    key = BASE_KEY + config_id if seed is None else seed
    return np.random.Generator(np.random.Philox(key=key))


def haar(g: np.random.Generator, d: int) -> np.ndarray:
    """Haar-distributed orthogonal matrix (QR of a Gaussian, sign-fixed)."""
    Z = g.standard_normal((d, d))
    Q, R = np.linalg.qr(Z)
    return Q * np.sign(np.diag(R))[None, :]


def spd(g: np.random.Generator, d: int, lo: float = -1.0, hi: float = 1.0) -> np.ndarray:
    Q = haar(g, d)
    lam = 10.0 ** g.uniform(lo, hi, size=d)
    S = (Q * lam[None, :]) @ Q.T
    return 0.5 * (S + S.T)


def gauss_m2lnp(X: np.ndarray, mean: np.ndarray, cov: np.ndarray) -> np.ndarray:
    L = np.linalg.cholesky(cov)
    Z = np.linalg.solve(L, (X - mean).T)
    lndet = 2.0 * np.sum(np.log(np.diag(L)))
    return np.sum(Z * Z, axis=0) + lndet + X.shape[1] * LN2PI


class Gaussian:
    def __init__(self, d: int, g: np.random.Generator, cov: np.ndarray | None = None):
        self.d = d
        self.mean = np.zeros(d)
        self.cov = spd(g, d) if cov is None else cov
        self.L = np.linalg.cholesky(self.cov)

    def draw(self, g: np.random.Generator, n: int) -> np.ndarray:
        return self.mean + g.standard_normal((n, self.d)) @ self.L.T

    def m2lnp(self, X: np.ndarray) -> np.ndarray:
        return gauss_m2lnp(X, self.mean, self.cov)


class Banana:
    def __init__(self, d: int, b: float = 0.5):
        self.d, self.b = d, b

    def draw(self, g: np.random.Generator, n: int) -> np.ndarray:
        X = g.standard_normal((n, self.d))
        X[:, 1] += self.b * (X[:, 0] ** 2 - 1.0)
        return X

    def m2lnp(self, X: np.ndarray) -> np.ndarray:
        u = X.copy()
        u[:, 1] -= self.b * (X[:, 0] ** 2 - 1.0)
        return np.sum(u * u, axis=1) + self.d * LN2PI


class Mixture:
    def __init__(self, d: int, g: np.random.Generator, ncomp: int = 4):
        self.d = d
        self.means = g.uniform(-5.0, 5.0, size=(ncomp, d))
        self.covs = [spd(g, d) for _ in range(ncomp)]
        self.Ls = [np.linalg.cholesky(c) for c in self.covs]
        self.w = np.full(ncomp, 1.0 / ncomp)

    def draw(self, g: np.random.Generator, n: int) -> np.ndarray:
        comp = g.choice(len(self.w), size=n, p=self.w)
        Z = g.standard_normal((n, self.d))
        X = np.empty((n, self.d))
        for c in range(len(self.w)):
            sel = comp == c
            X[sel] = self.means[c] + Z[sel] @ self.Ls[c].T
        return X

    def m2lnp(self, X: np.ndarray) -> np.ndarray:
        terms = np.stack([np.log(self.w[c]) - 0.5 * gauss_m2lnp(X, self.means[c], self.covs[c])
                          for c in range(len(self.w))])
        mx = terms.max(axis=0)
        return -2.0 * (mx + np.log(np.sum(np.exp(terms - mx), axis=0)))


def cond_gaussian(d: int, g: np.random.Generator, cond: float = 1.0e4) -> Gaussian:
    Q = haar(g, d)
    lam = np.logspace(0.0, np.log10(cond), d)
    S = (Q * lam[None, :]) @ Q.T
    return Gaussian(d, g, cov=0.5 * (S + S.T))


def workload(name: str, n: int, m: int, d: int, config_id: int, seed: int | None = None):
    """Return (target, X_sample, m2lnp_sample, X_eval)."""
    g = rng(config_id, seed)
    if name == "gaussian":
        tgt = Gaussian(d, g)
    elif name == "banana":
        tgt = Banana(d)
    elif name == "mixture":
        tgt = Mixture(d, g)
    elif name == "cond":
        tgt = cond_gaussian(d, g)
    else:
        raise ValueError(name)
    X = tgt.draw(g, n)
    Xe = tgt.draw(g, m)
    return tgt, X, tgt.m2lnp(X), Xe


class ServedStream:
    """Served request stream: Poisson arrivals over prepared sessions.

    Sessions have n log-uniform on [n_lo, n_hi] and d uniform on ``dims``;
    each targets a seeded correlated Gaussian (covariance as ``gaussian``)
    and is fitted with VKDE + prepare_interp on n draws.  Requests pick a
    session uniformly (so a request's n is log-uniform and its d uniform),
    with ops eval_m2lnp / sample / prepare_interp in proportions ``mix``;
    eval and sample requests carry m = m_factor * n points or draws, and a
    prepare_interp request refits its session on a fresh sample of n.
    Per-session point and draw pools are generated once and shared by the
    session's requests.
    """

    OPS = ("eval", "sample", "prepare_interp")

    def __init__(self, n_req=512, rate=20.0, n_sessions=48, n_lo=500, n_hi=20000, dims=(4, 8, 16, 32),
                 mix=(0.6, 0.3, 0.1), m_factor=10, seed=None, nu=3.0):
        g = rng(4, seed)
        self.nu = nu
        self.m_factor = m_factor
        self.sessions = []
        for s in range(n_sessions):
            n = int(round(np.exp(g.uniform(np.log(n_lo), np.log(n_hi)))))
            d = int(g.choice(dims))
            self.sessions.append({"n": n, "d": d, "seed": int(g.integers(1 << 62))})
        self.arrival = np.cumsum(g.exponential(1.0 / rate, size=n_req)) if rate > 0 else np.zeros(n_req)
        self.op = g.choice(3, size=n_req, p=np.asarray(mix) / np.sum(mix))
        self.session = g.integers(0, n_sessions, size=n_req)
        self._pools = {}

    def target(self, s):
        S = self.sessions[s]
        return Gaussian(S["d"], rng(seed=S["seed"]))

    def fit_sample(self, s, refit=0):
        """(X, m2lnp) of the session's fit number ``refit`` (0: initial)."""
        S = self.sessions[s]
        g = rng(seed=S["seed"] + 1 + refit)
        tgt = self.target(s)
        X = tgt.draw(g, S["n"])
        return X, tgt.m2lnp(X)

    def pool(self, s):
        """Evaluation points and (u, z, chi2) draws of session s."""
        if s not in self._pools:
            S = self.sessions[s]
            n, d = S["n"], S["d"]
            m = self.m_factor * n
            g = rng(seed=S["seed"] + 1_000_003)
            tgt = self.target(s)
            Xe = tgt.draw(g, m)
            u = g.random(m)
            z = g.standard_normal((m, d))
            c2 = g.chisquare(self.nu, size=m)
            self._pools[s] = (Xe, u, z, c2)
        return self._pools[s]
