"""APES ensemble sampler on the GPU kernel-density engine.

Mirrors numcosmo_py's ``APES`` (numcosmo_py/sampling/apes.py:39-172), which
drives NcmFitESMCMC with NcmFitESMCMCWalkerAPES: ``APES(nwalkers=...,
ndim=..., log_prob=...)`` then ``run_mcmc(initial_sample, niter)``.  The
walker itself is native (nckde_apes_* in include/nckde.h).

``log_prob(theta, args)`` is called per point as in the reference; passing
``log_prob_batch(thetas, args) -> array`` instead evaluates a whole
half-ensemble of proposals per call.
"""
from __future__ import annotations

import ctypes as C
import enum
from typing import Callable, Optional, Tuple

import numpy as np

from ._lib import M2LNL_BATCH_FN, check, lib


class InterpolationMethod(enum.IntEnum):
    """NcmFitESMCMCWalkerAPESMethod (ncm_fit_esmcmc_walker_apes.h:51-57)."""
    KDE = 0
    VKDE = 1


class InterpolationKernel(enum.IntEnum):
    """NcmFitESMCMCWalkerAPESKType (ncm_fit_esmcmc_walker_apes.h:68-75)."""
    CAUCHY = 0
    ST3 = 1
    GAUSS = 2


class APES:
    """Ensemble sampler: 2 x (nwalkers/2) walkers, each half moved with a VKDE
    of the other half."""

    def __init__(self, *, nwalkers: int, ndim: int,
                 log_prob: Optional[Callable] = None, log_prob_batch: Optional[Callable] = None,
                 args: Tuple = (), use_interpolation: bool = True,
                 interpolation_method: InterpolationMethod = InterpolationMethod.VKDE,
                 interpolation_kernel: InterpolationKernel = InterpolationKernel.CAUCHY,
                 over_smooth: float = 0.2, local_fraction: Optional[float] = None,
                 shrink: float = 0.01, random_walk_prob: float = 0.02,
                 bounds: Optional[Tuple[np.ndarray, np.ndarray]] = None,
                 precision: int = 0, seed: int = 0):
        if (log_prob is None) == (log_prob_batch is None):
            raise ValueError("exactly one of log_prob and log_prob_batch must be given")
        self.nwalkers, self.ndim, self.args = int(nwalkers), int(ndim), args
        self.log_prob, self.log_prob_batch = log_prob, log_prob_batch
        h = C.c_void_p()
        check(lib().nckde_apes_create(C.byref(h), self.nwalkers, self.ndim, int(interpolation_method),
                                      int(interpolation_kernel), int(seed)))
        self._h = h
        L = lib()
        check(L.nckde_apes_set_over_smooth(h, float(over_smooth)))
        check(L.nckde_apes_set_shrink(h, float(shrink)))
        if local_fraction is not None:
            check(L.nckde_apes_set_local_frac(h, float(local_fraction)))
        check(L.nckde_apes_set_random_walk_prob(h, float(random_walk_prob)))
        check(L.nckde_apes_set_use_interp(h, int(bool(use_interpolation))))
        check(L.nckde_apes_set_precision(h, int(precision)))
        if bounds is not None:
            lb = np.ascontiguousarray(bounds[0], dtype=np.float64)
            ub = np.ascontiguousarray(bounds[1], dtype=np.float64)
            check(L.nckde_apes_set_bounds(h, lb, ub))
        self._error = None
        self._cb = M2LNL_BATCH_FN(self._callback)
        self.chain = None
        self.trace = None
        self._catalog = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().nckde_apes_destroy(h)
            self._h = None

    def m2lnL(self, thetas: np.ndarray) -> np.ndarray:
        if self.log_prob_batch is not None:
            return -2.0 * np.asarray(self.log_prob_batch(thetas, *self.args), dtype=np.float64)
        return np.array([-2.0 * self.log_prob(t, *self.args) for t in thetas])

    def _callback(self, theta_p, m, d, out_p, _user):
        try:
            th = np.ctypeslib.as_array(theta_p, shape=(m, d)).copy()
            out = np.ctypeslib.as_array(out_p, shape=(m,))
            out[:] = self.m2lnL(th)
            return 0
        except Exception as e:  # reported after the native call returns
            self._error = e
            return 1

    def run_mcmc(self, initial_sample: np.ndarray, niter: int, keep_chain: bool = True,
                 trace: bool = False) -> None:
        x0 = np.ascontiguousarray(initial_sample, dtype=np.float64)
        if x0.shape != (self.nwalkers, self.ndim):
            raise ValueError("Initial sample has wrong shape.")
        m2_0 = np.ascontiguousarray(self.m2lnL(x0))
        check(lib().nckde_apes_init(self._h, x0, m2_0))
        chain = np.empty((niter, self.nwalkers, 1 + self.ndim)) if keep_chain else None
        # optional record of every move: rows (half, walker, exploration,
        # m2lnL, m2lnL*, m2lnp*, m2lnp_cur, jump, accepted, theta, theta*)
        tbuf = np.empty((niter * self.nwalkers, 2 * self.ndim + 9)) if trace else None
        check(lib().nckde_apes_set_trace(self._h, tbuf.ctypes.data if trace else None, tbuf.shape[0] if trace else 0))
        st = lib().nckde_apes_run(self._h, int(niter), C.cast(self._cb, C.c_void_p), None,
                                  chain.ctypes.data if chain is not None else None)
        rows = C.c_size_t()
        check(lib().nckde_apes_get_trace_rows(self._h, C.byref(rows)))
        check(lib().nckde_apes_set_trace(self._h, None, 0))
        if self._error is not None:
            e, self._error = self._error, None
            raise e
        check(st)
        self.chain = chain
        self.trace = tbuf[:rows.value] if trace else None
        if keep_chain:
            # as numcosmo_py's APES.run_mcmc: the catalog starts with the
            # initial sample (numcosmo_py/sampling/apes.py:146-167)
            from .catalog import MSetCatalog

            if self._catalog is None:
                self._catalog = MSetCatalog([f"APES:theta_{i}" for i in range(self.ndim)], nchains=self.nwalkers)
                self._catalog.add_rows(np.hstack([m2_0[:, None], x0]))
            self._catalog.add_rows(chain.reshape(-1, 1 + self.ndim))

    def get_catalog(self):
        """numcosmo_py APES.get_catalog: every walker row so far (the initial
        sample first), as an NcmMSetCatalog-layout ``MSetCatalog``."""
        if self._catalog is None:
            raise ValueError("no catalog: run_mcmc with keep_chain=True first")
        return self._catalog

    def write_catalog(self, path: str, param_names=None, param_symbols=None) -> None:
        """``get_catalog()`` as an NcmMSetCatalog FITS file: one row per
        walker and step (NcmFit:m2lnL, then the parameters), NCHAINS =
        nwalkers, the initial sample in the first step."""
        cat = self.get_catalog()
        if param_names is not None:
            cat.param_names = list(param_names)
            cat.param_symbols = list(param_symbols) if param_symbols is not None else list(param_names)
        cat.write(path)

    def set_exploration(self, steps: int) -> None:
        """ncm_fit_esmcmc_walker_apes_set_exploration: the next ``steps``
        half-step setups count down an exploration phase in which the
        acceptance uses the likelihood ratio alone (ncm_fit_esmcmc_walker_apes.c:815-816,
        895-919); the countdown runs before each half's acceptance, so the
        first ``steps - 1`` half-steps explore."""
        check(lib().nckde_apes_set_exploration(self._h, int(steps)))

    def get_state(self):
        th = np.empty((self.nwalkers, self.ndim))
        m2 = np.empty(self.nwalkers)
        check(lib().nckde_apes_get_state(self._h, th, m2))
        return th, m2

    def get_stats(self):
        acc, prop = C.c_uint64(), C.c_uint64()
        secs = np.zeros(4)
        check(lib().nckde_apes_get_stats(self._h, C.byref(acc), C.byref(prop), secs))
        return {"accepted": acc.value, "proposed": prop.value,
                "acceptance": acc.value / max(prop.value, 1),
                "seconds": dict(zip(["prepare", "sample", "eval", "likelihood"], secs.tolist()))}
