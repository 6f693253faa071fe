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
import os
from typing import Callable, Optional, Tuple

import numpy as np

from ._lib import M2LNL_BATCH_FN, check, lib
from .catalog import MSetCatalog

# column of -2 ln L in an ESMCMC catalog (NCM_FIT_ESMCMC_M2LNL_ID)
M2LNL_ID = 0


class InterpolationMethod(enum.IntEnum):
    """NcmFitESMCMCWalkerAPESMethod (ncm_fit_esmcmc_walker_apes.h:51-57)."""
    KDE = 0
    VKDE = 1


class InterpolationKernel(enum.IntEnum):
    """NcmFitESMCMCWalkerAPESKType (ncm_fit_esmcmc_walker_apes.h:68-75)."""
    CAUCHY = 0
    ST3 = 1
    GAUSS = 2


def catalog_state(cat: MSetCatalog, nwalkers: int) -> Tuple[np.ndarray, np.ndarray]:
    """Walker positions and -2 ln L at the end of a catalog, as
    ncm_fit_esmcmc_start_run reads them (ncm_fit_esmcmc.c:1826-1880).

    With c = cur_id + 1 samples recorded, t, ki = divmod(c, nwalkers):
    walker k < ki is at sample id t nwalkers + k and walker k >= ki at
    (t - 1) nwalkers + k.  While the first ensemble is incomplete
    (c < nwalkers) only walkers 0..c-1 are known and fewer rows come back;
    a catalog trimmed past id 0 (first_id > 0) is then an error."""
    cur = cat.get_cur_id()
    rows = cat.as_array()
    nadd = len(cat.add_names)
    if cur + 1 < nwalkers:
        if cat.get_first_id() > 0:
            raise ValueError("cannot use catalogs with first_id > 0 without a complete first ensemble.")
        return rows[:cur + 1, nadd:].copy(), rows[:cur + 1, M2LNL_ID].copy()
    t, ki = divmod(cur + 1, nwalkers)
    ids = [cat.get_row_from_time(nwalkers * t + k) for k in range(ki)]
    ids += [cat.get_row_from_time(nwalkers * (t - 1) + k) for k in range(ki, nwalkers)]
    return rows[ids, nadd:].copy(), rows[ids, M2LNL_ID].copy()


def validate_rows(cat: MSetCatalog, m2lnL: Callable[[np.ndarray], np.ndarray], start: int, end: int) -> bool:
    """ncm_fit_esmcmc_validate over rows [start, end)
    (ncm_fit_esmcmc.c:2526-2560, 2676-2712): recompute -2 ln L at each row's
    parameters; the rows fail when |(row - recomputed) / row| > 1e-3 for any
    of them (a NaN difference does not fail, as in the reference)."""
    rows = cat.as_array()[start:end]
    nadd = len(cat.add_names)
    m = np.asarray(m2lnL(np.ascontiguousarray(rows[:, nadd:])), dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        diff = np.abs((rows[:, M2LNL_ID] - m) / rows[:, M2LNL_ID])
    return not bool(np.any(diff > 1.0e-3))


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

    def _run_native(self, k_first: int, nsteps: int, keep_chain: bool, trace: bool):
        """nsteps ensemble steps through nckde_apes_run_from; the first moves
        walkers [k_first, nwalkers).  Returns (chain, trace rows)."""
        chain = np.empty((nsteps, self.nwalkers, 1 + self.ndim)) if keep_chain else None
        # optional record of every move: rows (half, walker, exploration,
        # m2lnL, m2lnL*, m2lnp*, m2lnp_cur, jump, accepted, theta, theta*)
        tbuf = np.empty((nsteps * self.nwalkers, 2 * self.ndim + 9)) if trace else None
        check(lib().nckde_apes_set_trace(self._h, tbuf.ctypes.data if trace else None, tbuf.shape[0] if trace else 0))
        st = lib().nckde_apes_run_from(self._h, int(k_first), int(nsteps), C.cast(self._cb, C.c_void_p), None,
                                       chain.ctypes.data if chain is not None else None)
        rows = C.c_size_t()
        check(lib().nckde_apes_get_trace_rows(self._h, C.byref(rows)))
        check(lib().nckde_apes_set_trace(self._h, None, 0))
        if self._error is not None:
            e, self._error = self._error, None
            raise e
        check(st)
        return chain, (tbuf[:rows.value] if trace else None)

    def run_mcmc(self, initial_sample: np.ndarray, niter: int, keep_chain: bool = True,
                 trace: bool = False) -> None:
        x0 = np.ascontiguousarray(initial_sample, dtype=np.float64)
        if x0.shape != (self.nwalkers, self.ndim):
            raise ValueError("Initial sample has wrong shape.")
        m2_0 = np.ascontiguousarray(self.m2lnL(x0))
        check(lib().nckde_apes_init(self._h, x0, m2_0))
        chain, self.trace = self._run_native(0, niter, keep_chain, trace)
        self.chain = chain
        if keep_chain:
            # as numcosmo_py's APES.run_mcmc: the catalog starts with the
            # initial sample (numcosmo_py/sampling/apes.py:146-167)
            if self._catalog is None:
                self._catalog = MSetCatalog([f"APES:theta_{i}" for i in range(self.ndim)], nchains=self.nwalkers)
                self._catalog.add_rows(np.hstack([m2_0[:, None], x0]))
            self._catalog.add_rows(chain.reshape(-1, 1 + self.ndim))

    def resume(self, catalog, n_total: int, initial_sample: Optional[np.ndarray] = None,
               skip_check: bool = False, trace: bool = False) -> None:
        """Continue a run recorded in a catalog until it holds ``n_total``
        ensembles (sample ids up to n_total x nwalkers - 1, counted from id 0
        as ncm_fit_esmcmc_run (ncm_fit_esmcmc.c:2012-2034) counts them).

        ``catalog`` is an ``MSetCatalog``, which is extended in place, or the
        path of one in FITS, which is written back with the new rows at the
        end, as NcmMSetCatalog keeps its file in sync.  The walkers start where ncm_fit_esmcmc_start_run
        (ncm_fit_esmcmc.c:1808-1910) puts them: each at its latest row, so a
        catalog that ends inside an ensemble resumes with that ensemble's
        remaining walkers (``catalog_state``).  Unless ``skip_check``, the
        last ensemble's -2 ln L are recomputed; a relative difference above
        1e-3 in any row removes that ensemble (``remove_last_ensemble``,
        with the file backed up as ``<path>.<k>.bak``) and starts over.  A
        catalog that holds less than one whole ensemble is completed from
        ``initial_sample`` (the reference draws those walkers from its
        initial-point sampler, which this sampler does not have).

        The walker's random streams continue from this sampler's state (its
        seed and the draws made so far), not from a state stored in the
        catalog: the reference keeps its GSL generator's state in the file
        (ncm_mset_catalog.c:1106-1150)."""
        path = None
        if isinstance(catalog, (str, os.PathLike)):
            path = os.fspath(catalog)
            cat = MSetCatalog.read(path)
        else:
            cat = catalog
        if cat.nchains != self.nwalkers or len(cat.param_names) != self.ndim or len(cat.add_names) != 1:
            raise ValueError(f"catalog of {cat.nchains} chains, {len(cat.param_names)} parameters and "
                             f"{len(cat.add_names)} additional values; the sampler needs {self.nwalkers}, "
                             f"{self.ndim} and 1 (-2 ln L)")
        nw = self.nwalkers
        while True:
            theta, m2 = catalog_state(cat, nw)
            read_from_cat = theta.shape[0] > 0
            if theta.shape[0] < nw:
                if initial_sample is None:
                    raise ValueError(f"the catalog holds {theta.shape[0]} of the first ensemble's {nw} walkers; "
                                     "pass initial_sample for the rest")
                x = np.ascontiguousarray(initial_sample, dtype=np.float64)
                if x.shape != (nw, self.ndim):
                    raise ValueError("Initial sample has wrong shape.")
                x = x[theta.shape[0]:]
                mx = np.asarray(self.m2lnL(x), dtype=np.float64)
                if not np.all(np.isfinite(mx)):
                    raise ValueError("non-finite -2 ln L at an initial point")
                cat.add_rows(np.hstack([mx[:, None], x]))
                theta, m2 = np.vstack([theta, x]), np.concatenate([m2, mx])
            if read_from_cat and not skip_check:
                n = len(cat)
                if not validate_rows(cat, self.m2lnL, n - nw if n > nw else 0, n):
                    cat.remove_last_ensemble(path)
                    continue
            break
        ti, ki = divmod(cat.get_cur_id() + 1, nw)
        check(lib().nckde_apes_init(self._h, np.ascontiguousarray(theta), np.ascontiguousarray(m2)))
        self._catalog = cat
        self.chain, self.trace = None, None
        if n_total > ti:
            chain, self.trace = self._run_native(ki, n_total - ti, True, trace)
            self.chain = chain
            cat.add_rows(chain.reshape(-1, 1 + self.ndim)[ki:])
        if path is not None:
            cat.write(path)

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
