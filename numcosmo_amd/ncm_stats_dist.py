"""Python mirror of NumCosmo's NcmStatsDist API over the nckde C-ABI.

Names, argument meaning and errors follow the GObject-Introspection binding
of the reference (numcosmo_py exposes them as ``Ncm.StatsDistVKDE.new(...)``,
``sd.add_obs(v)``, ``sd.prepare_interp(m2lnp)``, ``sd.eval_m2lnp(v)`` ...):

* ncm_stats_dist.h:87-141            NcmStatsDist methods
* ncm_stats_dist_kde.h:74-86         NcmStatsDistKDE
* ncm_stats_dist_vkde.h:53-62        NcmStatsDistVKDE
* ncm_stats_dist_kernel_{gauss,st}.h NcmStatsDistKernelGauss / ST

Vectors are numpy arrays (the reference's NcmVector), matrices are 2-D
numpy arrays (NcmMatrix, row-major).  Batch extensions ``eval_m2lnp_batch``,
``eval_batch``, ``kernel_choose_batch`` and ``sample_batch`` submit many
points in one call; the single-point forms route to the same device engine.

Sampling draws its randomness from an ``RNG`` in the reference's call order
(one uniform, d standard normals, then one chi-squared for Student-t); the
stream is numpy's, not GSL's mt19937, so seeded draws differ from the
reference's while the transform of given draws is identical.
"""
from __future__ import annotations

import ctypes as C
import enum

import numpy as np

from ._lib import NckdeError, check, lib

__all__ = [
    "StatsDistCV", "StatsDistKDECovType", "Precision", "NNLSUMethod", "RNG", "NckdeError",
    "StatsDistKernelGauss", "StatsDistKernelST", "StatsDistKDE", "StatsDistVKDE", "deserialize",
    "eval_m2lnp_multi",
]


def eval_m2lnp_multi(sds, Xs, precision=1):
    """eval_m2lnp of several objects at once (nckde_eval_m2lnp_multi): mixed
    evals of objects with the same dimension class and kernel type share one
    launch per pass; results equal separate evals bit for bit."""
    n = len(sds)
    Xs = [np.ascontiguousarray(np.atleast_2d(x), dtype=np.float64) for x in Xs]
    outs = [np.empty(x.shape[0]) for x in Xs]
    hs = (C.c_void_p * n)(*[sd._h.value for sd in sds])
    xp = (C.c_void_p * n)(*[x.ctypes.data for x in Xs])
    mp = (C.c_size_t * n)(*[x.shape[0] for x in Xs])
    op = (C.c_void_p * n)(*[o.ctypes.data for o in outs])
    check(lib().nckde_eval_m2lnp_multi(hs, xp, mp, None, op, n, int(precision)))
    return outs


class StatsDistCV(enum.IntEnum):
    """NcmStatsDistCV (ncm_stats_dist.h:77-85)."""
    NONE = 0
    SPLIT = 1
    SPLIT_NOFIT = 2
    LOO = 3


class StatsDistKDECovType(enum.IntEnum):
    """NcmStatsDistKDECovType (ncm_stats_dist_kde.h:64-72)."""
    SAMPLE = 0
    FIXED = 1
    ROBUST_DIAG = 2
    ROBUST = 3


class Precision(enum.IntEnum):
    FP64 = 0
    MIXED = 1
    FP32 = 2


class NNLSUMethod(enum.IntEnum):
    """NcmNNLSUMethod (ncm_nnls.h:51-60) of the weight fit's unconstrained
    solve; DGELSD and GSL are not implemented."""
    NORMAL = 0
    NORMAL_LU = 1
    QR = 2


class RNG:
    """Stand-in for NcmRNG: uniform, standard normal and chi-squared draws."""

    def __init__(self, seed: int | None = None):
        self.g = np.random.default_rng(seed)

    def uniform_gen(self, a=0.0, b=1.0):
        return self.g.uniform(a, b)

    def ugaussian_gen(self):
        return self.g.standard_normal()

    def chisq_gen(self, nu):
        return self.g.chisquare(nu)


class _Kernel:
    kind = 0

    def __init__(self, dim: int, nu: float = 0.0):
        self.dim = int(dim)
        self.nu = float(nu)

    def get_dim(self):
        return self.dim


class StatsDistKernelGauss(_Kernel):
    """NcmStatsDistKernelGauss (ncm_stats_dist_kernel_gauss.h:41)."""
    kind = 0

    @classmethod
    def new(cls, dim: int):
        return cls(dim)


class StatsDistKernelST(_Kernel):
    """NcmStatsDistKernelST (ncm_stats_dist_kernel_st.h:41); nu >= 1."""
    kind = 1

    @classmethod
    def new(cls, dim: int, nu: float):
        return cls(dim, nu)

    def get_nu(self):
        return self.nu


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class _StatsDist:
    _kind = 0

    def __init__(self, sdk: _Kernel, cv_type: StatsDistCV = StatsDistCV.NONE):
        h = C.c_void_p()
        check(lib().nckde_create(C.byref(h), self._kind, sdk.kind, sdk.nu, sdk.dim, int(cv_type)))
        self._h = h
        self.kernel = sdk
        self.d = sdk.dim
        self._over_smooth = 1.0
        self.precision = Precision.FP64

    @classmethod
    def new(cls, sdk: _Kernel, cv_type: StatsDistCV = StatsDistCV.NONE):
        return cls(sdk, cv_type)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().nckde_destroy(h)
            self._h = None

    # ---- properties ----------------------------------------------------
    def peek_kernel(self):
        return self.kernel

    def get_dim(self):
        return self.d

    def set_over_smooth(self, v):
        check(lib().nckde_set_over_smooth(self._h, float(v)))
        self._over_smooth = float(v)

    def get_over_smooth(self):
        v = C.c_double()
        check(lib().nckde_get_over_smooth(self._h, C.byref(v)))
        self._over_smooth = v.value
        return v.value

    def set_split_frac(self, v):
        check(lib().nckde_set_split_frac(self._h, float(v)))

    def get_split_frac(self):
        v = C.c_double()
        check(lib().nckde_get_split_frac(self._h, C.byref(v)))
        return v.value

    def set_print_fit(self, v):
        check(lib().nckde_set_print_fit(self._h, int(bool(v))))

    def get_print_fit(self):
        v = C.c_int()
        check(lib().nckde_get_print_fit(self._h, C.byref(v)))
        return bool(v.value)

    def set_cv_type(self, cv_type):
        check(lib().nckde_set_cv_type(self._h, int(cv_type)))

    def get_cv_type(self):
        v = C.c_int()
        check(lib().nckde_get_cv_type(self._h, C.byref(v)))
        return StatsDistCV(v.value)

    def get_cv_stats(self):
        """(objective evaluations, seconds) of the last cross-validation fit;
        no reference counterpart."""
        n, t = C.c_uint(), C.c_double()
        check(lib().nckde_get_cv_stats(self._h, C.byref(n), C.byref(t)))
        return n.value, t.value

    def get_mixed_check(self):
        """(max relative m2lnp error of the mixed-precision probes, whether
        mixed evals fall back to FP64); no reference counterpart."""
        e, f = C.c_double(), C.c_int()
        check(lib().nckde_get_mixed_check(self._h, C.byref(e), C.byref(f)))
        return e.value, bool(f.value)

    def set_eval_split(self, nsplit):
        """Force the eval's kernel-range split count (0: tuned table or
        default rule); tuning hook, no reference counterpart."""
        check(lib().nckde_set_eval_split(self._h, int(nsplit)))

    def set_graphs(self, on):
        """Replay repeated evals of one shape as captured HIP graphs; no
        reference counterpart."""
        check(lib().nckde_set_graphs(self._h, int(bool(on))))

    def set_mixed_tolerance(self, tol):
        """Probe error above which mixed evals run in FP64 (default 5e-7);
        no reference counterpart."""
        check(lib().nckde_set_mixed_tolerance(self._h, float(tol)))

    def set_nnls_umethod(self, umethod):
        """ncm_nnls_set_umethod (ncm_nnls.h:66) of the weight fit (NORMAL in
        the reference, ncm_stats_dist.c:861)."""
        check(lib().nckde_set_nnls_umethod(self._h, int(umethod)))

    def get_nnls_umethod(self):
        v = C.c_int()
        check(lib().nckde_get_nnls_umethod(self._h, C.byref(v)))
        return NNLSUMethod(v.value)

    def set_nnls_refine(self, max_steps):
        """Refinement steps after a badly spread Cholesky solve (0 disables);
        no reference counterpart."""
        check(lib().nckde_set_nnls_refine(self._h, int(max_steps)))

    def get_nnls_refine(self):
        v = C.c_uint()
        check(lib().nckde_get_nnls_refine(self._h, C.byref(v)))
        return v.value

    def get_nnls_rungs(self):
        """Unconstrained solves of the last prepare_interp per rung:
        (Cholesky, symmetric solve, QR, refinement steps)."""
        c = (C.c_uint * 4)()
        check(lib().nckde_get_nnls_rungs(self._h, c))
        return tuple(int(v) for v in c)

    def set_shrink(self, v):
        check(lib().nckde_set_shrink(self._h, float(v)))

    def get_shrink(self):
        """ncm_stats_dist_get_shrink (ncm_stats_dist.h:107)"""
        v = C.c_double()
        check(lib().nckde_get_shrink(self._h, C.byref(v)))
        return v.value

    def set_use_threads(self, v):
        check(lib().nckde_set_use_threads(self._h, int(bool(v))))

    def get_use_threads(self):
        """ncm_stats_dist_get_use_threads (ncm_stats_dist.h:116)"""
        v = C.c_int()
        check(lib().nckde_get_use_threads(self._h, C.byref(v)))
        return bool(v.value)

    def set_cov_type(self, cov_type):
        check(lib().nckde_set_cov_type(self._h, int(cov_type)))

    def get_cov_type(self):
        """ncm_stats_dist_kde_get_cov_type (ncm_stats_dist_kde.h:83)"""
        v = C.c_int()
        check(lib().nckde_get_cov_type(self._h, C.byref(v)))
        return v.value

    def set_cov_fixed(self, cov):
        check(lib().nckde_set_cov_fixed(self._h, _f64(cov)))

    def peek_cov_fixed(self):
        """ncm_stats_dist_kde_peek_cov_fixed (ncm_stats_dist_kde.h:86); None
        when no fixed covariance was set (the reference returns NULL)"""
        out = np.empty((self.d, self.d))
        st = lib().nckde_get_cov_fixed(self._h, out)
        if st == -3:
            return None
        check(st)
        return out

    def set_nearPD_maxiter(self, v):
        check(lib().nckde_set_nearPD_maxiter(self._h, int(v)))

    def get_nearPD_maxiter(self):
        """ncm_stats_dist_kde_get_nearPD_maxiter (ncm_stats_dist_kde.h:80)"""
        v = C.c_uint()
        check(lib().nckde_get_nearPD_maxiter(self._h, C.byref(v)))
        return v.value

    # ---- sample --------------------------------------------------------
    def add_obs(self, x):
        x = _f64(x)
        if x.shape != (self.d,):
            raise ValueError("add_obs: vector of length d expected")
        check(lib().nckde_add_obs(self._h, x))

    def add_obs_array(self, X):
        X = _f64(X)
        check(lib().nckde_add_obs_batch(self._h, X, X.shape[0], self.d))

    def reset(self):
        check(lib().nckde_reset(self._h))

    # ---- prepare -------------------------------------------------------
    def prepare(self):
        check(lib().nckde_prepare(self._h))

    def prepare_interp(self, m2lnp):
        m2lnp = _f64(m2lnp)
        check(lib().nckde_prepare_interp(self._h, m2lnp, m2lnp.shape[0]))

    def compute_IM(self):
        IM = np.empty((self.get_sample_size(), self.get_n_kernels()))
        check(lib().nckde_compute_IM(self._h, IM))
        return IM

    # ---- evaluation ----------------------------------------------------
    def eval_m2lnp_batch(self, X, precision=None):
        X = _f64(np.atleast_2d(X))
        out = np.empty(X.shape[0])
        p = int(self.precision if precision is None else precision)
        check(lib().nckde_eval_m2lnp(self._h, X, X.shape[0], X.shape[1], out, p))
        return out

    def eval_batch(self, X, precision=None):
        X = _f64(np.atleast_2d(X))
        out = np.empty(X.shape[0])
        p = int(self.precision if precision is None else precision)
        check(lib().nckde_eval(self._h, X, X.shape[0], X.shape[1], out, p))
        return out

    def eval_m2lnp(self, x):
        return float(self.eval_m2lnp_batch(_f64(x)[None, :])[0])

    def eval(self, x):
        return float(self.eval_batch(_f64(x)[None, :])[0])

    # ---- sampling ------------------------------------------------------
    def kernel_choose_batch(self, u):
        u = _f64(u)
        idx = np.empty(u.shape[0], dtype=np.uint32)
        check(lib().nckde_kernel_choose(self._h, u, u.shape[0], idx))
        return idx

    def kernel_choose(self, rng: RNG):
        return int(self.kernel_choose_batch(np.array([rng.uniform_gen(0.0, 1.0)]))[0])

    def sample_batch(self, u, z, chi2=None):
        u = _f64(u)
        z = _f64(z).reshape(u.shape[0], self.d)
        out = np.empty((u.shape[0], self.d))
        idx = np.empty(u.shape[0], dtype=np.uint32)
        c2 = None
        if chi2 is not None:
            c2 = _f64(chi2)
        check(lib().nckde_sample(self._h, u, z, c2.ctypes.data if c2 is not None else None,
                                 u.shape[0], out, idx.ctypes.data))
        return out, idx

    def sample(self, x, rng: RNG):
        """ncm_stats_dist_sample: writes a draw into the vector x."""
        u = rng.uniform_gen(0.0, 1.0)
        z = np.array([rng.ugaussian_gen() for _ in range(self.d)])
        c2 = np.array([rng.chisq_gen(self.kernel.nu)]) if self.kernel.kind == 1 else None
        out, _ = self.sample_batch(np.array([u]), z[None, :], c2)
        x[:] = out[0]

    # ---- peeks ---------------------------------------------------------
    def _sz(self, fn):
        v = C.c_size_t()
        check(fn(self._h, C.byref(v)))
        return v.value

    def get_sample_size(self):
        return self._sz(lib().nckde_get_sample_size)

    def get_n_kernels(self):
        return self._sz(lib().nckde_get_n_kernels)

    def get_n_samples(self):
        return self._sz(lib().nckde_get_n_samples)

    def get_href(self):
        v = C.c_double()
        check(lib().nckde_get_href(self._h, C.byref(v)))
        return v.value

    def get_rnorm(self):
        v = C.c_double()
        check(lib().nckde_get_rnorm(self._h, C.byref(v)))
        return v.value

    def get_lnnorm(self, i):
        v = C.c_double()
        check(lib().nckde_get_lnnorm(self._h, int(i), C.byref(v)))
        return v.value

    def peek_weights(self):
        w = np.empty(self.get_n_kernels())
        check(lib().nckde_get_weights(self._h, w))
        return w

    def peek_cov_decomp(self, i):
        U = np.empty((self.d, self.d))
        check(lib().nckde_get_cov_decomp(self._h, int(i), U))
        return U

    def peek_full_cov(self):
        c = np.empty((self.d, self.d))
        check(lib().nckde_get_full_cov(self._h, c))
        return c

    def peek_full_cov_decomp(self):
        c = np.empty((self.d, self.d))
        check(lib().nckde_get_full_cov_decomp(self._h, c))
        return c

    def peek_sample_array(self):
        X = np.empty((self.get_n_samples(), self.d))
        check(lib().nckde_get_sample(self._h, X))
        return X

    def peek_invUsample(self):
        Y = np.empty((self.get_sample_size(), self.d))
        check(lib().nckde_get_invUsample(self._h, Y))
        return Y

    def get_k(self):
        """VKDE: number of neighbours of each local covariance."""
        k = C.c_size_t()
        check(lib().nckde_get_knn(self._h, C.byref(k), None))
        return k.value

    def peek_nnls_state(self):
        """(x, g, r) of the last NNLS fit: the solution before shrink, the
        gradient A^T (f - A x) and the residual f - A x (nckde_get_nnls_state;
        no reference counterpart)."""
        nk, ns = self.get_n_kernels(), self.get_sample_size()
        x, g, r = np.empty(nk), np.empty(nk), np.empty(ns)
        check(lib().nckde_get_nnls_state(self._h, x.ctypes.data, g.ctypes.data, r.ctypes.data))
        return x, g, r

    def peek_knn(self):
        k = C.c_size_t()
        check(lib().nckde_get_knn(self._h, C.byref(k), None))
        idx = np.empty((self.get_n_kernels(), k.value), dtype=np.int32)
        if k.value:
            check(lib().nckde_get_knn(self._h, C.byref(k), idx.ctypes.data))
        return idx

    def get_Ki(self, i):
        """ncm_stats_dist_get_Ki (ncm_stats_dist.c:1810-1828): kernel centre,
        covariance h^2 U^T U, normalisation exp(lnnorm_i) and weight."""
        y = np.empty(self.d)
        cov = np.empty((self.d, self.d))
        n_i, w_i = C.c_double(), C.c_double()
        check(lib().nckde_get_Ki(self._h, int(i), y, cov, C.byref(n_i), C.byref(w_i)))
        return y, cov, n_i.value, w_i.value


class StatsDistKDE(_StatsDist):
    """NcmStatsDistKDE (ncm_stats_dist_kde.h:74-86)."""
    _kind = 0


def _serialize(self, with_state: bool = True) -> bytes:
    """Configuration and sample, plus the prepared state when with_state
    (nckde_serialize); restore with ``deserialize``."""
    n = C.c_size_t()
    check(lib().nckde_serialize(self._h, int(bool(with_state)), None, 0, C.byref(n)))
    buf = C.create_string_buffer(n.value)
    check(lib().nckde_serialize(self._h, int(bool(with_state)), buf, n.value, C.byref(n)))
    return buf.raw[: n.value]


_StatsDist.serialize = _serialize


def deserialize(blob: bytes):
    """Object restored from ``serialize`` output (a StatsDistKDE or
    StatsDistVKDE, prepared if the blob carries the state)."""
    import struct

    # header of nckde_serialize: magic[8], int32 x 10, double x 5 (layout in csrc/api.cpp)
    kind, kernel, d, cv, rot, _ct, _mi, _hf, _st, _pf = struct.unpack_from("=10i", blob, 8)
    nu, over_smooth, _shrink, _local_frac, _sf = struct.unpack_from("=5d", blob, 8 + 40)
    h = C.c_void_p()
    buf = C.create_string_buffer(blob, len(blob))
    check(lib().nckde_deserialize(C.byref(h), buf, len(blob)))
    cls = StatsDistVKDE if kind == 1 else StatsDistKDE
    sd = cls.__new__(cls)
    sd._h = h
    sd.kernel = StatsDistKernelST(d, nu) if kernel == 1 else StatsDistKernelGauss(d)
    sd.d = d
    sd._over_smooth = over_smooth
    sd.precision = Precision.FP64
    return sd


class StatsDistVKDE(StatsDistKDE):
    """NcmStatsDistVKDE (ncm_stats_dist_vkde.h:53-62)."""
    _kind = 1

    def __init__(self, sdk, cv_type=StatsDistCV.NONE):
        super().__init__(sdk, cv_type)

    def set_local_frac(self, v):
        check(lib().nckde_set_local_frac(self._h, float(v)))

    def get_local_frac(self):
        """ncm_stats_dist_vkde_get_local_frac (ncm_stats_dist_vkde.h:59)"""
        v = C.c_double()
        check(lib().nckde_get_local_frac(self._h, C.byref(v)))
        return v.value

    def set_use_rot_href(self, v):
        check(lib().nckde_set_use_rot_href(self._h, int(bool(v))))

    def get_use_rot_href(self):
        """ncm_stats_dist_vkde_get_use_rot_href (ncm_stats_dist_vkde.h:62)"""
        v = C.c_int()
        check(lib().nckde_get_use_rot_href(self._h, C.byref(v)))
        return bool(v.value)
