"""ctypes wrapper of the FP64 CPU oracle (liboracle_nckde.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` as the checker.  The product package
``numcosmo_amd`` never imports this module.

The object API mirrors NumCosmo's ``NcmStatsDist`` (numcosmo/ncm/stats/
ncm_stats_dist.h:168-222) so that tests read like the reference's own.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_nckde.so")
# the same restatement built -O3 with FMA contraction and AVX2 (x86-64-v3):
# the CPU-baseline build, timed by bench.py, never the parity checker
_FAST_PATH = os.path.join(_HERE, "liboracle_nckde_fast.so")
_lib = None
_fast = None

KDE, VKDE = 0, 1
GAUSS, ST = 0, 1
COV_SAMPLE, COV_FIXED, COV_ROBUST_DIAG, COV_ROBUST = 0, 1, 2, 3
CV_NONE, CV_SPLIT, CV_SPLIT_NOFIT, CV_LOO = 0, 1, 2, 3

FN1 = C.CFUNCTYPE(C.c_double, C.c_void_p, C.c_double)
RESFN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_double, C.POINTER(C.c_double))

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build() -> str:
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = _bind(C.CDLL(_LIB_PATH))
    return _lib


def fast_lib():
    """The CPU-baseline build of the restatement (timing only)."""
    global _fast
    if _fast is None:
        if not os.path.exists(_FAST_PATH):
            build()
        _fast = _bind(C.CDLL(_FAST_PATH))
    return _fast


def _bind(L):
    sig = {
        "ok_kernel_rot_bandwidth": (C.c_double, [C.c_int, C.c_double, C.c_int, C.c_double]),
        "ok_kernel_lnnorm": (C.c_double, [C.c_int, C.c_double, C.c_int, _dp]),
        "ok_kernel_eval_unnorm": (C.c_double, [C.c_int, C.c_double, C.c_int, C.c_double]),
        "ok_kernel_sum0": (None, [C.c_int, C.c_double, C.c_int, C.c_int, _dp, _dp, _dp,
                                  C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "ok_kernel_sum1": (None, [C.c_int, C.c_double, C.c_int, C.c_int, _dp, _dp, C.c_double,
                                  C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "ok_cholesky_upper": (C.c_int, [C.c_int, _dp, C.c_int]),
        "ok_cholesky_lndet": (C.c_double, [C.c_int, _dp, C.c_int]),
        "ok_nearPD": (C.c_int, [C.c_int, _dp, C.c_int]),
        "ok_cholesky_decomp_fallback": (None, [C.c_int, _dp, _dp, C.c_int]),
        "ok_jacobi_eigen": (C.c_int, [C.c_int, _dp, _dp, _dp]),
        "ok_cov_welford": (None, [C.c_int, C.c_int, _dp, C.c_void_p, _dp]),
        "ok_knn": (None, [C.c_int, C.c_int, _dp, C.c_int, C.c_int, _ip, C.c_void_p]),
        "ok_nnls": (C.c_double, [C.c_int, C.c_int, _dp, _dp, _dp]),
        "ok_nnls_m": (C.c_double, [C.c_int, C.c_int, _dp, _dp, _dp, C.c_int]),
        "ok_nnls_rung_counts": (None, [np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS"),
                                       C.c_int]),
        "ok_lstsq_qr": (None, [C.c_int, C.c_int, _dp, _dp, _dp]),
        "ok_set_nnls_umethod": (None, [C.c_void_p, C.c_int]),
        "ok_set_num_threads": (None, [C.c_int]),
        "ok_get_max_threads": (C.c_int, []),
        "ok_new": (C.c_void_p, [C.c_int, C.c_int, C.c_double, C.c_int]),
        "ok_free": (None, [C.c_void_p]),
        "ok_set_over_smooth": (None, [C.c_void_p, C.c_double]),
        "ok_set_shrink": (None, [C.c_void_p, C.c_double]),
        "ok_set_local_frac": (None, [C.c_void_p, C.c_double]),
        "ok_set_use_rot_href": (None, [C.c_void_p, C.c_int]),
        "ok_set_nearPD_maxiter": (None, [C.c_void_p, C.c_int]),
        "ok_set_use_threads": (None, [C.c_void_p, C.c_int]),
        "ok_set_cov_type": (C.c_int, [C.c_void_p, C.c_int]),
        "ok_set_cov_fixed": (C.c_int, [C.c_void_p, _dp]),
        "ok_reset": (None, [C.c_void_p]),
        "ok_add_obs": (None, [C.c_void_p, _dp]),
        "ok_prepare": (C.c_int, [C.c_void_p]),
        "ok_prepare_interp": (C.c_int, [C.c_void_p, _dp, C.c_int]),
        "ok_compute_IM": (C.c_int, [C.c_void_p, _dp]),
        "ok_eval_m2lnp": (C.c_double, [C.c_void_p, _dp]),
        "ok_eval": (C.c_double, [C.c_void_p, _dp]),
        "ok_eval_m2lnp_batch": (None, [C.c_void_p, _dp, C.c_int, _dp]),
        "ok_eval_batch": (None, [C.c_void_p, _dp, C.c_int, _dp]),
        "ok_kernel_choose": (C.c_int, [C.c_void_p, C.c_double]),
        "ok_sample": (C.c_int, [C.c_void_p, C.c_double, _dp, C.c_double, _dp]),
        "ok_set_weights": (None, [C.c_void_p, _dp]),
        "ok_set_prepared_vkde": (None, [C.c_void_p, C.c_int, _dp, _dp, _dp, C.c_double]),
        "ok_get_dim": (C.c_int, [C.c_void_p]),
        "ok_get_sample_size": (C.c_int, [C.c_void_p]),
        "ok_get_n_kernels": (C.c_int, [C.c_void_p]),
        "ok_get_n_samples": (C.c_int, [C.c_void_p]),
        "ok_get_k": (C.c_int, [C.c_void_p]),
        "ok_get_href": (C.c_double, [C.c_void_p]),
        "ok_get_rnorm": (C.c_double, [C.c_void_p]),
        "ok_get_lnnorm": (C.c_double, [C.c_void_p, C.c_int]),
        "ok_get_weights": (None, [C.c_void_p, _dp]),
        "ok_get_cov_decomp": (None, [C.c_void_p, C.c_int, _dp]),
        "ok_get_full_cov": (None, [C.c_void_p, _dp]),
        "ok_get_full_cov_decomp": (None, [C.c_void_p, _dp]),
        "ok_get_sample": (None, [C.c_void_p, _dp]),
        "ok_get_invUsample": (None, [C.c_void_p, _dp]),
        "ok_get_knn": (None, [C.c_void_p, _ip]),
        "ok_last_error": (C.c_char_p, []),
        "ok_qn": (C.c_double, [C.c_int, _dp]),
        "ok_cov_robust_diag": (C.c_int, [C.c_int, C.c_int, _dp, C.c_void_p, _dp]),
        "ok_cov_robust_ogk": (C.c_int, [C.c_int, C.c_int, _dp, C.c_void_p, _dp]),
        "ok_set_cv_type": (None, [C.c_void_p, C.c_int]),
        "ok_set_split_frac": (None, [C.c_void_p, C.c_double]),
        "ok_get_over_smooth": (C.c_double, [C.c_void_p]),
        "ok_cv_objective": (C.c_double, [C.c_void_p, C.c_double]),
        "ok_nm_minimize1": (C.c_int, [FN1, C.c_void_p, C.POINTER(C.c_double), C.c_double, C.c_double,
                                      C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
        "ok_lm_dif1": (C.c_int, [RESFN, C.c_void_p, C.POINTER(C.c_double), C.c_int, C.c_int, _dp,
                                 C.POINTER(C.c_int)]),
        "ok_pair_stats_update": (None, [_dp, C.c_double, C.c_double]),
        "ok_draw": (None, [C.c_uint64, C.c_uint64, C.c_uint32, C.c_int, C.c_int, C.POINTER(C.c_double), _dp,
                           C.POINTER(C.c_double)]),
        "ok_host_normal": (C.c_double, [C.c_uint64, C.c_uint64]),
        "ok_philox4x32_10": (None, [np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS"),
                                    np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")]),
        "ok_sample_at": (None, [C.c_void_p, C.c_int, _dp, C.c_double, _dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class OracleError(RuntimeError):
    pass


# ---------------------------------------------------------------- kernels --
def rot_bandwidth(kernel, nu, d, n):
    return lib().ok_kernel_rot_bandwidth(kernel, nu, d, n)


def kernel_lnnorm(kernel, nu, d, U):
    return lib().ok_kernel_lnnorm(kernel, nu, d, _f64(U))


def eval_unnorm(kernel, nu, d, chi2):
    return lib().ok_kernel_eval_unnorm(kernel, nu, d, chi2)


def sum0(kernel, nu, d, chi2, w, lnnorms):
    g, lam = C.c_double(), C.c_double()
    chi2 = _f64(chi2)
    lib().ok_kernel_sum0(kernel, nu, d, len(chi2), chi2, _f64(w), _f64(lnnorms), C.byref(g), C.byref(lam))
    return g.value, lam.value


def sum1(kernel, nu, d, chi2, w, lnnorm):
    g, lam = C.c_double(), C.c_double()
    chi2 = _f64(chi2)
    lib().ok_kernel_sum1(kernel, nu, d, len(chi2), chi2, _f64(w), lnnorm, C.byref(g), C.byref(lam))
    return g.value, lam.value


# ---------------------------------------------------------------- algebra --
def cholesky_upper(A):
    U = _f64(A).copy()
    info = lib().ok_cholesky_upper(U.shape[0], U, U.shape[0])
    return info, U


def cholesky_lndet(U):
    U = _f64(U)
    return lib().ok_cholesky_lndet(U.shape[0], U, U.shape[0])


def nearPD(A, maxiter=200):
    U = _f64(A).copy()
    info = lib().ok_nearPD(U.shape[0], U, maxiter)
    return info, U


def cholesky_decomp_fallback(cov, maxiter=200):
    cov = _f64(cov)
    U = np.empty_like(cov)
    lib().ok_cholesky_decomp_fallback(cov.shape[0], cov, U, maxiter)
    return U


def jacobi_eigen(A):
    A = _f64(A)
    d = A.shape[0]
    eva = np.empty(d)
    eve = np.empty((d, d))
    lib().ok_jacobi_eigen(d, A, eva, eve)
    return eva, eve


def cov_welford(X, rows=None):
    X = _f64(X)
    d = X.shape[1]
    cov = np.empty((d, d))
    if rows is None:
        lib().ok_cov_welford(d, X.shape[0], X, None, cov)
    else:
        r = np.ascontiguousarray(rows, dtype=np.int32)
        lib().ok_cov_welford(d, len(r), X, r.ctypes.data, cov)
    return cov


def qn(x):
    """Qn scale estimator (ncm_stats_vec.c:1849-1855, gsl_stats_Qn_from_sorted_data)."""
    x = _f64(x)
    return lib().ok_qn(x.shape[0], x)


def cov_robust(X, kind, rows=None):
    """ROBUST_DIAG (kind 2) or OGK (kind 3) covariance of the rows of X."""
    X = _f64(X)
    d = X.shape[1]
    cov = np.empty((d, d))
    r = None if rows is None else np.ascontiguousarray(rows, dtype=np.int32)
    n = X.shape[0] if r is None else len(r)
    f = lib().ok_cov_robust_diag if kind == COV_ROBUST_DIAG else lib().ok_cov_robust_ogk
    if f(d, n, X, None if r is None else r.ctypes.data, cov) != 0:
        raise OracleError(lib().ok_last_error().decode())
    return cov


def knn(Y, q, k):
    Y = _f64(Y)
    idx = np.empty(k, dtype=np.int32)
    d2 = np.empty(k)
    lib().ok_knn(Y.shape[0], Y.shape[1], Y, q, k, idx, d2.ctypes.data)
    return idx, d2


NNLS_NORMAL, NNLS_NORMAL_LU, NNLS_QR = 0, 1, 2   # NcmNNLSUMethod (ncm_nnls.h:51-60)


def nnls(A, f, umethod=NNLS_NORMAL):
    A = _f64(A)
    x = np.zeros(A.shape[1])
    rnorm = lib().ok_nnls_m(A.shape[0], A.shape[1], A, _f64(f), x, int(umethod))
    return x, rnorm


def nnls_rung_counts(reset=True):
    """Unconstrained solves per rung (Cholesky, symmetric solve, QR) since the
    last reset, in this thread."""
    c = np.zeros(3, dtype=np.uint32)
    lib().ok_nnls_rung_counts(c, int(reset))
    return [int(v) for v in c]


def set_num_threads(n):
    """OpenMP team size of the oracle's batched calls (both builds)."""
    lib().ok_set_num_threads(int(n))


def get_max_threads():
    return lib().ok_get_max_threads()


def lstsq_qr(A, f):
    """min |A x - f| by the oracle's Householder QR rung (dgels)."""
    A = _f64(A)
    x = np.zeros(A.shape[1])
    lib().ok_lstsq_qr(A.shape[0], A.shape[1], A, _f64(f), x)
    return x


# ------------------------------------------------------ cross-validation --
def nm_minimize1(f, x0, step=0.1, epsabs=1e-3, maxiter=1000):
    """GSL nmsimplex2 in one dimension (the minimiser of ncm_stats_dist.c:664-701).

    Returns (x, fval, iterations, status); status 0 = size below epsabs."""
    cb = FN1(lambda ctx, x: float(f(x)))
    x = C.c_double(x0)
    it, fv = C.c_int(), C.c_double()
    st = lib().ok_nm_minimize1(cb, None, C.byref(x), step, epsabs, maxiter, C.byref(it), C.byref(fv))
    return x.value, fv.value, it.value, st


def lm_dif1(res, p0, n, itmax=10000, opts=(1e-3, 1e-7, 1e-7, 1e-10, 1e-6)):
    """levmar dlevmar_dif for one parameter (ncm_stats_dist.c:1063-1065).

    res(p) -> n residuals; returns (p, stop reason, iterations)."""
    def cb(ctx, p, hx):
        r = np.asarray(res(p), dtype=np.float64)
        for i in range(n):
            hx[i] = r[i]
        return 0

    fn = RESFN(cb)
    p = C.c_double(p0)
    it = C.c_int()
    stop = lib().ok_lm_dif1(fn, None, C.byref(p), n, itmax, _f64(opts), C.byref(it))
    return p.value, stop, it.value


def pair_stats(x0, x1):
    """NcmStatsVec COV statistics of two series: (mean0, mean1, var0, var1, cov01)."""
    st = np.zeros(8)
    for a, b in zip(x0, x1):
        lib().ok_pair_stats_update(st, float(a), float(b))
    return st[3], st[4], st[5] * st[2], st[6] * st[2], st[7] * st[2]


def draw(key, ctr, j, d, nu_int=0):
    """Draw j of the build's counter-based stream (key, ctr): (u, z[d], chi2)."""
    u, c = C.c_double(), C.c_double()
    z = np.empty(d)
    lib().ok_draw(key, ctr, j, d, nu_int, C.byref(u), z, C.byref(c))
    return u.value, z, c.value


def philox4x32_10(ctr, key):
    """One Philox4x32-10 block: 4 counter words, 2 key words -> 4 output words."""
    c = np.array(ctr, dtype=np.uint32)
    lib().ok_philox4x32_10(c, np.array(key, dtype=np.uint32))
    return [int(v) for v in c]


def host_normal(key, ctr):
    return lib().ok_host_normal(key, ctr)


# ------------------------------------------------------------ stats dist --
class StatsDist:
    """Oracle mirror of NcmStatsDistKDE / NcmStatsDistVKDE.

    fast=True binds the CPU-baseline build (liboracle_nckde_fast.so, -O3 with
    FMA contraction): for timing, never for parity."""

    def __init__(self, kind, kernel, d, nu=3.0, cv_type=CV_NONE, fast=False):
        self._L = fast_lib() if fast else lib()
        self._h = self._L.ok_new(kind, kernel, float(nu), int(d))
        self.kind, self.kernel, self.d, self.nu = kind, kernel, int(d), float(nu)
        self._L.ok_set_cv_type(self._h, int(cv_type))

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.ok_free(self._h)
            self._h = None

    def _check(self, ret):
        if ret != 0:
            raise OracleError(self._L.ok_last_error().decode())

    # configuration
    def set_over_smooth(self, v): self._L.ok_set_over_smooth(self._h, v)
    def set_shrink(self, v): self._L.ok_set_shrink(self._h, v)
    def set_local_frac(self, v): self._L.ok_set_local_frac(self._h, v)
    def set_use_rot_href(self, v): self._L.ok_set_use_rot_href(self._h, int(v))
    def set_nearPD_maxiter(self, v): self._L.ok_set_nearPD_maxiter(self._h, int(v))
    def set_use_threads(self, v): self._L.ok_set_use_threads(self._h, int(v))
    def set_nnls_umethod(self, v): self._L.ok_set_nnls_umethod(self._h, int(v))
    def set_cov_type(self, v): self._check(self._L.ok_set_cov_type(self._h, int(v)))
    def set_cov_fixed(self, cov): self._check(self._L.ok_set_cov_fixed(self._h, _f64(cov)))
    def set_cv_type(self, v): self._L.ok_set_cv_type(self._h, int(v))
    def set_split_frac(self, v): self._L.ok_set_split_frac(self._h, float(v))
    def get_over_smooth(self): return self._L.ok_get_over_smooth(self._h)

    def cv_objective(self, lnos):
        """Objective of the CV type at ln(over_smooth) on the prepared object."""
        return self._L.ok_cv_objective(self._h, float(lnos))

    def sample_at(self, i, z, chi2=1.0):
        x = np.empty(self.d)
        self._L.ok_sample_at(self._h, int(i), _f64(z), float(chi2), x)
        return x

    def reset(self): self._L.ok_reset(self._h)

    def add_obs(self, x):
        self._L.ok_add_obs(self._h, _f64(x))

    def add_obs_array(self, X):
        for row in _f64(X):
            self._L.ok_add_obs(self._h, row)

    def prepare(self):
        self._check(self._L.ok_prepare(self._h))

    def prepare_interp(self, m2lnp):
        m2lnp = _f64(m2lnp)
        self._check(self._L.ok_prepare_interp(self._h, m2lnp, len(m2lnp)))

    def compute_IM(self):
        n = self.get_sample_size()
        IM = np.empty((n, self.get_n_kernels()))
        self._check(self._L.ok_compute_IM(self._h, IM))
        return IM

    def eval_m2lnp(self, x):
        return self._L.ok_eval_m2lnp(self._h, _f64(x))

    def eval(self, x):
        return self._L.ok_eval(self._h, _f64(x))

    def eval_m2lnp_batch(self, X):
        X = _f64(X)
        out = np.empty(X.shape[0])
        self._L.ok_eval_m2lnp_batch(self._h, X, X.shape[0], out)
        return out

    def eval_batch(self, X):
        X = _f64(X)
        out = np.empty(X.shape[0])
        self._L.ok_eval_batch(self._h, X, X.shape[0], out)
        return out

    def kernel_choose(self, u):
        return self._L.ok_kernel_choose(self._h, float(u))

    def sample(self, u, z, chi2=1.0):
        x = np.empty(self.d)
        i = self._L.ok_sample(self._h, float(u), _f64(z), float(chi2), x)
        return i, x

    def set_weights(self, w):
        self._L.ok_set_weights(self._h, _f64(w))

    def set_prepared_vkde(self, X, U, w, href):
        """Checker hook: a prepared VKDE holding the given centres X (n x d),
        upper factors U (n x d x d), weights and bandwidth; the oracle
        recomputes the lnnorms from U."""
        X = _f64(X)
        self._L.ok_set_prepared_vkde(self._h, X.shape[0], X, _f64(U), _f64(w), float(href))

    # peeks
    def get_dim(self): return self._L.ok_get_dim(self._h)
    def get_sample_size(self): return self._L.ok_get_sample_size(self._h)
    def get_n_kernels(self): return self._L.ok_get_n_kernels(self._h)
    def get_n_samples(self): return self._L.ok_get_n_samples(self._h)
    def get_k(self): return self._L.ok_get_k(self._h)
    def get_href(self): return self._L.ok_get_href(self._h)
    def get_rnorm(self): return self._L.ok_get_rnorm(self._h)
    def get_lnnorm(self, i): return self._L.ok_get_lnnorm(self._h, int(i))

    def get_weights(self):
        w = np.empty(self.get_n_kernels())
        self._L.ok_get_weights(self._h, w)
        return w

    def get_cov_decomp(self, i):
        U = np.empty((self.d, self.d))
        self._L.ok_get_cov_decomp(self._h, int(i), U)
        return U

    def get_lnnorms_raw(self):
        """lnnorm_i without the d*ln(href) term (the kernel's own)."""
        lnh = self.d * np.log(self.get_href())
        return np.array([self.get_lnnorm(i) for i in range(self.get_n_kernels())]) - lnh

    def get_full_cov(self):
        c = np.empty((self.d, self.d))
        self._L.ok_get_full_cov(self._h, c)
        return c

    def get_full_cov_decomp(self):
        c = np.empty((self.d, self.d))
        self._L.ok_get_full_cov_decomp(self._h, c)
        return c

    def get_sample(self):
        X = np.empty((self.get_n_samples(), self.d))
        self._L.ok_get_sample(self._h, X)
        return X

    def get_invUsample(self):
        Y = np.empty((self.get_sample_size(), self.d))
        self._L.ok_get_invUsample(self._h, Y)
        return Y

    def get_knn(self):
        k = self.get_k()
        idx = np.empty((self.get_n_kernels(), k), dtype=np.int32)
        self._L.ok_get_knn(self._h, idx)
        return idx


# ---- APES transition densities (pure Python; small ensembles) -----------------
ERF_BOUND = 3.0


def log_normal_gaussian_integral(zl, zu):
    """ln |integral of the standard normal over [zl, zu]| by the branches of
    ncm_util.c:646-690 (erfc in either tail, log1p when both ends lie
    beyond ERF_BOUND / sqrt 2 on opposite sides, erf otherwise)."""
    import math
    if zl == zu:
        return -math.inf
    ul, uu = (zl, zu) if zl < zu else (zu, zl)
    ul *= math.sqrt(0.5)
    uu *= math.sqrt(0.5)
    if ul > ERF_BOUND:
        return math.log(abs(0.5 * (math.erfc(ul) - math.erfc(uu))))
    if uu < -ERF_BOUND:
        return math.log(abs(0.5 * (math.erfc(-ul) - math.erfc(-uu))))
    if uu > ERF_BOUND and ul < -ERF_BOUND:
        return math.log1p(-0.5 * (math.erfc(uu) + math.erfc(-ul)))
    return math.log(abs(0.5 * (math.erf(uu) - math.erf(ul))))


def apes_rw_std(full_cov):
    """Random-walk scales of ncm_fit_esmcmc_walker_apes.c:647-690: a quarter
    of the square root of the diagonal of the sample covariance."""
    v = np.diag(np.asarray(full_cov, dtype=np.float64))
    if np.any(v <= 0.0):
        raise OracleError("non-positive covariance diagonal")
    return 0.25 * np.sqrt(v)


def apes_transition_m2lnp(sd, theta, thetastar, lb, ub, std, rw_prob):
    """-2 ln q(theta -> thetastar) of ncm_fit_esmcmc_walker_apes.c:819-859:
    the density ``sd`` at thetastar mixed, with weight rw_prob, with a
    Gaussian random walk centred at theta and truncated to [lb, ub]."""
    import math
    m_sd = sd.eval_m2lnp(np.asarray(thetastar, dtype=np.float64))
    if rw_prob <= 0.0:
        return m_sd
    m_rw = 0.0
    for i in range(len(theta)):
        ln_norm = 0.5 * math.log(2.0 * math.pi) + math.log(std[i]) + \
            log_normal_gaussian_integral((lb[i] - theta[i]) / std[i], (ub[i] - theta[i]) / std[i])
        m_rw += ((thetastar[i] - theta[i]) / std[i]) ** 2 + 2.0 * ln_norm
    m_rw += -2.0 * math.log(rw_prob)
    m_sd += -2.0 * math.log1p(-rw_prob)
    if m_sd < m_rw:
        return m_sd - 2.0 * math.log1p(math.exp(-0.5 * (m_rw - m_sd)))
    return m_rw - 2.0 * math.log1p(math.exp(-0.5 * (m_sd - m_rw)))
