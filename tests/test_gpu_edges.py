"""Edge cases of the device engine against the FP64 oracle.

- empty point and draw batches;
- point counts that leave partial tiles, waves and blocks (1 ... 1025);
- kernel counts that leave padding kernels in the last MFMA group;
- d = 1 (one live row of the smallest tile).

Bars as in test_gpu_parity.py: FP64 eval <= 1e-12 relative, mixed eval
<= 1e-6 relative (|dm2lnp| / max(|m2lnp|, 1)), choices bit-exact.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GAUSS, ST = 0, 1


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _pair(oracle, kind, kernel, d, X, nu=3.0, local_frac=None):
    import numcosmo_amd as na
    sdk = na.StatsDistKernelGauss.new(d) if kernel == GAUSS else na.StatsDistKernelST.new(d, nu)
    sd = (na.StatsDistVKDE if kind == 1 else na.StatsDistKDE).new(sdk, na.StatsDistCV.NONE)
    od = oracle.StatsDist(kind, kernel, d, nu)
    if local_frac is not None:
        sd.set_local_frac(local_frac)
        od.set_local_frac(local_frac)
    sd.add_obs_array(X)
    od.add_obs_array(X)
    sd.prepare()
    od.prepare()
    return sd, od


def _rel(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1.0)


def test_empty_batches(oracle):
    g = np.random.default_rng(1)
    X = g.normal(size=(40, 3))
    sd, _ = _pair(oracle, 1, ST, 3, X, local_frac=0.2)
    E = np.empty((0, 3))
    for p in (0, 1):
        assert sd.eval_m2lnp_batch(E, precision=p).shape == (0,)
        assert sd.eval_batch(E, precision=p).shape == (0,)
    assert sd.kernel_choose_batch(np.empty(0)).shape == (0,)
    out, idx = sd.sample_batch(np.empty(0), np.empty((0, 3)), np.empty(0))
    assert out.shape == (0, 3) and idx.shape == (0,)


@pytest.mark.parametrize("kind,kernel,d,n,lf", [(1, ST, 8, 300, 0.1), (0, GAUSS, 3, 200, None),
                                                (1, GAUSS, 16, 200, 0.2)])
def test_ragged_point_counts(oracle, kind, kernel, d, n, lf):
    """Point counts around the tile (32), wave and block sizes of both eval
    paths: every point of a partial tile or block is computed and stored, and
    no padding lane leaks into a real point's result."""
    g = np.random.default_rng(10 + d)
    X = g.normal(size=(n, d))
    sd, od = _pair(oracle, kind, kernel, d, X, local_frac=lf)
    for m in (1, 2, 31, 33, 127, 129, 255, 257, 511, 513, 1025):
        Y = g.normal(size=(m, d)) * 1.5
        ref = od.eval_m2lnp_batch(Y)
        e64 = _rel(sd.eval_m2lnp_batch(Y, precision=0), ref).max()
        emx = _rel(sd.eval_m2lnp_batch(Y, precision=1), ref).max()
        assert e64 <= 1e-12, (m, e64)
        assert emx <= 1e-6, (m, emx)
    _say(f"   kind={kind} d={d}: ragged m ok")


@pytest.mark.parametrize("kind,kernel,d,n,lf", [(0, ST, 3, 5, None), (1, GAUSS, 2, 12, 0.5),
                                                (1, ST, 8, 21, 0.6), (0, GAUSS, 16, 19, None)])
def test_padding_kernels(oracle, kind, kernel, d, n, lf):
    """Kernel counts that are not a multiple of the kernels per MFMA group
    (8 at d <= 4, 4 at d <= 8, 2 at d <= 16): the padding kernels carry zero
    weight and change no result."""
    g = np.random.default_rng(20 + n)
    X = g.normal(size=(n, d))
    sd, od = _pair(oracle, kind, kernel, d, X, local_frac=lf)
    Y = np.vstack([X, g.normal(size=(300, d))])
    ref = od.eval_m2lnp_batch(Y)
    assert _rel(sd.eval_m2lnp_batch(Y, precision=0), ref).max() <= 1e-12
    assert _rel(sd.eval_m2lnp_batch(Y, precision=1), ref).max() <= 1e-6
    u = np.concatenate([[0.0, np.nextafter(1.0, 0.0)], g.uniform(size=200)])
    np.testing.assert_array_equal(sd.kernel_choose_batch(u), [od.kernel_choose(p) for p in u])


@pytest.mark.parametrize("kind,kernel,nu", [(0, GAUSS, 3.0), (1, ST, 1.0)])
def test_one_dimension(oracle, kind, kernel, nu):
    g = np.random.default_rng(30 + kind)
    X = g.standard_t(4.0, size=(200, 1))
    sd, od = _pair(oracle, kind, kernel, 1, X, nu=nu)
    Y = np.linspace(-30.0, 30.0, 401).reshape(-1, 1)
    ref = od.eval_m2lnp_batch(Y)
    assert _rel(sd.eval_m2lnp_batch(Y, precision=0), ref).max() <= 1e-12
    assert _rel(sd.eval_m2lnp_batch(Y, precision=1), ref).max() <= 1e-6
