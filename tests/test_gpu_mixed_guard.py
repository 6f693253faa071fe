"""The mixed-precision eval against the oracle where its split products are
weakest: narrow kernels far from the sample mean.

The mixed form computes z' = A_i (s y) - c_i with FP16 hi/lo operands and
FP32 accumulation, so its absolute error grows with |B_i| |y| rather than
with |z|.  When the factors are built, probe points next to the most exposed
kernels are evaluated in both precisions, and a probe error above 5e-7
sends the handle's mixed evals to the FP64 path.  Either way every result
must stay within the 1e-6 bar against the FP64 oracle.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1.0)


def _cluster_target(seed, d, n, frac, centre, width):
    rng = np.random.default_rng(seed)
    nc = int(n * frac)
    X = np.concatenate([rng.standard_normal((n - nc, d)),
                        centre + width * rng.standard_normal((nc, d))])
    rng.shuffle(X)
    return rng, X


@pytest.mark.parametrize("d,n,local_frac,centre,width", [
    (2, 3000, 0.01, 8.0, 0.02),     # tight cluster 8 sigma out, k = 30
    (2, 3000, 0.005, 3.0, 0.05),    # the advisor's regime: width ~0.2 kernels at |y| ~ 2.5
    (3, 4000, 0.01, 5.0, 0.01),
])
def test_mixed_narrow_kernels_far_out(oracle, d, n, local_frac, centre, width):
    import numcosmo_amd as na
    _say(f"[mixed guard] d={d} n={n} local_frac={local_frac} cluster at {centre} width {width}")
    rng, X = _cluster_target(17 + d, d, n, 0.1, centre, width)
    sd = na.StatsDistVKDE.new(na.StatsDistKernelST.new(d, 3.0))
    sd.set_local_frac(local_frac)
    sd.add_obs_array(X)
    sd.prepare()
    od = oracle.StatsDist(oracle.VKDE, oracle.ST, d, 3.0)
    od.set_local_frac(local_frac)
    od.add_obs_array(X)
    od.prepare()
    # points at and around the cluster, and around the bulk
    Xc = X[np.argsort(np.abs(X - centre).sum(1))[:300]]
    Xe = np.concatenate([Xc + 0.5 * width * rng.standard_normal(Xc.shape),
                         rng.standard_normal((300, d))])
    ref = od.eval_m2lnp_batch(Xe)
    got = sd.eval_m2lnp_batch(Xe, precision=na.Precision.MIXED)
    err, fb = sd.get_mixed_check()
    e = np.max(_rel(got, ref))
    _say(f"   probe error {err:.2e}, fallback {fb}; mixed-path error vs oracle {e:.2e}")
    assert e <= 1e-6


def test_mixed_stays_mixed_on_the_banana():
    """The guard does not fire on a well-resolved target (d=8 banana,
    n=4000, local_frac 0.05): the probes stay far below the bar."""
    import numcosmo_amd as na
    from numcosmo_amd import synth
    tgt, X, m2, Xe = synth.workload("banana", 4000, 1000, 8, 2)
    sd = na.StatsDistVKDE.new(na.StatsDistKernelST.new(8, 3.0))
    sd.add_obs_array(X)
    sd.prepare_interp(m2)
    err, fb = sd.get_mixed_check()
    _say(f"   banana d=8: probe error {err:.2e}, fallback {fb}")
    assert not fb and err < 5e-7
