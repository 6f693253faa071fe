"""Captured eval launch sequences (nckde_set_graphs).

An eval repeated with one shape on one set of buffers is captured into a HIP
graph at its second call and replayed afterwards.  Every replay is compared
bit for bit with the eager eval of the same points, with other eager work
(another handle's evals, synchronisations) between the replays, in both
precisions; a refit and a change of the weights must reach the replays.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("prec", [1, 0, 2], ids=["MIXED", "FP64", "FP32"])
def test_graph_replays_match_eager(prec):
    import numcosmo_amd as na
    from numcosmo_amd import synth

    d, n, m = 8, 2000, 20000
    tgt, X, m2, Xe = synth.workload("banana", n, m, d, 2)
    _say(f"[graphs] prec={prec} d={d} n={n} m={m}")
    sd = na.StatsDistVKDE.new(na.StatsDistKernelST.new(d, 3.0))
    eager = na.StatsDistVKDE.new(na.StatsDistKernelST.new(d, 3.0))
    other = na.StatsDistKDE.new(na.StatsDistKernelGauss.new(d))
    for h in (sd, eager):
        h.add_obs_array(X)
        h.prepare_interp(m2)
    other.add_obs_array(X[:500])
    other.prepare()
    sd.set_graphs(True)
    ref = eager.eval_m2lnp_batch(Xe, precision=prec)
    for rep in range(6):
        got = sd.eval_m2lnp_batch(Xe, precision=prec)
        assert np.array_equal(got, ref), f"replay {rep}"
        other.eval_m2lnp_batch(Xe[:3000], precision=1)     # eager work between replays
    # new weights (a refit on another sample) reach the replays
    _, X2, m22, _ = synth.workload("banana", n, 10, d, 7)
    for h in (sd, eager):
        h.reset()
        h.add_obs_array(X2)
        h.prepare_interp(m22)
    ref2 = eager.eval_m2lnp_batch(Xe, precision=prec)
    for rep in range(4):
        got = sd.eval_m2lnp_batch(Xe, precision=prec)
        assert np.array_equal(got, ref2), f"replay {rep} after the refit"
    assert not np.array_equal(ref2, ref)
    _say(f"[graphs] prec={prec}: 10 replays bit-identical to the eager evals")
