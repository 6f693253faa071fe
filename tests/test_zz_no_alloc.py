"""nckde_eval_m2lnp_device_noalloc: the allocation-free eval against the
allocating nckde_eval_m2lnp_device, through the C-ABI.

Both run the same kernels with the same split counts, so the results are
bit-identical.  Two calls in a row, the second with fewer points (the
shared scratch keeps the first call's contents), in both precisions and for
KDE and VKDE; an oversized batch is refused on the host.
"""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _calls(sd, Xd, outd, prec, stream, noalloc):
    from numcosmo_amd._lib import lib
    f = lib().nckde_eval_m2lnp_device_noalloc if noalloc else lib().nckde_eval_m2lnp_device
    return f(sd._h, C.c_void_p(Xd.data_ptr()), Xd.shape[0], C.c_void_p(outd.data_ptr()), int(prec),
             C.c_void_p(stream.cuda_stream))


@pytest.mark.parametrize("kind", ["VKDE", "KDE"])
@pytest.mark.parametrize("prec", [1, 0], ids=["MIXED", "FP64"])
def test_noalloc_matches_allocating_eval(kind, prec):
    import torch

    import numcosmo_amd as na
    from numcosmo_amd import synth
    from numcosmo_amd._lib import check

    d, n = 8, 3000
    tgt, X, m2, Xe = synth.workload("banana", n, 40000, d, 2)
    _say(f"[noalloc] {kind} prec={prec} d={d} n={n}: m=40000 then m=7001")
    cls = na.StatsDistVKDE if kind == "VKDE" else na.StatsDistKDE
    sd = cls.new(na.StatsDistKernelST.new(d, 3.0))
    sd.add_obs_array(X)
    sd.prepare_interp(m2)
    if prec == 1:
        sd.get_mixed_check()          # builds the mixed factors
    s = torch.cuda.Stream()
    for m in (40000, 7001):
        Xd = torch.tensor(Xe[:m], dtype=torch.float64, device="cuda")
        ref = torch.full((m,), float("nan"), dtype=torch.float64, device="cuda")
        got = torch.full((m,), float("nan"), dtype=torch.float64, device="cuda")
        check(_calls(sd, Xd, ref, prec, s, noalloc=False))
        s.synchronize()
        check(_calls(sd, Xd, got, prec, s, noalloc=True))
        s.synchronize()
        a, b = ref.cpu().numpy(), got.cpu().numpy()
        assert np.all(np.isfinite(b))
        assert np.array_equal(a, b), f"m={m}: max diff {np.max(np.abs(a - b))}"
    _say(f"[noalloc] {kind} prec={prec}: both calls bit-identical to the allocating eval")


def test_noalloc_refuses_oversized_batches():
    import torch

    import numcosmo_amd as na
    from numcosmo_amd import synth
    from numcosmo_amd._lib import lib

    tgt, X, m2, Xe = synth.workload("banana", 500, 10, 4, 2)
    sd = na.StatsDistVKDE.new(na.StatsDistKernelGauss.new(4))
    sd.add_obs_array(X)
    sd.prepare()
    cap = lib().nckde_noalloc_scratch_bytes()
    assert cap == 64 << 20
    # FP64 needs at least 16 bytes per point: this many can never fit; the
    # host check refuses before anything touches the (tiny) device buffers
    m = cap // 16 + 1
    Xd = torch.zeros((1, 4), dtype=torch.float64, device="cuda")
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    r = lib().nckde_eval_m2lnp_device_noalloc(sd._h, C.c_void_p(Xd.data_ptr()), m, C.c_void_p(out.data_ptr()),
                                              0, None)
    assert r == -1   # NCKDE_EINVAL
    # mixed before the mixed factors exist: a state error, not an allocation
    r = lib().nckde_eval_m2lnp_device_noalloc(sd._h, C.c_void_p(Xd.data_ptr()), 1, C.c_void_p(out.data_ptr()),
                                              1, None)
    assert r == -3   # NCKDE_ESTATE
