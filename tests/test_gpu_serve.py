"""Served requests (continuous batching) on the device engine.

Each request's result in a concurrent mix must equal the same request run
alone, in the same per-session order, on sessions built the same way:
bit-exact for eval_m2lnp (FP64 and mixed), sample and prepare_interp, since
every kernel is deterministic and no kernel depends on what else runs.
"""
import sys

import numpy as np
import pytest

from numcosmo_amd import synth

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _sessions(stream):
    from numcosmo_amd import StatsDistKernelST, StatsDistVKDE
    out = []
    for s, S in enumerate(stream.sessions):
        sd = StatsDistVKDE.new(StatsDistKernelST.new(S["d"], stream.nu))
        sd.set_over_smooth(1.0)
        X, m2 = stream.fit_sample(s)
        sd.add_obs_array(X)
        sd.prepare_interp(m2)
        out.append(sd)
    return out


def _plan(stream, precisions=(1,)):
    """Per request: (op, session, payload) with refit counters per session."""
    refits = {}
    plan = []
    for r in range(len(stream.op)):
        s = int(stream.session[r])
        op = stream.OPS[int(stream.op[r])]
        Xe, u, z, c2 = stream.pool(s)
        if op == "eval":
            prec = precisions[r % len(precisions)]
            plan.append((op, s, (Xe, prec)))
        elif op == "sample":
            plan.append((op, s, (u, z, c2)))
        else:
            refits[s] = refits.get(s, 0) + 1
            plan.append((op, s, stream.fit_sample(s, refits[s])))
    return plan


def _submit(srv, sds, item):
    op, s, p = item
    if op == "eval":
        return srv.submit_eval(sds[s], p[0], precision=p[1])
    if op == "sample":
        return srv.submit_sample(sds[s], *p)
    X, m2 = p
    return srv.submit_prepare_interp(sds[s], m2, X)


def _alone(sds, item):
    op, s, p = item
    sd = sds[s]
    if op == "eval":
        return sd.eval_m2lnp_batch(p[0], precision=p[1])
    if op == "sample":
        return sd.sample_batch(*p)
    X, m2 = p
    sd.reset()
    sd.add_obs_array(X)
    sd.prepare_interp(m2)
    w = np.full(len(m2), np.nan)
    pw = sd.peek_weights()
    w[:len(pw)] = pw
    return w


def _same(a, b):
    if isinstance(a, tuple):
        return all(_same(x, y) for x, y in zip(a, b))
    return np.array_equal(a, b, equal_nan=True)


def _stream(seed, n_req=40, n_sessions=6):
    return synth.ServedStream(n_req=n_req, rate=0, n_sessions=n_sessions, n_lo=300, n_hi=2500,
                              dims=(4, 8, 16, 32), seed=seed, m_factor=10)


def test_results_alone_equal_results_in_the_mix():
    from numcosmo_amd.serve import Server
    stream = _stream(7)
    plan = _plan(stream, precisions=(1, 0))
    _say(f"[serve] mix of {len(plan)} requests over {len(stream.sessions)} sessions, 4 slots")
    sds = _sessions(stream)
    with Server(slots=4, hbm_budget=8 << 30) as srv:
        tickets = [_submit(srv, sds, it) for it in plan]
        got = [srv.wait(t) for t in tickets]
    # some requests ran at the same time
    iv = sorted((i["t_start"], i["t_end"]) for _, i in got)
    overlaps = sum(1 for a, b in zip(iv, iv[1:]) if b[0] < a[1])
    _say(f"   overlapping request pairs: {overlaps}")
    assert overlaps > 0

    _say("[serve] replay alone")
    sds2 = _sessions(stream)
    for k, it in enumerate(plan):
        ref = _alone(sds2, it)
        assert _same(got[k][0], ref), f"request {k} ({it[0]}, session {it[1]}) differs from its run alone"


def _factors(sd):
    n = sd.get_n_kernels()
    return np.stack([sd.peek_cov_decomp(i) for i in range(n)])


def test_served_stream_at_full_range_against_the_oracle(oracle):
    """Sessions over the whole served range (n log-uniform on [500, 20000],
    d in {4, 8, 16, 32}), with one session per d at n >= 10^4: the mix equals
    the requests run alone, and every eval result (FP64 and mixed) is checked
    against the oracle holding the session's device factors and weights at
    that request, on a subset of its points."""
    from numcosmo_amd.serve import Server
    stream = synth.ServedStream(n_req=48, rate=0, n_sessions=8, n_lo=500, n_hi=20000, dims=(4, 8, 16, 32),
                                seed=21, m_factor=10)
    for s, (n, d) in enumerate([(10500, 4), (12000, 8), (16000, 16), (20000, 32)]):
        stream.sessions[s] = {"n": n, "d": d, "seed": 9000 + s}
    # every large session gets an eval in each precision
    for k, s in enumerate([0, 1, 2, 3]):
        stream.session[2 * k], stream.op[2 * k] = s, 0
        stream.session[2 * k + 1], stream.op[2 * k + 1] = s, 0
    plan = _plan(stream, precisions=(1, 0))
    _say(f"[serve C4] {len(plan)} requests over {len(stream.sessions)} sessions "
         f"(n {[S['n'] for S in stream.sessions]}, d {[S['d'] for S in stream.sessions]})")
    sds = _sessions(stream)
    with Server(slots=4, hbm_budget=64 << 30) as srv:
        tickets = [_submit(srv, sds, it) for it in plan]
        got = [srv.wait(t) for t in tickets]

    _say("[serve C4] replay alone, with the oracle on eval results")
    sds2 = _sessions(stream)
    cache = {}
    refits = {}
    worst = {0: 0.0, 1: 0.0}
    for k, it in enumerate(plan):
        ref = _alone(sds2, it)
        assert _same(got[k][0], ref), f"request {k} ({it[0]}, session {it[1]}) differs from its run alone"
        op, s, p = it
        if op == "prepare_interp":
            refits[s] = refits.get(s, 0) + 1
            continue
        if op != "eval":
            continue
        key = (s, refits.get(s, 0))
        if key not in cache:
            sd = sds2[s]
            od = oracle.StatsDist(oracle.VKDE, oracle.ST, sd.d, stream.nu)
            od.set_prepared_vkde(sd.peek_sample_array()[:sd.get_n_kernels()], _factors(sd), sd.peek_weights(),
                                 sd.get_href())
            cache = {key: od}
        sub = np.arange(0, len(p[0]), max(1, len(p[0]) // 64))
        oref = cache[key].eval_m2lnp_batch(p[0][sub])
        err = np.max(np.abs(ref[sub] - oref) / np.maximum(np.abs(oref), 1.0))
        worst[p[1]] = max(worst[p[1]], err)
        assert err <= (1e-6 if p[1] == 1 else 1e-12), (k, s, p[1], err)
    _say(f"   eval vs oracle: FP64 max rel err {worst[0]:.2e}, mixed {worst[1]:.2e}")


def test_admission_respects_the_device_budget():
    """With a budget below every request, requests run one at a time and
    still give the same results."""
    from numcosmo_amd.serve import Server
    stream = _stream(8, n_req=16, n_sessions=4)
    plan = _plan(stream)
    _say("[serve] budget of 1 byte: serial admission")
    sds = _sessions(stream)
    with Server(slots=4, hbm_budget=1) as srv:
        tickets = [_submit(srv, sds, it) for it in plan]
        got = [srv.wait(t) for t in tickets]
    iv = sorted((i["t_start"], i["t_end"]) for _, i in got)
    assert all(b[0] >= a[1] for a, b in zip(iv, iv[1:]))
    assert all(i["bytes"] > 0 for _, i in got)
    sds2 = _sessions(stream)
    for k, it in enumerate(plan):
        assert _same(got[k][0], _alone(sds2, it))


def test_errors_come_back_on_the_ticket():
    from numcosmo_amd import NckdeError, StatsDistKernelGauss, StatsDistVKDE
    from numcosmo_amd.serve import Server
    sd = StatsDistVKDE.new(StatsDistKernelGauss.new(3))
    with Server(slots=2) as srv:
        t = srv.submit_eval(sd, np.zeros((4, 3)))     # not prepared
        with pytest.raises(NckdeError):
            srv.wait(t)


def test_fused_multi_session_eval_is_bit_identical():
    """nckde_eval_m2lnp_multi: mixed evals of sessions that share the MFMA
    dimension and kernel type run as one launch per pass over a segment
    table; every result equals the session's own eval bit for bit, for
    groups of one and several sessions, mixed dimensions and kernels, and
    in FP64 (run one by one)."""
    import numcosmo_amd as na
    from numcosmo_amd.ncm_stats_dist import eval_m2lnp_multi

    g = np.random.default_rng(77)
    specs = [(8, 3.0, 1200), (8, 3.0, 2500), (8, 3.0, 700), (16, 1.0, 1800), (16, 3.0, 900), (4, 0.0, 600),
             (32, 5.0, 1000)]
    sds, Xs = [], []
    for d, nu, n in specs:
        X = g.standard_normal((n, d))
        k = na.StatsDistKernelGauss.new(d) if nu == 0.0 else na.StatsDistKernelST.new(d, nu)
        sd = na.StatsDistVKDE.new(k)
        sd.add_obs_array(X)
        sd.prepare()
        sds.append(sd)
        Xs.append(g.standard_normal((int(g.integers(500, 9000)), d)))
    for prec in (1, 0):
        fused = eval_m2lnp_multi(sds, Xs, prec)
        for sd, X, got in zip(sds, Xs, fused):
            assert np.array_equal(got, sd.eval_m2lnp_batch(X, precision=prec))
    one = eval_m2lnp_multi(sds[:1], Xs[:1], 1)[0]
    assert np.array_equal(one, sds[0].eval_m2lnp_batch(Xs[0], precision=1))


def test_server_fuses_waiting_evals():
    """A worker that takes a mixed eval also takes the waiting mixed evals
    of other idle sessions with the same MFMA dimension and kernel type and
    runs them as one fused launch per pass: with one slot held by a refit,
    the evals queued behind it run together, and each result equals the
    session's own eval bit for bit.  An FP64 eval never joins a batch."""
    import numcosmo_amd as na
    from numcosmo_amd.serve import Server

    g = np.random.default_rng(91)
    sds, Xs, alone = [], [], []
    for s in range(6):
        n = int(g.integers(600, 1200))
        sd = na.StatsDistVKDE.new(na.StatsDistKernelST.new(8, 3.0))
        sd.add_obs_array(g.standard_normal((n, 8)))
        sd.prepare()
        X = g.standard_normal((int(g.integers(2000, 6000)), 8))
        alone.append(sd.eval_m2lnp_batch(X, precision=1))      # also builds the mixed factors
        sds.append(sd)
        Xs.append(X)
    X64 = g.standard_normal((1000, 8))
    alone64 = sds[0].eval_m2lnp_batch(X64, precision=0)
    big = na.StatsDistVKDE.new(na.StatsDistKernelST.new(8, 3.0))
    Xb = g.standard_normal((3000, 8))
    big.add_obs_array(Xb)
    big.prepare()
    m2b = np.einsum("ij,ij->i", Xb, Xb)
    with Server(slots=1) as srv:
        t_big = srv.submit_prepare_interp(big, m2b)
        t64 = srv.submit_eval(sds[0], X64, precision=0)
        ts = [srv.submit_eval(sd, X, precision=1) for sd, X in zip(sds, Xs)]
        srv.wait(t_big)
        got64, i64 = srv.wait(t64)
        res = [srv.wait(t) for t in ts]
    batches = [info["batch"] for _, info in res]
    _say(f"   fused batch sizes {batches}, FP64 eval batch {i64['batch']}")
    assert i64["batch"] == 1
    assert np.array_equal(got64, alone64)
    assert max(batches) > 1
    for (got, info), a in zip(res, alone):
        assert np.array_equal(got, a)
