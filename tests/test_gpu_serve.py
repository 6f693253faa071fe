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
