"""Served requests over prepared kernel-density sessions (continuous batching).

A ``Server`` owns native worker threads (``nckde_server_*`` in
``include/nckde.h``).  Each session is a ``StatsDistKDE``/``StatsDistVKDE``
object whose factors stay in HBM; requests (eval_m2lnp, sample,
prepare_interp) are queued with ``submit_*`` and return a ``Ticket`` at once.
Requests on different sessions run concurrently, each on its session's HIP
stream; requests on one session run in submission order, so every result is
the same as the same request run alone after the same earlier requests on
its session.

There is no counterpart in the reference, where each caller evaluates its own
NcmStatsDist object (ncm_stats_dist.h:122-126); the per-request work is
exactly the corresponding NcmStatsDist call.
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass, field

import numpy as np

from ._lib import NckdeError, Request, RequestInfo, lib
from .ncm_stats_dist import Precision, _StatsDist

__all__ = ["Op", "Server", "Ticket"]


class Op(enum.IntEnum):
    EVAL_M2LNP = 0
    SAMPLE = 1
    PREPARE_INTERP = 2


@dataclass
class Ticket:
    id: int
    op: Op
    sd: _StatsDist
    m: int
    result: object = None
    keep: list = field(default_factory=list)   # buffers alive until the wait


def _ptr(a):
    return None if a is None else a.ctypes.data


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Server:
    def __init__(self, slots: int = 8, hbm_budget: int = 64 << 30):
        h = C.c_void_p()
        st = lib().nckde_server_create(C.byref(h), slots, hbm_budget)
        if st != 0:
            raise NckdeError(st, lib().nckde_last_error().decode(errors="replace"))
        self._h = h
        self.slots = slots
        self.hbm_budget = hbm_budget

    def close(self):
        if getattr(self, "_h", None):
            lib().nckde_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()

    def clock(self) -> float:
        return lib().nckde_server_clock(self._h)

    def _submit(self, req: Request, t: Ticket) -> Ticket:
        tid = C.c_uint64()
        st = lib().nckde_server_submit(self._h, C.byref(req), C.byref(tid))
        if st != 0:
            raise NckdeError(st, lib().nckde_last_error().decode(errors="replace"))
        t.id = tid.value
        return t

    def request_bytes(self, op: Op, sd: _StatsDist, m: int, precision=Precision.MIXED) -> int:
        req = Request(op=int(op), h=sd._h, m=m, precision=int(precision))
        b = C.c_size_t()
        st = lib().nckde_request_bytes(C.byref(req), C.byref(b))
        if st != 0:
            raise NckdeError(st, lib().nckde_last_error().decode(errors="replace"))
        return b.value

    def submit_eval(self, sd: _StatsDist, X, precision=Precision.MIXED) -> Ticket:
        X = _f64(np.atleast_2d(X))
        if X.shape[1] != sd.d:
            raise ValueError("submit_eval: points must have d columns")
        out = np.empty(X.shape[0])
        req = Request(op=int(Op.EVAL_M2LNP), h=sd._h, m=X.shape[0], X=_ptr(X), out=_ptr(out),
                      precision=int(precision))
        return self._submit(req, Ticket(0, Op.EVAL_M2LNP, sd, X.shape[0], out, [X, out]))

    def submit_sample(self, sd: _StatsDist, u, z, chi2=None) -> Ticket:
        u = _f64(u)
        m = u.shape[0]
        z = _f64(z).reshape(m, sd.d)
        c2 = None if chi2 is None else _f64(chi2)
        out = np.empty((m, sd.d))
        idx = np.empty(m, dtype=np.uint32)
        req = Request(op=int(Op.SAMPLE), h=sd._h, m=m, u=_ptr(u), z=_ptr(z), chi2=_ptr(c2), out=_ptr(out),
                      idx=_ptr(idx))
        return self._submit(req, Ticket(0, Op.SAMPLE, sd, m, (out, idx), [u, z, c2, out, idx]))

    def submit_prepare_interp(self, sd: _StatsDist, m2lnp, X=None) -> Ticket:
        m2lnp = _f64(m2lnp)
        if X is not None:
            X = _f64(X)
            if X.shape != (m2lnp.shape[0], sd.d):
                raise ValueError("submit_prepare_interp: X must be len(m2lnp) x d")
        w = np.empty(m2lnp.shape[0])
        req = Request(op=int(Op.PREPARE_INTERP), h=sd._h, m=m2lnp.shape[0], X=_ptr(X), m2lnp=_ptr(m2lnp),
                      out=_ptr(w))
        return self._submit(req, Ticket(0, Op.PREPARE_INTERP, sd, m2lnp.shape[0], w, [X, m2lnp, w]))

    def wait(self, t: Ticket):
        """Block until the request is done; returns (result, info dict).

        result: eval -> m2lnp array; sample -> (points, kernel indices);
        prepare_interp -> the fitted weights (NaN beyond the kernels kept by
        the range cut)."""
        info = RequestInfo()
        st = lib().nckde_server_wait(self._h, t.id, C.byref(info))
        t.keep = []
        if st != 0:
            raise NckdeError(st, info.message.decode(errors="replace"))
        return t.result, {"t_submit": info.t_submit, "t_start": info.t_start, "t_end": info.t_end,
                     "latency": info.t_end - info.t_submit, "bytes": info.bytes, "batch": info.batch}
