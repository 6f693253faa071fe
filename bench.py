#!/usr/bin/env python
"""Kernel-pair evaluations per second of VKDE eval_m2lnp on one MI355X.

Headline workload (one step = one eval_m2lnp pass over the evaluation points):
  VKDE, Student-t kernel nu=3, d=16, n=20000 kernels with NNLS weights
  (prepare_interp on the exact m2lnp of a seeded correlated Gaussian target),
  eval_m2lnp at m=200000 points already resident in HBM: 4e9 pairs per step.
  This is where the mixed-precision path is required to beat the plain FP64
  path by 4x at relative m2lnp error <= 1e-6; both are timed, the mixed one
  is `value`.  The mixed result is checked against the FP64 oracle on 1e4
  points and against the device FP64 result on all 2e5.

Also measured: the banana VKDE at d=8, n=4000, m=1e5; the large VKDE
(d=64, n=50000, nu=5, condition number 1e4, m=1e6) checked against the
oracle on 1e4 points; 200 APES ensemble steps (d=16, 2x3200 walkers); the
cross-validated bandwidth fits; and a served stream (512 requests, Poisson
arrivals, eval/sample/prepare_interp 60/30/10 over VKDE sessions with n
log-uniform on [500, 20000], d in {4, 8, 16, 32}, m = 10 n) whose p50/p99
request latency goes into "latency".

Timing: W warm-up steps, then K steps bracketed by torch.cuda.synchronize();
the dominant kernel's duration comes from HIP events recorded on the stream
the engine launches on.  The CPU baseline is the oracle (a restatement of
the reference algorithm) built -O3 for AVX2/FMA, timed on a bounded sample
of the same workload at all cores and at one core.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"f64": 78.6, "f16": 2500.0}  # MI355X dense TFLOP/s (FP64 vector=matrix, FP16 MFMA)


def flop_per_pair(d: int) -> int:
    # whitening product d^2 MAC, subtract, square-accumulate
    return 2 * d * d + 3 * d


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed_eval(sd, Xd, outd, prec, steps, warmup, stream):
    """Returns (event ms per step, wall ms per step)."""
    import torch

    from numcosmo_amd._lib import check, lib

    L = lib()
    m = Xd.shape[0]
    sptr = C.c_void_p(stream.cuda_stream)

    def step():
        check(L.nckde_eval_m2lnp_device(sd._h, C.c_void_p(Xd.data_ptr()), m, C.c_void_p(outd.data_ptr()),
                                        int(prec), sptr))

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    step_ms = e0.elapsed_time(e1) / steps

    # the dominant kernel alone: HIP events the engine records around its
    # main evaluation kernel on the same stream, one step at a time
    check(L.nckde_set_profiling(sd._h, 1))
    kms = []
    f = C.c_float()
    for _ in range(steps):
        step()
        check(L.nckde_get_kernel_ms(sd._h, C.byref(f)))
        kms.append(f.value)
    check(L.nckde_set_profiling(sd._h, 0))
    torch.cuda.synchronize()
    return step_ms, wall, float(np.mean(kms))


def rel_err(got, ref):
    e = np.abs(got - ref) / np.maximum(np.abs(ref), 1.0)
    return float(e.max()), float(np.quantile(e, 0.99))


def supports(sd, prec, Xd, outd):
    from numcosmo_amd._lib import lib
    r = lib().nckde_eval_m2lnp_device(sd._h, C.c_void_p(Xd.data_ptr()), 1, C.c_void_p(outd.data_ptr()),
                                      int(prec), None)
    import torch
    torch.cuda.synchronize()
    return r == 0


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(oracle_mod, d, nu, X, w, Xe, n, threads, target_s):
    """FP64 eval_m2lnp of the CPU-baseline build of the restatement (-O3,
    FMA, AVX2) on `threads` OpenMP threads, over a bounded sample of the
    points sized to about target_s seconds; returns (pairs/s, sample text)."""
    od = oracle_mod.StatsDist(oracle_mod.VKDE, oracle_mod.ST, d, nu, fast=True)
    od.set_use_threads(1)
    od.add_obs_array(X)
    od.prepare()
    od.set_weights(w)
    oracle_mod.set_num_threads(threads)
    try:
        pts = max(threads, 8)
        while True:
            t0 = time.perf_counter()
            od.eval_m2lnp_batch(Xe[:pts])
            dt = time.perf_counter() - t0
            if dt >= 0.5 or pts >= len(Xe):
                break
            pts = min(len(Xe), pts * 4)
        pts = int(min(len(Xe), max(pts, pts * target_s / max(dt, 1e-3))))
        t0 = time.perf_counter()
        od.eval_m2lnp_batch(Xe[:pts])
        dt = time.perf_counter() - t0
    finally:
        oracle_mod.set_num_threads(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    return pts * n / dt, f"{pts} points x {n} kernels, {dt:.1f} s"


def run_case(name, kind_target, d, n, m, nu, steps, warmup, fit, check_pts, config_id, oracle_mod, cores):
    import torch

    import numcosmo_amd as na
    from numcosmo_amd import synth

    log(f"[{name}] generating d={d} n={n} m={m}")
    tgt, X, m2lnp, Xe = synth.workload(kind_target, n, m, d, config_id)
    sd = na.StatsDistVKDE.new(na.StatsDistKernelST.new(d, nu), na.StatsDistCV.NONE)
    sd.add_obs_array(X)
    t0 = time.perf_counter()
    if fit:
        sd.prepare_interp(m2lnp)
    else:
        sd.prepare()
    prep_s = time.perf_counter() - t0
    log(f"[{name}] {'prepare_interp' if fit else 'prepare'} {prep_s:.3f} s")

    Xd = torch.tensor(Xe, dtype=torch.float64, device="cuda")
    outd = torch.empty(m, dtype=torch.float64, device="cuda")
    stream = torch.cuda.Stream()
    res = {"prepare_s": prep_s, "n": n, "m": m, "d": d}
    w = sd.peek_weights()

    # FP64 oracle (the parity build, all threads) on the first check_pts
    # points, with the device-fitted weights
    od = oracle_mod.StatsDist(oracle_mod.VKDE, oracle_mod.ST, d, nu)
    od.set_use_threads(1)
    od.add_obs_array(X)
    t0 = time.perf_counter()
    od.prepare()
    log(f"[{name}] oracle prepare {time.perf_counter() - t0:.1f} s")
    od.set_weights(w)
    sub = Xe[:check_pts]
    t0 = time.perf_counter()
    ref = od.eval_m2lnp_batch(sub)
    log(f"[{name}] oracle eval of {check_pts} points {time.perf_counter() - t0:.1f} s")
    del od

    # CPU baseline: the -O3 build of the same restatement, all cores and one
    cpu_all, sample_all = cpu_baseline(oracle_mod, d, nu, X, w, Xe, n, cores, 10.0)
    cpu_one, sample_one = cpu_baseline(oracle_mod, d, nu, X, w, Xe, n, 1, 10.0)
    res["cpu"] = {"pairs_per_s": cpu_all, "cores": cores, "sample": sample_all,
                  "pairs_per_s_1core": cpu_one, "sample_1core": sample_one, "cpu_model": cpu_model(),
                  "build": "oracle restatement, gcc -O3 -march=x86-64-v3 -ffp-contract=fast, OpenMP over points"}
    log(f"[{name}] cpu baseline {cpu_all:.3e} pairs/s on {cores} threads, {cpu_one:.3e} on 1 ({cpu_model()})")

    full = {}
    for prec, key in ((na.Precision.FP64, "fp64"), (na.Precision.MIXED, "mixed")):
        if not supports(sd, prec, Xd, outd):
            log(f"[{name}] {key}: not available")
            continue
        ms, wall, kms = timed_eval(sd, Xd, outd, prec, steps, warmup, stream)
        out_all = outd.cpu().numpy()
        full[key] = out_all
        emax, e99 = rel_err(out_all[:check_pts], ref)
        pairs = n * m
        res[key] = {"ms_per_step": ms, "wall_ms_per_step": wall, "pairs_per_s": pairs / (ms * 1e-3),
                    "rel_err_max": emax, "rel_err_p99": e99, "checked_points": check_pts, "kernel_ms": kms,
                    "tflops": pairs * flop_per_pair(d) / (ms * 1e-3) / 1e12,
                    "kernel_tflops": pairs * flop_per_pair(d) / (kms * 1e-3) / 1e12}
        log(f"[{name}] {key}: {ms:.3f} ms/step (main kernel {kms:.3f} ms), {pairs / (ms * 1e-3):.3e} pairs/s, "
            f"rel err vs oracle on {check_pts} points: max {emax:.2e} p99 {e99:.2e}")
    if "fp64" in full and "mixed" in full:
        emax, e99 = rel_err(full["mixed"], full["fp64"])
        res["mixed"]["vs_fp64_all_points"] = {"points": m, "rel_err_max": emax, "rel_err_p99": e99}
        res["mixed_over_fp64"] = res["fp64"]["ms_per_step"] / res["mixed"]["ms_per_step"]
        log(f"[{name}] mixed vs GPU FP64 on all {m} points: max {emax:.2e}; mixed is "
            f"{res['mixed_over_fp64']:.1f}x FP64")
    return res


def serve_case(n_req, rate, n_sessions, slots, seed=None):
    """Served stream: Poisson arrivals of eval / sample / prepare_interp
    requests (60/30/10) over VKDE sessions with n log-uniform on [500, 20000]
    and d in {4, 8, 16, 32}, m = 10 n points or draws per request; p50/p99 of
    the server-side latency (submit to results in host memory)."""
    import threading

    import numcosmo_amd as na
    from numcosmo_amd import synth
    from numcosmo_amd.serve import Server

    stream = synth.ServedStream(n_req=n_req, rate=rate, n_sessions=n_sessions, seed=seed)
    log(f"[serve] {n_sessions} sessions, {n_req} requests at {rate}/s, {slots} slots")
    t0 = time.perf_counter()
    sds = []
    for s, S in enumerate(stream.sessions):
        sd = na.StatsDistVKDE.new(na.StatsDistKernelST.new(S["d"], stream.nu), na.StatsDistCV.NONE)
        X, m2 = stream.fit_sample(s)
        sd.add_obs_array(X)
        sd.prepare_interp(m2)
        sds.append(sd)
    setup_s = time.perf_counter() - t0
    log(f"[serve] sessions fitted in {setup_s:.1f} s")
    # payloads (generated before the stream starts: a request's inputs exist at its arrival)
    refits = {}
    payload = []
    for r in range(n_req):
        s = int(stream.session[r])
        op = stream.OPS[int(stream.op[r])]
        if op == "prepare_interp":
            refits[s] = refits.get(s, 0) + 1
            payload.append(stream.fit_sample(s, refits[s]))
        else:
            payload.append(stream.pool(s))
    # warm-up: one eval and one sample per session (device work buffers)
    for s, sd in enumerate(sds):
        Xe, u, z, c2 = stream.pool(s)
        sd.eval_m2lnp_batch(Xe, precision=na.Precision.MIXED)
        sd.sample_batch(u, z, c2)
    log(f"[serve] payloads and warm-up done")

    srv = Server(slots=slots, hbm_budget=64 << 30)
    tickets = [None] * n_req
    lateness = np.zeros(n_req)

    def submitter():
        base = srv.clock()
        for r in range(n_req):
            target = base + stream.arrival[r]
            while True:
                dt = target - srv.clock()
                if dt <= 0:
                    break
                time.sleep(min(dt, 0.002))
            s = int(stream.session[r])
            op = stream.OPS[int(stream.op[r])]
            if op == "eval":
                tickets[r] = srv.submit_eval(sds[s], payload[r][0], precision=na.Precision.MIXED)
            elif op == "sample":
                _, u, z, c2 = payload[r]
                tickets[r] = srv.submit_sample(sds[s], u, z, c2)
            else:
                X, m2 = payload[r]
                tickets[r] = srv.submit_prepare_interp(sds[s], m2, X)
            lateness[r] = srv.clock() - target

    th = threading.Thread(target=submitter)
    th.start()
    infos = []
    for r in range(n_req):
        while tickets[r] is None:
            time.sleep(0.0005)
        _, info = srv.wait(tickets[r])
        infos.append(info)
    th.join()
    srv.close()

    lat = np.array([i["latency"] for i in infos]) * 1e3
    t_first = min(i["t_submit"] for i in infos)
    t_last = max(i["t_end"] for i in infos)
    ops = np.array([stream.OPS[int(o)] for o in stream.op])
    pairs = sum(10 * stream.sessions[int(stream.session[r])]["n"] ** 2 for r in range(n_req) if ops[r] == "eval")
    res = {"requests": n_req, "rate_per_s": rate, "sessions": n_sessions, "slots": slots,
           "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
           "mean_ms": float(lat.mean()), "max_ms": float(lat.max()),
           "requests_per_s": n_req / (t_last - t_first), "eval_pairs_per_s": pairs / (t_last - t_first),
           "submit_lateness_p99_ms": float(np.percentile(lateness, 99) * 1e3), "setup_s": setup_s}
    wait = np.array([i["t_start"] - i["t_submit"] for i in infos]) * 1e3
    run = np.array([i["t_end"] - i["t_start"] for i in infos]) * 1e3
    sizes = np.array([stream.sessions[int(stream.session[r])]["n"] for r in range(n_req)])
    dims = np.array([stream.sessions[int(stream.session[r])]["d"] for r in range(n_req)])
    for op in stream.OPS:
        sel = ops == op
        if sel.any():
            res[op] = {"count": int(sel.sum()), "p50_ms": float(np.percentile(lat[sel], 50)),
                       "p99_ms": float(np.percentile(lat[sel], 99)),
                       "queue_p99_ms": float(np.percentile(wait[sel], 99)),
                       "run_p50_ms": float(np.percentile(run[sel], 50)),
                       "run_p99_ms": float(np.percentile(run[sel], 99))}
    # the slowest requests: what they were and where their time went
    worst = np.argsort(lat)[-8:][::-1]
    res["slowest"] = [{"op": str(ops[r]), "n": int(sizes[r]), "d": int(dims[r]), "latency_ms": float(lat[r]),
                       "queue_ms": float(wait[r]), "run_ms": float(run[r])} for r in worst]
    log(f"[serve] p50 {res['p50_ms']:.2f} ms p99 {res['p99_ms']:.2f} ms, {res['requests_per_s']:.1f} req/s, "
        f"per op {[(op, res[op]) for op in stream.OPS if op in res]}")
    return res


def apes_case(steps, precision):
    """APES ensemble steps: d=16, 2 x 3200 walkers on a seeded 4-component
    Gaussian mixture (equal weights, means U[-5,5]^16, random SPD
    covariances), VKDE with the Cauchy kernel, over_smooth 1; each half-step
    fits a VKDE of the other half (prepare_interp), samples 3200 proposals and
    evaluates both transition densities in one batch."""
    import numcosmo_amd as na
    from numcosmo_amd import synth
    from numcosmo_amd.apes import APES, InterpolationKernel

    d, nw = 16, 6400
    g = synth.rng(3)
    tgt = synth.Mixture(d, g)
    x0 = tgt.draw(g, nw)
    s = APES(nwalkers=nw, ndim=d, log_prob_batch=lambda th: -0.5 * tgt.m2lnp(th),
             interpolation_kernel=InterpolationKernel.CAUCHY, over_smooth=1.0, precision=int(precision), seed=7)
    log(f"[apes] d={d} walkers={nw} steps={steps} precision={int(precision)}")
    t0 = time.perf_counter()
    s.run_mcmc(x0, steps)
    wall = time.perf_counter() - t0
    st = s.get_stats()
    pairs = steps * 2 * (2 * (nw // 2) * (nw // 2))
    res = {"d": d, "walkers": nw, "steps": steps, "precision": "mixed" if int(precision) else "fp64",
           "s_per_step": wall / steps, "acceptance": st["acceptance"],
           "seconds": {k: v / steps for k, v in st["seconds"].items()},
           "transition_pairs_per_s": pairs / max(st["seconds"]["eval"], 1e-9)}
    log(f"[apes] {wall / steps * 1e3:.1f} ms/step, acceptance {st['acceptance']:.3f}, per step "
        f"{ {k: round(v * 1e3, 2) for k, v in res['seconds'].items()} } ms")
    return res


def c5_case(m, m_fp64, check_pts, oracle_mod):
    """Large VKDE: d=64, n=50000, Student-t nu=5, Gaussian target with
    condition number 1e4; prepare_interp (50000^2 interpolation matrix, NNLS
    weights) then eval_m2lnp at m points in mixed precision and at the first
    m_fp64 of them in FP64.  Both are checked against the FP64 oracle
    evaluating the device's factors and weights on the first check_pts
    points, and the mixed result against the device FP64 one on m_fp64."""
    import torch

    import numcosmo_amd as na
    from numcosmo_amd import synth

    d, n = 64, 50000
    log(f"[C5] generating d={d} n={n} m={m}")
    tgt, X, m2lnp, Xe = synth.workload("cond", n, m, d, 5)
    sd = na.StatsDistVKDE.new(na.StatsDistKernelST.new(d, 5.0), na.StatsDistCV.NONE)
    sd.set_nnls_refine(2)       # the workload fits its NNLS weights with FP64 refinement
    sd.add_obs_array(X)
    t0 = time.perf_counter()
    sd.prepare_interp(m2lnp)
    prep = time.perf_counter() - t0
    log(f"[C5] prepare_interp {prep:.1f} s, {sd.get_n_kernels()} kernels, NNLS rungs {sd.get_nnls_rungs()}")
    Xd = torch.tensor(Xe, dtype=torch.float64, device="cuda")
    outm = torch.empty(m, dtype=torch.float64, device="cuda")
    out64 = torch.empty(m_fp64, dtype=torch.float64, device="cuda")
    stream = torch.cuda.Stream()
    nk = sd.get_n_kernels()
    ms_m, _, k_m = timed_eval(sd, Xd, outm, na.Precision.MIXED, 3, 1, stream)
    X64 = Xd[:m_fp64].contiguous()
    ms_f, _, k_f = timed_eval(sd, X64, out64, na.Precision.FP64, 1, 1, stream)
    got_m, got_f = outm.cpu().numpy(), out64.cpu().numpy()
    ms_32, _, k_32 = timed_eval(sd, X64, out64, na.Precision.FP32, 1, 1, stream)
    got_32 = out64.cpu().numpy()
    emax, e99 = rel_err(got_m[:m_fp64], got_f)

    # the oracle holding the device's factors and weights (lnnorms recomputed)
    U = np.stack([sd.peek_cov_decomp(i) for i in range(nk)])
    od = oracle_mod.StatsDist(oracle_mod.VKDE, oracle_mod.ST, d, 5.0)
    od.set_prepared_vkde(X, U, sd.peek_weights(), sd.get_href())
    del U
    t0 = time.perf_counter()
    ref = od.eval_m2lnp_batch(Xe[:check_pts])
    t_or = time.perf_counter() - t0
    om_max, om_99 = rel_err(got_m[:check_pts], ref)
    of_max, of_99 = rel_err(got_f[:check_pts], ref)
    o32_max, o32_99 = rel_err(got_32[:check_pts], ref)
    res = {"d": d, "n": n, "m": m, "prepare_interp_s": prep, "nnls_rungs": list(sd.get_nnls_rungs()),
           "mixed": {"ms_per_step": ms_m, "pairs_per_s": nk * m / (ms_m * 1e-3), "kernel_ms": k_m,
                     "rel_err_vs_oracle_max": om_max, "rel_err_vs_oracle_p99": om_99},
           "fp32": {"m": m_fp64, "ms_per_step": ms_32, "pairs_per_s": nk * m_fp64 / (ms_32 * 1e-3),
                    "kernel_ms": k_32, "rel_err_vs_oracle_max": o32_max, "rel_err_vs_oracle_p99": o32_99},
           "fp64": {"m": m_fp64, "ms_per_step": ms_f, "pairs_per_s": nk * m_fp64 / (ms_f * 1e-3), "kernel_ms": k_f,
                    "rel_err_vs_oracle_max": of_max, "rel_err_vs_oracle_p99": of_99},
           "oracle_checked_points": check_pts, "oracle_s": t_or,
           "mixed_vs_fp64_rel_err_max": emax, "mixed_vs_fp64_rel_err_p99": e99}
    log(f"[C5] mixed {ms_m:.1f} ms ({res['mixed']['pairs_per_s']:.3e} pairs/s), fp32 {ms_32:.1f} ms and fp64 "
        f"{ms_f:.1f} ms for {m_fp64} points ({res['fp32']['pairs_per_s']:.3e}, {res['fp64']['pairs_per_s']:.3e} "
        f"pairs/s); vs oracle on {check_pts} points ({t_or:.0f} s): mixed {om_max:.2e}, fp32 {o32_max:.2e}, "
        f"fp64 {of_max:.2e}; mixed vs fp64 {emax:.2e}")
    return res


def cv_case(oracle_mod):
    """Cross-validated bandwidths on VKDE Student-t nu=3, d=8, banana, n=2000
    (local_frac 0.05): CV_LOO (Nelder-Mead on the Monte-Carlo AMISE),
    CV_SPLIT_NOFIT (Nelder-Mead on the held-out m2lnp) and CV_SPLIT (restarts
    and levmar on NNLS fits), each as one prepare_interp on the device.  The
    oracle times one LOO objective and one SPLIT_NOFIT objective at
    over_smooth 1 on one host core (the reference evaluates them serially)."""
    import numcosmo_amd as na
    from numcosmo_amd import synth

    d, n = 8, 2000
    tgt, X, m2, _ = synth.workload("banana", n, 10, d, 2)
    res = {"workload": f"VKDE Student-t nu=3, d={d}, banana, n={n}"}
    for cv in ("LOO", "SPLIT_NOFIT", "SPLIT"):
        sd = na.StatsDistVKDE.new(na.StatsDistKernelST.new(d, 3.0), getattr(na.StatsDistCV, cv))
        sd.add_obs_array(X)
        t0 = time.perf_counter()
        sd.prepare_interp(m2)
        wall = time.perf_counter() - t0
        nobj, secs = sd.get_cv_stats()
        res[cv] = {"prepare_interp_s": wall, "fit_s": secs, "objectives": nobj,
                   "s_per_objective": secs / max(nobj, 1), "over_smooth": sd.get_over_smooth()}
        log(f"[cv] {cv}: prepare_interp {wall:.2f} s, fit {secs:.2f} s over {nobj} objectives, "
            f"over_smooth {sd.get_over_smooth():.6f}")
    if oracle_mod is not None:
        od = oracle_mod.StatsDist(oracle_mod.VKDE, oracle_mod.ST, d, 3.0)
        od.add_obs_array(X)
        od.prepare()
        for cv, key in ((oracle_mod.CV_LOO, "LOO"), (oracle_mod.CV_SPLIT_NOFIT, "SPLIT_NOFIT")):
            if cv == oracle_mod.CV_SPLIT_NOFIT:
                od = oracle_mod.StatsDist(oracle_mod.VKDE, oracle_mod.ST, d, 3.0, cv)
                od.add_obs_array(X)
                od.prepare()
            od.set_cv_type(cv)
            t0 = time.perf_counter()
            od.cv_objective(0.0)
            res[key]["cpu_s_per_objective"] = time.perf_counter() - t0
            log(f"[cv] oracle {key} objective {res[key]['cpu_s_per_objective']:.3f} s (1 core)")
        res["cpu_cores"] = 1
    return res


def pmc_probe(case, steps):
    """Short run for hardware-counter passes: one workload's mixed and FP64
    evals (case "d16": d=16, n=20000, m=2e5, the headline; "c2": d=8,
    n=4000, m=1e5), then a plain device copy of 256 MiB (the calibration
    kernel of the FETCH_SIZE / WRITE_SIZE counters)."""
    import torch

    import numcosmo_amd as na
    from numcosmo_amd import synth

    if case == "d16":
        d, n, m, target, cid = 16, 20000, 200000, "gaussian", 4
    else:
        d, n, m, target, cid = 8, 4000, 100000, "banana", 2
    tgt, X, m2lnp, Xe = synth.workload(target, n, m, d, cid)
    sd = na.StatsDistVKDE.new(na.StatsDistKernelST.new(d, 3.0), na.StatsDistCV.NONE)
    sd.add_obs_array(X)
    sd.prepare_interp(m2lnp)
    Xd = torch.tensor(Xe, dtype=torch.float64, device="cuda")
    outd = torch.empty(Xe.shape[0], dtype=torch.float64, device="cuda")
    stream = torch.cuda.Stream()
    for prec in (na.Precision.MIXED, na.Precision.FP64):
        ms, _, kms = timed_eval(sd, Xd, outd, prec, steps, 1, stream)
        log(f"pmc probe {case}: precision {int(prec)} {ms:.3f} ms/step, main kernel {kms:.3f} ms")
    src = torch.empty(32 << 20, dtype=torch.float64, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    log("pmc probe done: copy kernel moves 256 MiB read + 256 MiB written per launch")


def prep_probe(case, steps):
    """Short run for kernel traces of prepare_interp: "apes" fits the VKDE of
    one APES half-ensemble (Cauchy kernel, d=16, n=3200, the seeded mixture),
    "d4" the largest served refit (Student-t nu=3, d=4, n=20000, seeded
    correlated Gaussian) and "c5" the large VKDE (Student-t nu=5, d=64,
    n=50000, Gaussian target of condition number 1e4, two refinement steps),
    each `steps` times on one handle after a warm-up."""
    import torch

    import numcosmo_amd as na
    from numcosmo_amd import synth

    if case == "apes":
        tgt, X, m2lnp, _ = synth.workload("mixture", 3200, 10, 16, 3)
        kern = na.StatsDistKernelST.new(16, 1.0)
    elif case == "c5":
        tgt, X, m2lnp, _ = synth.workload("cond", 50000, 10, 64, 5)
        kern = na.StatsDistKernelST.new(64, 5.0)
    else:
        tgt, X, m2lnp, _ = synth.workload("gaussian", 20000, 10, 4, 4)
        kern = na.StatsDistKernelST.new(4, 3.0)
    sd = na.StatsDistVKDE.new(kern, na.StatsDistCV.NONE)
    if case == "c5":
        sd.set_nnls_refine(2)
    sd.add_obs_array(X)
    sd.prepare_interp(m2lnp)
    for r in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sd.prepare_interp(m2lnp)
        torch.cuda.synchronize()
        log(f"prep probe {case}: prepare_interp {(time.perf_counter() - t0) * 1e3:.2f} ms, NNLS rungs "
            f"{sd.get_nnls_rungs()}")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-c2", action="store_true", help="skip the d=8, n=4000 banana VKDE eval")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--prep-probe", choices=["apes", "d4", "c5"], default=None,
                    help="short run for kernel traces of prepare_interp")
    ap.add_argument("--pmc-probe", choices=["d16", "c2"], default=None,
                    help="short run for rocprofv3 --pmc passes")
    ap.add_argument("--apes-steps", type=int, default=200, help="APES ensemble steps (0: skip)")
    ap.add_argument("--no-c5", action="store_true", help="skip the d=64, n=50000 VKDE")
    ap.add_argument("--c5-check", type=int, default=10000, help="C5 points checked against the oracle")
    ap.add_argument("--no-cv", action="store_true", help="skip the cross-validated bandwidth fits")
    ap.add_argument("--only-c5", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=512)
    ap.add_argument("--serve-rate", type=float, default=10.0)
    ap.add_argument("--serve-sessions", type=int, default=48)
    ap.add_argument("--serve-slots", type=int, default=8)
    args = ap.parse_args(argv)

    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        return 2

    import torch

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.pmc_probe:
        pmc_probe(args.pmc_probe, max(2, min(args.steps, 5)))
        return 0
    if args.prep_probe:
        prep_probe(args.prep_probe, max(1, min(args.steps, 10)))
        return 0

    from oracle import oracle as O

    if args.only_c5:
        print(json.dumps({"c5": c5_case(1000000, 100000, args.c5_check, O)}))
        return 0

    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))

    # the headline: VKDE eval_m2lnp at d=16, n=20000, m=2e5, mixed vs FP64
    head = run_case("d16", "gaussian", 16, 20000, 200000, 3.0, args.steps, args.warmup, True, 10000, 4, O, cores)
    c2 = None
    if not args.no_c2:
        c2 = run_case("C2", "banana", 8, 4000, 100000, 3.0, args.steps, args.warmup, True, 4000, 2, O, cores)

    c5 = None if args.no_c5 else c5_case(1000000, 100000, args.c5_check, O)

    apes = None
    if args.apes_steps > 0:
        import numcosmo_amd as na
        apes = apes_case(args.apes_steps, na.Precision.MIXED)

    cv = None if args.no_cv else cv_case(O)

    latency = None
    if not args.no_serve:
        latency = serve_case(args.serve_requests, args.serve_rate, args.serve_sessions, args.serve_slots)

    # HBM bytes per launch of the dominant kernel, from the rocprofv3 --pmc
    # FETCH_SIZE / WRITE_SIZE passes (scaled by a 256 MiB copy)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "eval_m2lnp_vkde_d16_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")

    best_key = "mixed" if ("mixed" in head and head["mixed"]["rel_err_max"] <= 1e-6) else "fp64"
    best = head[best_key]
    tflops = best["kernel_tflops"]
    peak = PEAK["f16"] if best_key == "mixed" else PEAK["f64"]
    cpu = head["cpu"]
    line = {
        "metric": "kernel-pair evals/s, mixed vs FP64, at rel err<=1e-6; p50/p99 request latency",
        "value": best["pairs_per_s"],
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": best["wall_ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16split+f32" if best_key == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": "VKDE eval_m2lnp, Student-t nu=3, d=16, n=20000 kernels with NNLS-fitted weights "
                               "(seeded correlated Gaussian target), m=200000 eval points resident in HBM",
                   "n_kernels": 20000, "d": 16, "m": 200000, "precision": best_key},
        "roofline": {"bound": "mfma", "achieved": tflops, "peak": peak, "unit": "TFLOP/s",
                     "frac": tflops / peak, "traffic": traffic if best_key == "mixed" else None},
        "cpu_baseline": {"value": cpu["pairs_per_s"], "unit": "pairs/s", "cores": cpu["cores"], "kind": "port",
                         "sample": cpu["sample"] + " (FP64 eval_m2lnp of the -O3 restatement)",
                         "value_1core": cpu["pairs_per_s_1core"], "sample_1core": cpu["sample_1core"],
                         "cpu_model": cpu["cpu_model"]},
        "mixed_over_fp64": head.get("mixed_over_fp64"),
        "fp64": head.get("fp64"),
        "mixed": head.get("mixed"),
        "prepare_interp_s": head["prepare_s"],
        "c2_vkde_d8_n4000": c2,
        "latency": latency,
        "apes": apes,
        "c5": c5,
        "cv": cv,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
