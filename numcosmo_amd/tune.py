"""Tuner of the eval's kernel-range split count.

Usage (on the GPU the library will run on):

    python -m numcosmo_amd.tune [--out numcosmo_amd/csrc/eval_tune.inc] [--report FILE]

For every case of a grid covering the served sizes -- d in {4, 8, 16, 32},
n in {512, 2048, 8192, 20000} kernels and m = 10 n points, plus n = 4000 at
m = 25 n (1e5 points), the mixed and FP64 paths -- it fits nothing (a plain VKDE prepare of a seeded Gaussian
sample is enough: the timing depends on the sizes only), then times the eval
for split counts around the default rule (x1/4 ... x8, clamped), each the
median of five launches after a warm-up, and keeps the fastest.  The table
it writes (the fastest count of every case) is compiled into libnckde.so
(rebuild afterwards): an eval uses the entry of its precision and dimension
nearest in (log2 n, log2 m) within one octave, else the default rule.  The split count changes results only in the
last bits (the partial sums are merged in a fixed order), and a given
(n, m) always gets the same count, so evals stay reproducible.
"""
from __future__ import annotations

import argparse
import ctypes as C
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _mfma_dim(d):
    for p in (4, 8, 16, 32, 64):
        if d <= p:
            return p
    return -1


def _padded_dim(d):
    for p in (4, 8, 12, 16, 24, 32, 48, 64):
        if d <= p:
            return p
    return -1


def default_nsplit(prec, d, n, m):
    """The library's default rule (api.cpp mixed_nsplit / fp64_nsplit)."""
    if prec == 1:
        DP = _mfma_dim(d)
        kpt = 1 if DP >= 32 else 32 // DP
        ngroups = (n + kpt - 1) // kpt
        ppb = (2 if DP == 64 else (8 if DP == 16 else 4)) * 32 * (8 if DP == 64 else 4)
        bx = (m + ppb - 1) // ppb
        ns = (8192 + bx - 1) // bx
        return max(1, min(ns, max(1, ngroups // 8))), ngroups
    nblk = (m + 255) // 256
    ns = (16384 + nblk - 1) // nblk
    return max(1, min(ns, (n + 31) // 32)), n


def _time_eval(sd, Xd, outd, prec, stream, reps=5):
    import torch

    from ._lib import check, lib

    L = lib()

    def step():
        check(L.nckde_eval_m2lnp_device(sd._h, C.c_void_p(Xd.data_ptr()), Xd.shape[0], C.c_void_p(outd.data_ptr()),
                                        int(prec), C.c_void_p(stream.cuda_stream)))

    step()
    stream.synchronize()
    ts = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        step()
        e1.record(stream)
        e1.synchronize()
        ts.append(e0.elapsed_time(e1))
    return float(np.median(ts))


# (kernels, points) of the tuned cases
CASES = ((512, 5120), (2048, 20480), (4000, 100000), (8192, 81920), (20000, 200000))


def tune(dims=(4, 8, 16, 32), cases=CASES, precisions=(1, 0), log=print):
    import torch

    import numcosmo_amd as na

    torch.cuda.init()
    rows = []
    for d in dims:
        for n, m in cases:
            g = np.random.default_rng(1000 * d + n)
            X = g.standard_normal((n, d))
            Xe = g.standard_normal((m, d))
            sd = na.StatsDistVKDE.new(na.StatsDistKernelST.new(d, 3.0))
            sd.add_obs_array(X)
            sd.prepare()
            Xd = torch.tensor(Xe, dtype=torch.float64, device="cuda")
            outd = torch.empty(m, dtype=torch.float64, device="cuda")
            stream = torch.cuda.Stream()
            for prec in precisions:
                ns0, nmax = default_nsplit(prec, d, n, m)
                cands = sorted({max(1, min(nmax, int(round(ns0 * f)))) for f in (0.25, 0.5, 1, 2, 4, 8)})
                times = {}
                for ns in cands:
                    sd.set_eval_split(ns)
                    times[ns] = _time_eval(sd, Xd, outd, prec, stream)
                sd.set_eval_split(0)
                best = min(times, key=times.get)
                rows.append({"prec": prec, "d": d, "dim": _mfma_dim(d) if prec == 1 else _padded_dim(d), "n": n,
                             "m": m, "default": ns0, "default_ms": times[ns0], "best": best,
                             "best_ms": times[best], "times": times})
                log(f"[tune] prec={prec} d={d} n={n} m={m}: default {ns0} {times[ns0]:.3f} ms, best {best} "
                    f"{times[best]:.3f} ms")
    return rows


def write_table(rows, path, device_name):
    lines = [f"// Eval split counts measured by numcosmo_amd/tune.py on {device_name}:",
             "// {precision, dimension, log2 n, log2 m, split count}, the fastest count of",
             "// each tuned case."]
    for r in rows:
        lines.append(f"{{{r['prec']}, {r['dim']}, {int(round(math.log2(r['n'])))}, "
                     f"{int(round(math.log2(r['m'])))}, {r['best']}}},")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def write_report(rows, path, device_name, seconds):
    out = [f"# Eval split tuning ({device_name}, {seconds:.0f} s)", "",
           "`python -m numcosmo_amd.tune`: VKDE Student-t nu=3 on a seeded Gaussian sample, m = 10 n (25 n at "
           "n = 4000), each time the median of five event-timed launches after a warm-up.", "",
           "| precision | d | n | m | default splits | ms | best splits | ms | gain |",
           "|---|---:|---:|---:|---:|---:|---:|---:|---:|"]
    for r in rows:
        out.append(f"| {'mixed' if r['prec'] == 1 else 'FP64'} | {r['d']} | {r['n']} | {r['m']} | {r['default']} | "
                   f"{r['default_ms']:.3f} | {r['best']} | {r['best_ms']:.3f} | "
                   f"{r['default_ms'] / r['best_ms']:.2f}x |")
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "csrc", "eval_tune.inc"))
    ap.add_argument("--report", default=None)
    ap.add_argument("--fp64", action="store_true", help="also tune the FP64 path")
    args = ap.parse_args(argv)
    import torch
    t0 = time.perf_counter()
    rows = tune(precisions=(1, 0) if args.fp64 else (1,))
    name = torch.cuda.get_device_name(0)
    write_table(rows, args.out, name)
    if args.report:
        write_report(rows, args.report, name, time.perf_counter() - t0)
    print(f"wrote {args.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
