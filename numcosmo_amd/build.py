"""Build libnckde.so (gfx950) in-tree with hipcc.

Usage:  python -m numcosmo_amd.build [-j N]

Every HIP translation unit under csrc/ is compiled for gfx950 with packed
FP32 instructions disabled, then linked into numcosmo_amd/libnckde.so.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libnckde.so")
BUILD_DIR = os.path.join(HERE, "_build")

ARCH = os.environ.get("NCKDE_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 VALU instructions stay off in all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # MFMA sources and results in ordinary VGPRs (no accvgpr copies)
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I", os.path.join(HERE, "..", "include"),
]

SOURCES = ["api.cpp", "api_apes.cpp", "api_serve.cpp", "cv_min.cpp", "k_prepare.hip", "k_eval.hip", "k_eval_mixed.hip", "k_linalg.hip", "k_nnls.hip", "k_probe.hip", "k_sample.hip", "k_robust.hip"]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD_DIR, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *COMMON, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int | None = None, verbose: bool = False) -> str:
    os.makedirs(BUILD_DIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    # the offload linker leaves per-target intermediates beside its output:
    # link in the build folder, then move the library into place
    tmp = os.path.join(BUILD_DIR, "libnckde.so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=BUILD_DIR)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    args = ap.parse_args(argv)
    build(args.j, verbose=True)


if __name__ == "__main__":
    sys.exit(main())
