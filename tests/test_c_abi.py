"""The plain-C caller of the boundary (tests/c_abi/nckde_c_driver.c): it
includes include/nckde.h from C99 with -Werror -pedantic, links
libnckde.so, and on a GPU drives create -> add_obs -> prepare_interp ->
eval (FP64 and mixed) -> kernel_choose -> sample -> getters -> serialise,
checking each result against the oracle's C API."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CDIR = os.path.join(ROOT, "tests", "c_abi")
EXE = os.path.join(CDIR, "nckde_c_driver")


def _build():
    if not os.path.exists(os.path.join(ROOT, "numcosmo_amd", "libnckde.so")):
        from numcosmo_amd import build
        build.build()
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    r = subprocess.run(["make", "-s", "-C", CDIR], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr


def test_c_driver_compiles_and_links():
    _build()
    assert os.access(EXE, os.X_OK)
    # the binary resolves both libraries through its rpaths
    out = subprocess.run(["ldd", EXE], capture_output=True, text=True).stdout
    assert "libnckde.so" in out and "not found" not in out.split("libnckde.so")[1].splitlines()[0]


@pytest.mark.gpu
def test_c_driver_end_to_end():
    if not os.access(EXE, os.X_OK):
        _build()
    print("[c-abi] running tests/c_abi/nckde_c_driver", flush=True)
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().splitlines()[-1] == "nckde C driver: OK"
