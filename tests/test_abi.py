"""CPU-side checks of the C-ABI library: it is built for gfx950, it loads,
and it exports every function that include/nckde.h declares (no compute
call is made here)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nckde.h")
LIB = os.path.join(ROOT, "numcosmo_amd", "libnckde.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nckde_[A-Za-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def libso():
    if not os.path.exists(LIB):
        from numcosmo_amd import build
        build.build()
    return ctypes.CDLL(LIB)


def test_header_declares_the_api():
    names = declared_functions()
    for must in ["nckde_create", "nckde_prepare", "nckde_prepare_interp", "nckde_eval_m2lnp",
                 "nckde_kernel_choose", "nckde_sample", "nckde_compute_IM"]:
        assert must in names


def test_library_exports_every_declared_symbol(libso):
    missing = [n for n in declared_functions() if not hasattr(libso, n)]
    assert missing == []


def test_python_binding_covers_header():
    from numcosmo_amd._lib import SIGNATURES
    assert sorted(SIGNATURES) == declared_functions()


def test_library_carries_gfx950_code():
    out = subprocess.run(["/opt/rocm/llvm/bin/llvm-objdump", "--offloading", LIB],
                         capture_output=True, text=True)
    if out.returncode != 0:
        pytest.skip("llvm-objdump --offloading unavailable")
    assert "gfx950" in out.stdout


def test_create_without_device_fails_cleanly(libso):
    """No GPU in this container: nckde_create must return an error, not crash."""
    import torch  # noqa: F401  (only to know whether a device exists)
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from numcosmo_amd import StatsDistKernelGauss, StatsDistVKDE, NckdeError
    with pytest.raises(NckdeError):
        StatsDistVKDE.new(StatsDistKernelGauss.new(2))


def test_request_struct_layout_matches_the_header(tmp_path):
    """The ctypes mirrors of nckde_request / nckde_request_info have the C
    layout (sizes and field offsets as compiled from the header)."""
    import ctypes as C
    from numcosmo_amd._lib import Request, RequestInfo
    src = tmp_path / "layout.c"
    fields_r = [f for f, _ in Request._fields_]
    fields_i = [f for f, _ in RequestInfo._fields_]
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', 'int main(void){',
             'printf("%zu\\n", sizeof(nckde_request));']
    lines += [f'printf("%zu\\n", offsetof(nckde_request, {f}));' for f in fields_r]
    lines += ['printf("%zu\\n", sizeof(nckde_request_info));']
    lines += [f'printf("%zu\\n", offsetof(nckde_request_info, {f}));' for f in fields_i]
    lines += ['return 0;}']
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [C.sizeof(Request)] + [getattr(Request, f).offset for f in fields_r]
    want += [C.sizeof(RequestInfo)] + [getattr(RequestInfo, f).offset for f in fields_i]
    assert got == want


def test_apes_entry_points_reject_null_handles():
    """Handle checks of the APES walker run on the host: a NULL handle or
    output pointer is an error with a message, never a dereference."""
    from numcosmo_amd._lib import lib
    L = lib()
    rows = ctypes.c_size_t(7)
    assert L.nckde_apes_set_trace(None, None, 0) == -1
    assert b"NULL" in L.nckde_last_error()
    assert L.nckde_apes_get_trace_rows(None, ctypes.byref(rows)) == -1
    assert rows.value == 7
    assert L.nckde_apes_run(None, 1, None, None, None) != 0
    assert L.nckde_apes_run_from(None, 0, 1, None, None, None) != 0
    assert L.nckde_apes_destroy(None) == 0
