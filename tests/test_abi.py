"""CPU checks of the C ABI: every function include/ouc.h declares is exported
by ouc_amd/libouc.so, the header compiles as C99, the struct layouts agree
with the ctypes mirrors, and argument errors are reported before any device
call."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ouc.h")
LIB = os.path.join(ROOT, "ouc_amd", "libouc.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ouc_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for must in ("ouc_create", "ouc_forward_shot", "ouc_rtm_shot", "ouc_rtm_batch", "ouc_write_image_files",
                 "ouc_last_error", "ouc_destroy", "ouc_set_learned_boundary", "ouc_lb_train"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libouc.so not built")
def test_library_exports_every_declared_function():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s+(ouc_\w+)", out))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    C.CDLL(LIB)  # loads without a GPU


C_PROBE = r"""
#include <stddef.h>
#include <stdio.h>
#include "ouc.h"
int main(void) {
    printf("%zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(ouc_params), sizeof(ouc_model), sizeof(ouc_shot),
           sizeof(ouc_timings), offsetof(ouc_params, R), offsetof(ouc_timings, boundary_ratio),
           sizeof(ouc_segy_gather), offsetof(ouc_segy_gather, component));
    return 0;
}
"""


def test_header_compiles_as_c99_and_layouts_match_ctypes(tmp_path):
    from ouc_amd import abi

    src = tmp_path / "probe.c"
    src.write_text(C_PROBE)
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                    str(src), "-o", str(exe)], check=True)
    sizes = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert sizes == [C.sizeof(abi.Params), C.sizeof(abi.Model), C.sizeof(abi.Shot), C.sizeof(abi.Timings),
                     abi.Params.R.offset, abi.Timings.boundary_ratio.offset, C.sizeof(abi.SegyGather),
                     abi.SegyGather.component.offset]


@pytest.mark.skipif(not os.path.exists(LIB), reason="libouc.so not built")
def test_argument_errors_need_no_device():
    from ouc_amd import abi
    from ouc_amd.engine import lib

    L = lib()
    h = C.c_void_p()
    st = L.ouc_create(None, None, 0, C.byref(h))
    assert st == abi.OUC_INVALID_ARG
    assert L.ouc_last_error(None)
    # a bad grid is rejected during validation, before any HIP call
    p = abi.make_params(4, 4, 2, 5, 10, 5.0, 5.0, 1e-3)
    vp = (C.c_float * 16)(*([2000.0] * 16))
    m = abi.Model(vp, vp, None)
    st = L.ouc_create(C.byref(p), C.byref(m), 0, C.byref(h))
    assert st == abi.OUC_INVALID_ARG
    assert b"half_order" in L.ouc_last_error(None) or b"npml" in L.ouc_last_error(None)
    # the learned boundary needs the fused engine
    p = abi.make_params(64, 64, 20, 5, 10, 5.0, 5.0, 1e-3, boundary=abi.BOUNDARY_LEARNED)
    vp = (C.c_float * (64 * 64))(*([2000.0] * (64 * 64)))
    m = abi.Model(vp, vp, None)
    assert L.ouc_create(C.byref(p), C.byref(m), 0, C.byref(h)) == abi.OUC_INVALID_ARG
    assert b"fused" in L.ouc_last_error(None)


@pytest.mark.skipif(not os.path.exists(LIB), reason="libouc.so not built")
def test_argument_checks_need_no_gpu():
    """Entry points reject bad arguments before touching a device."""
    import numpy as np

    from ouc_amd import abi
    from ouc_amd.engine import lib

    ms = np.zeros(1, np.float32)
    assert lib().ouc_stream_probe(0, 0, 1, 10, 0, 0, 1, 1, abi.fptr(ms)) == abi.OUC_INVALID_ARG
    assert lib().ouc_stream_probe(0, 1, 17, 10, 0, 0, 1, 1, abi.fptr(ms)) == abi.OUC_INVALID_ARG
    assert lib().ouc_stream_probe(0, 2, 1, 10, 2, 5, 1, 1, abi.fptr(ms)) == abi.OUC_INVALID_ARG
    assert lib().ouc_stream_probe(0, 2, 1, 10, 0, 0, 3, 1, abi.fptr(ms)) == abi.OUC_INVALID_ARG
    v = np.zeros(4, np.float32)
    assert lib().ouc_model_update(0, 0, abi.fptr(v), abi.fptr(v), 30.0, 0, abi.fptr(v)) == abi.OUC_INVALID_ARG
    assert lib().ouc_fwi_gradient(None, None, 1, None, None, None, None, None) == abi.OUC_INVALID_ARG
    assert lib().ouc_rtm_grouped(None, 0, None, None, None, None) == abi.OUC_INVALID_ARG
    assert lib().ouc_rtm_concurrent(None, 0, None, 0, None, None, None) == abi.OUC_INVALID_ARG
    assert lib().ouc_rtm_concurrent(None, 3, None, 8, None, None, None) == abi.OUC_INVALID_ARG
