"""The allocation-free forward pass of the boundary network
(ouc_lb_forward_device: device buffers in and out, BF16 weights and hidden
activations in the library's per-GPU scratch array) against ouc_lb_forward,
which allocates its own buffers: bit-identical outputs on two calls in a row,
the second on a smaller input, so a stale scratch would show.

The caller's device buffers come from the HIP runtime libouc.so itself is
linked against (found in this process's mappings), not from torch, which
carries a runtime of its own."""
import ctypes as C

import numpy as np
import pytest

from ouc_amd import engine

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    engine.lib()
    path = None
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64.so" in line:
                path = line.split()[-1]
                break
    assert path, "libouc.so did not bring in a HIP runtime"
    h = C.CDLL(path)
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    return h


def _check(rc, what):
    assert rc == 0, f"{what} returned {rc}"


def _device_forward(h, x, w):
    M, _, H, W = x.shape
    out = np.empty((M, 2, H, W), np.float32)
    bufs = []
    try:
        for a in (x, w, out):
            p = C.c_void_p()
            _check(h.hipMalloc(C.byref(p), a.nbytes), "hipMalloc")
            bufs.append(p)
        _check(h.hipMemcpy(bufs[0], x.ctypes.data, x.nbytes, H2D), "hipMemcpy")
        _check(h.hipMemcpy(bufs[1], w.ctypes.data, w.nbytes, H2D), "hipMemcpy")
        _check(h.hipMemset(bufs[2], 0xFF, out.nbytes), "hipMemset")  # NaN: every output must be written
        _check(h.hipDeviceSynchronize(), "hipDeviceSynchronize")
        engine.lb_forward_device(M, H, W, bufs[0].value, bufs[1].value, bufs[2].value, device=0, stream=None)
        _check(h.hipDeviceSynchronize(), "hipDeviceSynchronize")
        _check(h.hipMemcpy(out.ctypes.data, bufs[2], out.nbytes, D2H), "hipMemcpy")
    finally:
        for p in bufs:
            h.hipFree(p)
    return out


def test_no_alloc_forward_equals_the_allocating_entry():
    h = _hip()
    rng = np.random.default_rng(11)
    w = (rng.standard_normal(3312) * 0.2).astype(np.float32)
    for M, H, W in ((40, 7, 70), (9, 5, 33)):
        x = rng.standard_normal((M, 5, H, W)).astype(np.float32)
        print(f"[case] lb_forward_device M={M} H={H} W={W}", flush=True)
        got = _device_forward(h, x, w)
        ref = engine.lb_forward(x, w)
        assert np.isfinite(got).all()
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_no_alloc_forward_refuses_inputs_beyond_the_scratch():
    with pytest.raises(engine.OucError) as e:
        engine.lb_forward_device(2, 256, 257, 1, 1, 1)
    assert e.value.status == 1 and "scratch maximum" in str(e.value)
