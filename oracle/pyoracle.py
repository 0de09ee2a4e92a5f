"""Python handle on the CPU oracle (oracle/liboucoracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the ouc_amd package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from ouc_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboucoracle.so")
_lib = None


def build(force=False):
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        dp = P(C.c_double)
        L.oracle_coeffs.argtypes = [C.c_int, dp]
        L.oracle_ricker.argtypes = [C.c_int, C.c_float, C.c_float, dp, P(C.c_int)]
        L.oracle_setup.argtypes = [P(abi.Params), P(abi.Model), dp, dp, dp, dp]
        L.oracle_dop.argtypes = [C.c_int, C.c_int, dp, C.c_int, C.c_int, C.c_double, dp]
        L.oracle_mute_thresholds.argtypes = [P(abi.Params), P(abi.Model), P(abi.Shot), dp]
        L.oracle_ring.argtypes = [P(abi.Params), P(C.c_int), P(C.c_int64)]
        L.oracle_fp16_encode.argtypes = [P(C.c_float), C.c_int64, C.c_int32, P(C.c_uint16), P(C.c_int8), P(C.c_float)]
        L.oracle_fp16_decode.argtypes = [P(C.c_uint16), P(C.c_int8), C.c_int64, P(C.c_float)]
        L.oracle_f32_to_f16.argtypes = [P(C.c_float), C.c_int64, P(C.c_uint16)]
        L.oracle_forward_shot.argtypes = [P(abi.Params), P(abi.Model), P(abi.Shot), C.c_int, dp]
        L.oracle_rtm_shot.argtypes = [P(abi.Params), P(abi.Model), P(abi.Shot), C.c_int, dp, dp, dp, dp, dp, dp]
        L.oracle_recon_error.argtypes = [P(abi.Params), P(abi.Model), P(abi.Shot), C.c_int, dp]
        L.oracle_forward_splitpml.argtypes = [P(abi.Params), P(abi.Model), P(abi.Shot), C.c_int, dp]
        L.oracle_rtm_splitpml.argtypes = [P(abi.Params), P(abi.Model), P(abi.Shot), C.c_int, dp, dp, dp, dp, dp]
        L.oracle_fwi_gradient.argtypes = [P(abi.Params), P(abi.Model), P(abi.Shot), C.c_int, dp, dp, dp, dp, dp]
        L.oracle_gradfun.argtypes = [P(abi.Params), P(abi.Model)] + [dp] * 11
        L.oracle_set_threads.argtypes = [C.c_int]
        L.oracle_rtm_trace.argtypes = [P(abi.Params), P(abi.Model), P(abi.Shot), C.c_int, dp, dp, dp, dp, dp, dp, dp,
                                       dp]
        L.oracle_forward_trace.argtypes = [P(abi.Params), P(abi.Model), P(abi.Shot), C.c_int, dp]
        L.oracle_laplace.argtypes = [C.c_int, C.c_int, C.c_double, dp, dp]
        L.oracle_set_mutation.argtypes = [C.c_int]
        _lib = L
    return _lib


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed: {abi.STATUS_NAMES.get(rc, rc)}")


def set_threads(n):
    return lib().oracle_set_threads(int(n))


def coeffs(N):
    out = np.zeros(N, np.float64)
    _check(lib().oracle_coeffs(N, abi.dptr(out)), "coeffs")
    return out


def ricker(nt, fm, dt):
    out = np.zeros(nt, np.float64)
    t0 = C.c_int(0)
    _check(lib().oracle_ricker(nt, fm, dt, abi.dptr(out), C.byref(t0)), "ricker")
    return out, t0.value


class _ModelRef:
    """Keeps the float32 arrays alive behind an abi.Model."""

    def __init__(self, vp, vs, rho=None):
        self.vp = np.ascontiguousarray(vp, np.float32)
        self.vs = np.ascontiguousarray(vs, np.float32)
        self.rho = None if rho is None else np.ascontiguousarray(rho, np.float32)
        self.m = abi.Model(abi.fptr(self.vp), abi.fptr(self.vs), abi.fptr(self.rho))


class _ShotRef:
    def __init__(self, sx, sz, rec_z, fm, wavelet=None, mute=0):
        self.w = None if wavelet is None else np.ascontiguousarray(wavelet, np.float32)
        self.s = abi.Shot(int(sx), int(sz), int(rec_z), float(fm), abi.fptr(self.w), int(mute))


def setup(p, vp, vs):
    Xn, Zn = p.nx + 2 * p.npml, p.nz + 2 * p.npml
    m = _ModelRef(vp, vs)
    out = [np.zeros((Xn, Zn), np.float64) for _ in range(4)]
    _check(lib().oracle_setup(C.byref(p), C.byref(m.m), *[abi.dptr(a) for a in out]), "setup")
    return out  # vp, vs, ax, az padded


def dop(N, kind, f, h):
    f = np.ascontiguousarray(f, np.float64)
    out = np.zeros_like(f)
    _check(lib().oracle_dop(N, kind, abi.dptr(f), f.shape[0], f.shape[1], h, abi.dptr(out)), "dop")
    return out


def mute_thresholds(p, vp, vs, shot):
    m = _ModelRef(vp, vs)
    s = _ShotRef(**shot)
    out = np.zeros(p.nx, np.float64)
    _check(lib().oracle_mute_thresholds(C.byref(p), C.byref(m.m), C.byref(s.s), abi.dptr(out)), "mute")
    return out


def ring(p):
    n = C.c_int64(0)
    lib().oracle_ring(C.byref(p), C.POINTER(C.c_int)(), C.byref(n))
    ij = np.zeros((n.value, 2), np.int32)
    lib().oracle_ring(C.byref(p), ij.ctypes.data_as(C.POINTER(C.c_int)), C.byref(n))
    return ij


def fp16_encode(x, decoded=False):
    """Strip encoder: 1-D x is one slot; 2-D x [steps][n] is successive slots
    of one field, coded predictively with error feedback.  Returns (residual
    bits, chunk exponents) and, with decoded=True, the values decoded
    backwards as the reconstruction does."""
    x = np.ascontiguousarray(x, np.float32)
    x2 = x.reshape(1, -1) if x.ndim == 1 else x
    steps, n = x2.shape
    h = np.zeros((steps, n), np.uint16)
    e = np.zeros((steps, (n + 63) // 64), np.int8)
    d = np.zeros((steps, n), np.float32)
    _check(lib().oracle_fp16_encode(abi.fptr(x2), n, steps, h.ctypes.data_as(C.POINTER(C.c_uint16)),
                                    e.ctypes.data_as(C.POINTER(C.c_int8)), abi.fptr(d)), "fp16_encode")
    out = (h, e, d) if decoded else (h, e)
    return tuple(a[0] for a in out) if x.ndim == 1 else out


def fp16_decode(h, e):
    h = np.ascontiguousarray(h, np.uint16)
    e = np.ascontiguousarray(e, np.int8)
    x = np.zeros(h.size, np.float32)
    _check(lib().oracle_fp16_decode(h.ctypes.data_as(C.POINTER(C.c_uint16)), e.ctypes.data_as(C.POINTER(C.c_int8)),
                                    h.size, abi.fptr(x)), "fp16_decode")
    return x


def f32_to_f16(x):
    x = np.ascontiguousarray(x, np.float32)
    h = np.zeros(x.size, np.uint16)
    _check(lib().oracle_f32_to_f16(abi.fptr(x), x.size, h.ctypes.data_as(C.POINTER(C.c_uint16))), "f32_to_f16")
    return h


def forward_shot(p, vp, vs, shot, precision=64):
    m = _ModelRef(vp, vs)
    s = _ShotRef(**shot)
    seis = np.zeros((6, p.nx, p.nt), np.float64)
    _check(lib().oracle_forward_shot(C.byref(p), C.byref(m.m), C.byref(s.s), precision, abi.dptr(seis)), "forward")
    return seis


def rtm_shot(p, vp, vs, shot, precision=64, rec_in=None):
    m = _ModelRef(vp, vs)
    s = _ShotRef(**shot)
    outs = [np.zeros((p.nx, p.nz), np.float64) for _ in range(4)]
    rec_out = np.zeros((2, p.nx, p.nt), np.float64)
    rin = None if rec_in is None else np.ascontiguousarray(rec_in, np.float64)
    _check(lib().oracle_rtm_shot(C.byref(p), C.byref(m.m), C.byref(s.s), precision, abi.dptr(rin),
                                 *[abi.dptr(a) for a in outs], abi.dptr(rec_out)), "rtm")
    return {"img_pp": outs[0], "img_ps": outs[1], "lap_pp": outs[2], "lap_ps": outs[3], "rec": rec_out}


def forward_splitpml(p, vp, vs, shot, precision=64):
    """Seismograms [6][nx][nt] of the first-order split-PML engine."""
    m = _ModelRef(vp, vs)
    s = _ShotRef(**shot)
    out = np.zeros((6, p.nx, p.nt), np.float64)
    _check(lib().oracle_forward_splitpml(C.byref(p), C.byref(m.m), C.byref(s.s), precision, abi.dptr(out)),
           "split-PML forward")
    return out


def rtm_splitpml(p, vp, vs, shot, precision=64, recon_err=False):
    """RTM with the first-order split-PML engine: images, Laplacians and,
    with recon_err, max |Vpx_recon - Vpx_fwd| / max |Vpx_fwd| per level."""
    m = _ModelRef(vp, vs)
    s = _ShotRef(**shot)
    outs = [np.zeros((p.nx, p.nz), np.float64) for _ in range(4)]
    err = np.zeros(p.nt, np.float64) if recon_err else None
    _check(lib().oracle_rtm_splitpml(C.byref(p), C.byref(m.m), C.byref(s.s), precision,
                                     *[abi.dptr(a) for a in outs], abi.dptr(err)), "split-PML rtm")
    out = {"img_pp": outs[0], "img_ps": outs[1], "lap_pp": outs[2], "lap_ps": outs[3]}
    if recon_err:
        out["recon_err"] = err
    return out


def fwi_gradient(p, vp, vs, shot, observed, precision=64):
    """gradfun accumulated over one shot's backward loop, driven by the
    residual modelled - observed (adsource); observed is [2][nx][nt]."""
    m = _ModelRef(vp, vs)
    s = _ShotRef(**shot)
    obs = np.ascontiguousarray(observed, np.float64)
    assert obs.shape == (2, p.nx, p.nt)
    outs = [np.zeros((p.nx, p.nz), np.float64) for _ in range(3)]
    mis = np.zeros(1, np.float64)
    _check(lib().oracle_fwi_gradient(C.byref(p), C.byref(m.m), C.byref(s.s), precision, abi.dptr(obs),
                                     *[abi.dptr(a) for a in outs], abi.dptr(mis)), "fwi")
    return {"grad_vp": outs[0], "grad_vs": outs[1], "illum": outs[2], "misfit": float(mis[0])}


def gradfun(p, vp, vs, fields, acc):
    """One gradfun step on padded (Xn, Zn) FP64 fields ux, uz, th, om, uxx,
    uxz, uzx, uzz; acc = [gvp, gvs, ill] is updated in place."""
    m = _ModelRef(vp, vs)
    f = [np.ascontiguousarray(fields[k], np.float64) for k in ("ux", "uz", "th", "om", "uxx", "uxz", "uzx", "uzz")]
    for a in acc:
        assert a.dtype == np.float64 and a.flags.c_contiguous
    _check(lib().oracle_gradfun(C.byref(p), C.byref(m.m), *[abi.dptr(a) for a in f], *[abi.dptr(a) for a in acc]),
           "gradfun")


def recon_error(p, vp, vs, shot, precision=64):
    m = _ModelRef(vp, vs)
    s = _ShotRef(**shot)
    err = np.zeros(p.nt, np.float64)
    _check(lib().oracle_recon_error(C.byref(p), C.byref(m.m), C.byref(s.s), precision, abi.dptr(err)), "recon")
    return err


def rtm_trace(p, vp, vs, shot, precision=64, rec_in=None):
    """RTM with the backward-loop traces (tiny grids only): images, forward
    source split velocities fwd_v [nt+1][4][nx][nz] (px, pz, sx, sz), the
    source P velocity as the backward loop holds it src_v [nt+1][2][nx][nz]
    (levels nt, nt-1 handed over, level t reconstructed at step t), receiver split parts after each
    backward step's velocity pass rcv [nt][4][nx][nz] (pxr, pzr, sxr, szr),
    accumulators acc [3][nx][nz] (num_pp, num_ps, den) and the receiver row
    right after injection inj [nt][2][Xn]."""
    m = _ModelRef(vp, vs)
    s = _ShotRef(**shot)
    nx, nz, nt = p.nx, p.nz, p.nt
    Xn = nx + 2 * p.npml
    out = {"img_pp": np.zeros((nx, nz)), "img_ps": np.zeros((nx, nz)),
           "fwd_v": np.zeros((nt + 1, 4, nx, nz)), "src_v": np.zeros((nt + 1, 2, nx, nz)), "rcv": np.zeros((nt, 4, nx, nz)),
           "acc": np.zeros((3, nx, nz)), "inj": np.zeros((nt, 2, Xn))}
    rin = None if rec_in is None else np.ascontiguousarray(rec_in, np.float64)
    _check(lib().oracle_rtm_trace(C.byref(p), C.byref(m.m), C.byref(s.s), precision, abi.dptr(rin),
                                  *[abi.dptr(out[k]) for k in ("img_pp", "img_ps", "fwd_v", "src_v", "rcv", "acc", "inj")]),
           "rtm trace")
    return out


def forward_trace(p, vp, vs, shot, precision=64):
    """Interior split velocities of every forward level: [nt+1][4][nx][nz]
    (px, pz, sx, sz; level 0 is zero).  Tiny grids only."""
    m = _ModelRef(vp, vs)
    s = _ShotRef(**shot)
    out = np.zeros((p.nt + 1, 4, p.nx, p.nz))
    _check(lib().oracle_forward_trace(C.byref(p), C.byref(m.m), C.byref(s.s), precision, abi.dptr(out)),
           "forward trace")
    return out


def laplace(img, dh):
    """The reference's Laplace filter (5-point / dh^2, 2-cell zero border) of
    an interior image, FP64."""
    img = np.ascontiguousarray(img, np.float64)
    out = np.zeros_like(img)
    _check(lib().oracle_laplace(img.shape[0], img.shape[1], float(dh), abi.dptr(img), abi.dptr(out)), "laplace")
    return out


class mutation:
    """Context manager planting mistake u in the oracle, for pins that must
    catch it (1..12: ax/az swapped in one CPML memory-variable update; 13, 14:
    the split-PML reconstruction's ring velocities one slot late, its source
    sample dropped)."""

    def __init__(self, u):
        self.u = int(u)

    def __enter__(self):
        _check(lib().oracle_set_mutation(self.u), "set_mutation")
        return self

    def __exit__(self, *exc):
        lib().oracle_set_mutation(0)
        return False
