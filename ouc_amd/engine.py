"""Python handle on the HIP engine (ouc_amd/libouc.so) through its C ABI.

Mirrors include/ouc.h: an Engine owns one ouc_ctx (one grid and model on one
GPU) and runs forward shots, RTM shots and batched RTM.  There is no CPU
fallback: if the HIP library is missing, importing this module's loader raises.
"""
import ctypes as C
import os
import warnings

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libouc.so")
_lib = None


class OucError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{abi.STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


def lib():
    """Load libouc.so (built by __graft_entry__.build()); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OucError(abi.OUC_UNSUPPORTED, f"HIP library {LIB_PATH} is not built; run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    fp = P(C.c_float)
    L.ouc_create.argtypes = [P(abi.Params), P(abi.Model), C.c_int32, P(C.c_void_p)]
    L.ouc_create.restype = C.c_int
    L.ouc_forward_shot.argtypes = [C.c_void_p, P(abi.Shot), fp]
    L.ouc_forward_shot.restype = C.c_int
    L.ouc_rtm_shot.argtypes = [C.c_void_p, P(abi.Shot), fp, fp, fp, fp, fp]
    L.ouc_rtm_shot.restype = C.c_int
    L.ouc_rtm_batch.argtypes = [C.c_void_p, P(abi.Shot), C.c_int32, fp, fp, fp]
    L.ouc_rtm_batch.restype = C.c_int
    L.ouc_rtm_grouped.argtypes = [P(C.c_void_p), C.c_int32, P(P(abi.Shot)), P(C.c_int32), P(fp), P(fp)]
    L.ouc_rtm_grouped.restype = C.c_int
    L.ouc_rtm_concurrent.argtypes = [P(C.c_void_p), C.c_int32, P(abi.Shot), C.c_int32, fp, fp, fp]
    L.ouc_rtm_concurrent.restype = C.c_int
    L.ouc_get_timings.argtypes = [C.c_void_p, P(abi.Timings)]
    L.ouc_get_timings.restype = C.c_int
    L.ouc_set_profiling.argtypes = [C.c_void_p, C.c_int32]
    L.ouc_set_profiling.restype = C.c_int
    L.ouc_write_image_files.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, fp, fp]
    L.ouc_write_image_files.restype = C.c_int
    L.ouc_strip_codec_fp16.argtypes = [C.c_int32, fp, C.c_int64, C.c_int32, P(C.c_uint16), P(C.c_int8), fp]
    L.ouc_strip_codec_fp16.restype = C.c_int
    L.ouc_write_model_files.argtypes = [C.c_char_p, C.c_int32, C.c_int32, fp, fp]
    L.ouc_write_model_files.restype = C.c_int
    L.ouc_write_segy.argtypes = [C.c_char_p, P(abi.SegyGather), fp]
    L.ouc_write_segy.restype = C.c_int
    L.ouc_read_segy.argtypes = [C.c_char_p, P(C.c_int32), P(C.c_int32), fp, fp, C.c_int64]
    L.ouc_read_segy.restype = C.c_int
    L.ouc_stream_probe.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int64, C.c_int32, C.c_int64, C.c_int32,
                                   C.c_int32, fp]
    L.ouc_stream_probe.restype = C.c_int
    L.ouc_fwi_gradient.argtypes = [C.c_void_p, P(abi.Shot), C.c_int32, fp, fp, fp, fp, P(C.c_double)]
    L.ouc_fwi_gradient.restype = C.c_int
    L.ouc_model_update.argtypes = [C.c_int32, C.c_int64, fp, fp, C.c_double, C.c_int32, fp]
    L.ouc_model_update.restype = C.c_int
    L.ouc_set_batch_capacity.argtypes = [C.c_void_p, C.c_int32, P(C.c_double)]
    L.ouc_set_batch_capacity.restype = C.c_int
    L.ouc_set_tile.argtypes = [C.c_void_p, P(abi.Tile)]
    L.ouc_set_tile.restype = C.c_int
    L.ouc_get_tile.argtypes = [C.c_void_p, P(abi.Tile)]
    L.ouc_get_tile.restype = C.c_int
    L.ouc_autotune.argtypes = [C.c_void_p, C.c_int32, C.c_int32, P(abi.Tile), fp]
    L.ouc_autotune.restype = C.c_int
    L.ouc_device_arch.argtypes = [C.c_int32, C.c_char_p, C.c_int32]
    L.ouc_device_arch.restype = C.c_int
    L.ouc_set_graphs.argtypes = [C.c_void_p, C.c_int32]
    L.ouc_set_graphs.restype = C.c_int
    L.ouc_set_learned_boundary.argtypes = [C.c_void_p, fp, C.c_int32]
    L.ouc_set_learned_boundary.restype = C.c_int
    L.ouc_set_learned_damping.argtypes = [C.c_void_p, C.c_float]
    L.ouc_set_learned_damping.restype = C.c_int
    L.ouc_get_learned_filter.argtypes = [C.c_void_p, fp]
    L.ouc_get_learned_filter.restype = C.c_int
    L.ouc_forward_snapshots.argtypes = [C.c_void_p, P(abi.Shot), C.c_int32, P(C.c_int32), C.c_int32, C.c_int32,
                                        C.c_int32, C.c_int32, fp]
    L.ouc_forward_snapshots.restype = C.c_int
    L.ouc_recon_snapshots.argtypes = L.ouc_forward_snapshots.argtypes
    L.ouc_recon_snapshots.restype = C.c_int
    L.ouc_lb_forward.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, fp, fp, fp]
    L.ouc_lb_forward.restype = C.c_int
    L.ouc_lb_forward_device.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_void_p]
    L.ouc_lb_forward_device.restype = C.c_int
    L.ouc_lb_loss_grad.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, fp, fp, fp, fp, fp, fp]
    L.ouc_lb_loss_grad.restype = C.c_int
    L.ouc_lb_train.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, fp, fp, fp, C.c_int32, C.c_float, fp, fp]
    L.ouc_lb_train.restype = C.c_int
    L.ouc_last_error.argtypes = [C.c_void_p]
    L.ouc_last_error.restype = C.c_char_p
    L.ouc_destroy.argtypes = [C.c_void_p]
    L.ouc_destroy.restype = None
    _lib = L
    return _lib


def _shot_struct(shot, nt):
    """dict(sx, sz, rec_z, fm, wavelet=None, mute=0) -> (abi.Shot, keepalive)."""
    w = shot.get("wavelet")
    nts = int(shot.get("nt", 0))
    if w is not None:
        w = np.ascontiguousarray(w, np.float32)
        if w.size < (nts or nt):
            raise ValueError("wavelet shorter than the shot's steps")
    s = abi.Shot(int(shot["sx"]), int(shot["sz"]), int(shot["rec_z"]), float(shot.get("fm", 0.0)), abi.fptr(w),
                 int(shot.get("mute", 0)), nts)
    return s, w


class Engine:
    """One grid + model on one GPU (ouc_create / ouc_destroy)."""

    def __init__(self, params, vp, vs, rho=None, device=0):
        self.p = params
        self.vp = np.ascontiguousarray(vp, np.float32)
        self.vs = np.ascontiguousarray(vs, np.float32)
        if self.vp.shape != (params.nx, params.nz) or self.vs.shape != (params.nx, params.nz):
            raise ValueError(f"model must be ({params.nx}, {params.nz}) x-major")
        self.rho = None if rho is None else np.ascontiguousarray(rho, np.float32)
        self._m = abi.Model(abi.fptr(self.vp), abi.fptr(self.vs), abi.fptr(self.rho))
        h = C.c_void_p()
        st = lib().ouc_create(C.byref(self.p), C.byref(self._m), int(device), C.byref(h))
        if st != abi.OUC_OK:
            raise OucError(st, lib().ouc_last_error(None).decode())
        self._h = h

    def _check(self, st):
        """Raise on failure.  OUC_NUMERIC_FALLBACK is not a failure: the
        learned boundary diverged and the results come from the CPML context;
        it sets self.fell_back and warns."""
        self.fell_back = False
        if st == abi.OUC_NUMERIC_FALLBACK:
            self.fell_back = True
            warnings.warn("ouc: " + lib().ouc_last_error(self._h).decode(), RuntimeWarning)
            return
        if st != abi.OUC_OK:
            raise OucError(st, lib().ouc_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            lib().ouc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_profiling(self, on=True):
        self._check(lib().ouc_set_profiling(self._h, int(bool(on))))

    def forward_shot(self, shot):
        """Seismograms [6, nx, nt]: vx, vz, vpx, vpz, vsx, vsz at row rec_z."""
        s, keep = _shot_struct(shot, self.p.nt)
        seis = np.zeros((6, self.p.nx, self.p.nt), np.float32)
        self._check(lib().ouc_forward_shot(self._h, C.byref(s), abi.fptr(seis)))
        return seis

    def rtm_shot(self, shot, rec_in=None):
        s, keep = _shot_struct(shot, self.p.nt)
        rin = None if rec_in is None else np.ascontiguousarray(rec_in, np.float32)
        outs = [np.zeros((self.p.nx, self.p.nz), np.float32) for _ in range(4)]
        self._check(lib().ouc_rtm_shot(self._h, C.byref(s), abi.fptr(rin), *[abi.fptr(a) for a in outs]))
        return {"img_pp": outs[0], "img_ps": outs[1], "lap_pp": outs[2], "lap_ps": outs[3]}

    def rtm_batch(self, shots, per_shot=False):
        n = len(shots)
        arr = (abi.Shot * n)()
        keep = []
        for k, sh in enumerate(shots):
            s, w = _shot_struct(sh, self.p.nt)
            arr[k] = s
            keep.append(w)
        spp = np.zeros((self.p.nx, self.p.nz), np.float32)
        sps = np.zeros_like(spp)
        ms = np.zeros(n, np.float32)
        self._check(lib().ouc_rtm_batch(self._h, arr, n, abi.fptr(spp), abi.fptr(sps), abi.fptr(ms)))
        out = {"stack_pp": spp, "stack_ps": sps}
        if per_shot:
            out["per_shot_ms"] = ms
        return out

    def fwi_gradient(self, shots, observed):
        """Elastic FWI gradient summed over `shots` (fused engine): observed is
        [nshots][2][nx][nt] (a single shot may pass [2][nx][nt]).  Returns
        grad_vp, grad_vs, illum (nx, nz) and the misfit 0.5 * sum(residual^2)."""
        n = len(shots)
        arr = (abi.Shot * n)()
        keep = []
        for k, sh in enumerate(shots):
            s, w = _shot_struct(sh, self.p.nt)
            arr[k] = s
            keep.append(w)
        obs = np.ascontiguousarray(observed, np.float32).reshape(n, 2, self.p.nx, self.p.nt)
        outs = [np.zeros((self.p.nx, self.p.nz), np.float32) for _ in range(3)]
        mis = C.c_double()
        self._check(lib().ouc_fwi_gradient(self._h, arr, n, abi.fptr(obs), *[abi.fptr(a) for a in outs],
                                           C.byref(mis)))
        return {"grad_vp": outs[0], "grad_vs": outs[1], "illum": outs[2], "misfit": mis.value}

    def set_batch_capacity(self, shots):
        """At most `shots` shots per fused launch (0: by free memory); returns
        the device bytes one shot of a batch holds."""
        b = C.c_double()
        self._check(lib().ouc_set_batch_capacity(self._h, int(shots), C.byref(b)))
        return b.value

    def set_tile(self, core_cols, prefetch):
        t = abi.Tile(int(core_cols), int(prefetch))
        self._check(lib().ouc_set_tile(self._h, C.byref(t)))

    def tile(self):
        t = abi.Tile()
        self._check(lib().ouc_get_tile(self._h, C.byref(t)))
        return int(t.core_cols), int(t.prefetch)

    def autotune(self, nshots, reps=5):
        """Times the 18 candidate tiles on this grid, keeps the fastest;
        returns ((core_cols, prefetch), {(cols, pf): ms per launch}).  The sixth
        width, labelled "fit", is the one at which a single round of resident
        blocks covers the launch (ouc_autotune)."""
        t = abi.Tile()
        ms = np.zeros(18, np.float32)
        self._check(lib().ouc_autotune(self._h, int(nshots), int(reps), C.byref(t), abi.fptr(ms)))
        cands = [(c, p) for c in (8, 16, 32, 64, 128, "fit") for p in (1, 2, 3)]
        return (int(t.core_cols), int(t.prefetch)), {k: float(v) for k, v in zip(cands, ms)}

    def set_graphs(self, on=True):
        """Replay captured time-loop graphs (fused engine; default on)."""
        self._check(lib().ouc_set_graphs(self._h, int(bool(on))))

    def set_learned_boundary(self, weights):
        """FP32 master weights of the boundary network (3312 values), or None
        for the pad without the network."""
        if weights is None:
            self._check(lib().ouc_set_learned_boundary(self._h, abi.fptr(None), 0))
            return
        w = np.ascontiguousarray(weights, np.float32).ravel()
        self._check(lib().ouc_set_learned_boundary(self._h, abi.fptr(w), w.size))

    def set_learned_damping(self, beta):
        """Weak sponge of the learned pad (increment scaled by 1 - beta r^2)."""
        self._check(lib().ouc_set_learned_damping(self._h, float(beta)))

    def learned_filter(self):
        """The composed filter [2][5][7][7] the ring kernel applies."""
        out = np.zeros((2, 5, 7, 7), np.float32)
        self._check(lib().ouc_get_learned_filter(self._h, abi.fptr(out)))
        return out

    def forward_snapshots(self, shot, steps, x0, z0, w, h):
        """V (vx, vz) over the padded-grid rectangle after each listed step:
        [len(steps), 2, w, h]."""
        s, keep = _shot_struct(shot, self.p.nt)
        st = np.ascontiguousarray(steps, np.int32)
        out = np.zeros((st.size, 2, w, h), np.float32)
        self._check(lib().ouc_forward_snapshots(self._h, C.byref(s), st.size,
                                                st.ctypes.data_as(C.POINTER(C.c_int32)), int(x0), int(z0), int(w),
                                                int(h), abi.fptr(out)))
        return out

    def recon_snapshots(self, shot, steps, x0, z0, w, h):
        """Reconstructed source V^t over the padded-grid rectangle at each
        listed (decreasing) time: [len(steps), 2, w, h]."""
        s, keep = _shot_struct(shot, self.p.nt)
        st = np.ascontiguousarray(steps, np.int32)
        out = np.zeros((st.size, 2, w, h), np.float32)
        self._check(lib().ouc_recon_snapshots(self._h, C.byref(s), st.size, st.ctypes.data_as(C.POINTER(C.c_int32)),
                                              int(x0), int(z0), int(w), int(h), abi.fptr(out)))
        return out

    def timings(self):
        t = abi.Timings()
        self._check(lib().ouc_get_timings(self._h, C.byref(t)))
        return {name: getattr(t, name) for name, _ in abi.Timings._fields_}


def rtm_grouped(engines, shot_lists):
    """Batched RTM over several grids in shared launches (ouc_rtm_grouped):
    engines[i] (its own grid and model, fused engine) runs shot_lists[i].  Every
    kernel of a time step is one launch over the shots of all grids.  Returns
    one dict(stack_pp, stack_ps) per engine, equal to that engine's rtm_batch of
    its shots.  OUC_NUMERIC_FALLBACK (a learned boundary diverged; the shots ran
    again on CPML) sets fell_back on the first engine and warns."""
    n = len(engines)
    if n != len(shot_lists):
        raise ValueError("one shot list per engine")
    hs = (C.c_void_p * n)(*[e._h for e in engines])
    arrs, keep = [], []
    for e, shots in zip(engines, shot_lists):
        a = (abi.Shot * max(1, len(shots)))()
        for k, sh in enumerate(shots):
            s, w = _shot_struct(sh, e.p.nt)
            a[k] = s
            keep.append(w)
        arrs.append(a)
    sp = (C.POINTER(abi.Shot) * n)(*[C.cast(a, C.POINTER(abi.Shot)) for a in arrs])
    ns = (C.c_int32 * n)(*[len(s) for s in shot_lists])
    outs = [{"stack_pp": np.zeros((e.p.nx, e.p.nz), np.float32), "stack_ps": np.zeros((e.p.nx, e.p.nz), np.float32)}
            for e in engines]
    fp = C.POINTER(C.c_float)
    pp = (fp * n)(*[abi.fptr(o["stack_pp"]) for o in outs])
    ps = (fp * n)(*[abi.fptr(o["stack_ps"]) for o in outs])
    engines[0]._check(lib().ouc_rtm_grouped(hs, n, sp, ns, pp, ps))
    return outs


def rtm_concurrent(engines, shots, per_shot=False):
    """One job's shots on several engines of the same grid at once
    (ouc_rtm_concurrent): batch j runs on engines[j % n], each engine on its
    own stream.  Returns dict(stack_pp, stack_ps), bit for bit the stacks of
    engines[0].rtm_batch(shots); OUC_NUMERIC_FALLBACK sets fell_back on the
    first engine and warns."""
    n = len(engines)
    hs = (C.c_void_p * n)(*[e._h for e in engines])
    m = len(shots)
    arr = (abi.Shot * max(1, m))()
    keep = []
    for k, sh in enumerate(shots):
        s, w = _shot_struct(sh, engines[0].p.nt)
        arr[k] = s
        keep.append(w)
    spp = np.zeros((engines[0].p.nx, engines[0].p.nz), np.float32)
    sps = np.zeros_like(spp)
    ms = np.zeros(max(1, m), np.float32)
    engines[0]._check(lib().ouc_rtm_concurrent(hs, n, arr, m, abi.fptr(spp), abi.fptr(sps), abi.fptr(ms)))
    out = {"stack_pp": spp, "stack_ps": sps}
    if per_shot:
        out["per_shot_ms"] = ms[:m]
    return out


def write_image_files(directory, pp_lap, ps_lap, shot):
    """Reference file format: <dir>/image_PP_lap_<nx>_<nz>_<shot>.dat (shot < 0: the stack names)."""
    pp = np.ascontiguousarray(pp_lap, np.float32)
    ps = np.ascontiguousarray(ps_lap, np.float32)
    st = lib().ouc_write_image_files(directory.encode(), pp.shape[0], pp.shape[1], int(shot), abi.fptr(pp),
                                     abi.fptr(ps))
    if st != abi.OUC_OK:
        raise OucError(st, lib().ouc_last_error(None).decode())


def strip_codec_fp16(x, device=0, decoded=False):
    """GPU scaled-FP16 strip codec on a host array.  A 1-D x is one slot; a
    2-D x [steps][n] is `steps` slots of one field stored in order (predictive
    coding with error feedback).  Returns (binary16 residual bits, chunk
    exponents) shaped like x and [steps][ceil(n/64)], and with decoded=True
    the values decoded backwards as the reconstruction does."""
    x = np.ascontiguousarray(x, np.float32)
    x2 = x.reshape(1, -1) if x.ndim == 1 else x
    steps, n = x2.shape
    h = np.zeros((steps, n), np.uint16)
    e = np.zeros((steps, (n + 63) // 64), np.int8)
    d = np.zeros((steps, n), np.float32)
    st = lib().ouc_strip_codec_fp16(int(device), abi.fptr(x2), n, steps, h.ctypes.data_as(C.POINTER(C.c_uint16)),
                                    e.ctypes.data_as(C.POINTER(C.c_int8)), abi.fptr(d))
    if st != abi.OUC_OK:
        raise OucError(st, lib().ouc_last_error(None).decode())
    out = (h, e, d) if decoded else (h, e)
    return tuple(a[0] for a in out) if x.ndim == 1 else out


def _io_check(st):
    if st != abi.OUC_OK:
        raise OucError(st, lib().ouc_last_error(None).decode())


def write_model_files(directory, vp=None, vs=None):
    """directory/Vp.dat and Vs.dat: raw float32 interiors, x-major."""
    arrs = [None if a is None else np.ascontiguousarray(a, np.float32) for a in (vp, vs)]
    shape = next(a.shape for a in arrs if a is not None)
    _io_check(lib().ouc_write_model_files(str(directory).encode(), shape[0], shape[1],
                                          *[abi.fptr(a) for a in arrs]))


def write_segy(path, traces, dt, shot=0, sx_m=0.0, sz_m=0.0, rec_x0_m=0.0, rec_dx_m=1.0, rec_z_m=0.0,
               component="VX"):
    """One component's shot gather [ntraces][nt] as a SEG-Y revision 1 file."""
    tr = np.ascontiguousarray(traces, np.float32)
    g = abi.SegyGather(int(shot), tr.shape[0], tr.shape[1], float(dt), float(sx_m), float(sz_m), float(rec_x0_m),
                       float(rec_dx_m), float(rec_z_m), component.encode()[:7])
    _io_check(lib().ouc_write_segy(str(path).encode(), C.byref(g), abi.fptr(tr)))


def read_segy(path):
    """(traces [ntraces][nt] float32, dt seconds) of a fixed-length IEEE-float SEG-Y file."""
    n, ns, dt = C.c_int32(), C.c_int32(), C.c_float()
    _io_check(lib().ouc_read_segy(str(path).encode(), C.byref(n), C.byref(ns), C.byref(dt), None, 0))
    out = np.zeros((n.value, ns.value), np.float32)
    _io_check(lib().ouc_read_segy(str(path).encode(), C.byref(n), C.byref(ns), C.byref(dt), abi.fptr(out), out.size))
    return out, dt.value


def stream_probe(nin, nout, n, reps=10, nshared=0, m=0, width=1, device=0):
    """Mean launch time (ms) of a kernel reading nin and writing nout arrays of
    n floats, `width` consecutive floats per thread; the last nshared inputs
    hold m floats, read at index mod m."""
    ms = np.zeros(1, np.float32)
    _io_check(lib().ouc_stream_probe(int(device), int(nin), int(nout), int(n), int(nshared), int(m), int(width),
                                     int(reps), abi.fptr(ms)))
    return float(ms[0])


def model_update(v, grad, step=30.0, iteration=0, device=0):
    """revise_model_end on the GPU: v - step * 0.95^iteration * grad, in double,
    rounded to float."""
    v = np.ascontiguousarray(v, np.float32)
    g = np.ascontiguousarray(grad, np.float32)
    assert v.shape == g.shape
    out = np.zeros_like(v)
    st = lib().ouc_model_update(int(device), v.size, abi.fptr(v), abi.fptr(g), float(step), int(iteration),
                                abi.fptr(out))
    if st != abi.OUC_OK:
        raise OucError(st, lib().ouc_last_error(None).decode())
    return out


def _lb_check(st):
    if st != abi.OUC_OK:
        raise OucError(st, lib().ouc_last_error(None).decode())


def lb_forward(x, weights, device=0):
    """Layered boundary network on the GPU: x [M][5][H][W] -> [M][2][H][W]."""
    x = np.ascontiguousarray(x, np.float32)
    w = np.ascontiguousarray(weights, np.float32).ravel()
    M, Cn, H, W = x.shape
    assert Cn == 5
    out = np.zeros((M, 2, H, W), np.float32)
    _lb_check(lib().ouc_lb_forward(int(device), M, H, W, abi.fptr(x), abi.fptr(w), abi.fptr(out)))
    return out


def lb_forward_device(M, H, W, x_ptr, weights_ptr, out_ptr, device=0, stream=None):
    """ouc_lb_forward_device on device addresses (integers, e.g. a torch
    tensor's data_ptr()): queues the forward pass on `stream` (a hipStream_t
    handle, None for the null stream) without allocating or waiting."""
    _lb_check(lib().ouc_lb_forward_device(int(device), int(M), int(H), int(W), C.c_void_p(x_ptr),
                                          C.c_void_p(weights_ptr), C.c_void_p(out_ptr),
                                          C.c_void_p(stream) if stream else None))


def lb_loss_grad(x, y, mask, weights, device=0):
    """(loss, gradient) of the masked MSE of the boundary network on the GPU."""
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    mk = np.ascontiguousarray(mask, np.float32)
    w = np.ascontiguousarray(weights, np.float32).ravel()
    M, _, H, W = x.shape
    loss = np.zeros(1, np.float32)
    grad = np.zeros_like(w)
    _lb_check(lib().ouc_lb_loss_grad(int(device), M, H, W, abi.fptr(x), abi.fptr(y), abi.fptr(mk), abi.fptr(w),
                                     abi.fptr(loss), abi.fptr(grad)))
    return float(loss[0]), grad


def lb_train(x, y, mask, weights, epochs, lr, device=0):
    """Adam on the GPU; returns (weights, per-epoch losses)."""
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    mk = np.ascontiguousarray(mask, np.float32)
    w = np.ascontiguousarray(weights, np.float32).ravel().copy()
    M, _, H, W = x.shape
    hist = np.zeros(int(epochs), np.float32)
    _lb_check(lib().ouc_lb_train(int(device), M, H, W, abi.fptr(x), abi.fptr(y), abi.fptr(mk), int(epochs),
                                 float(lr), abi.fptr(w), abi.fptr(hist)))
    return w, hist


def device_arch(device=0):
    buf = C.create_string_buffer(64)
    st = lib().ouc_device_arch(int(device), buf, 64)
    if st != abi.OUC_OK:
        raise OucError(st, lib().ouc_last_error(None).decode())
    return buf.value.decode()
