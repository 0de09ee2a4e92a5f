"""Writes the full-size FP64 image fixtures of the two RTM configurations.

    python tests/golden/make_config_fixtures.py [--threads T] [--only rtm1024|rtm2048]

rtm1024: 1024 x 512 at 5 m, 4000 steps, one shot at x = 512 (single-shot RTM).
rtm2048: 2048 x 1024 at 5 m, 6000 steps, shots s = 0 (x = 16) and s = 32
         (x = 1040) of the 64-shot layout x = 16 + 32 s (batched RTM).
Both: seeded six-layer models with a 3 % random perturbation
(ouc_amd.velocity_models), Ricker 20 Hz, dt 0.4 ms, N = 5, 30-cell CPML at
R = 1e-3, surface sources and receivers, muted records, FAITHFUL semantics.

The FP64 oracle (exact strips) runs each shot, from about ten minutes
(rtm1024) to about forty (rtm2048, per shot) on 8 host cores.  A full image
is 2-8 MB per array, so the fixture keeps, as float32:
  - the PP and PS images (num/den, image_fun, P-S-seperating.cu:1441-1455),
    their Laplacians (Laplace, P-S-seperating.cu:1457-1469, computed on the
    full image) and the source illumination den (fenmu_P of corr_v,
    P-S-seperating.cu:1426-1439) on the decimated grid x = 0, dx, 2 dx, ...,
    z = 0, dz, 2 dz, ... (dx = dz = 4 for rtm2048; dx = 4, dz = 2 for rtm1024);
  - rows 0..15 of the PP and PS images and of den at full x resolution.
A relative L2 over the decimated grid estimates the one over the full image;
the full-resolution top rows cover the receiver row and its neighbours.
The GPU tests compare the engine's images on the same grid
(tests/test_gpu_configs.py).
"""
import argparse
import ctypes as C
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

TOP_ROWS = 16
DECIMATION = {"rtm1024": (4, 2), "rtm2048": (4, 4)}


def workload(config):
    from ouc_amd import velocity_models as vm

    if config == "rtm1024":
        p, vp, vs, shot = vm.cfg_rtm_1024()
        return p, vp, vs, {0: shot}
    p, vp, vs, shots = vm.cfg_rtm_2048()
    return p, vp, vs, {0: shots[0], 32: shots[32]}


def oracle_images(p, vp, vs, shot):
    """FP64 images and accumulators (num_pp, num_ps, den) of one shot."""
    from oracle import pyoracle
    from ouc_amd import abi

    m = pyoracle._ModelRef(vp, vs)
    s = pyoracle._ShotRef(**shot)
    img_pp = np.zeros((p.nx, p.nz))
    img_ps = np.zeros((p.nx, p.nz))
    acc = np.zeros((3, p.nx, p.nz))
    null = C.POINTER(C.c_double)()
    rc = pyoracle.lib().oracle_rtm_trace(C.byref(p), C.byref(m.m), C.byref(s.s), 64, null, abi.dptr(img_pp),
                                         abi.dptr(img_ps), null, null, null, abi.dptr(acc), null)
    if rc != 0:
        raise RuntimeError(f"oracle rtm failed: {rc}")
    return img_pp, img_ps, acc


def make(config, out_dir):
    from oracle import pyoracle

    p, vp, vs, shots = workload(config)
    ddx, ddz = DECIMATION[config]
    arrays = {"decimation": np.array([ddx, ddz], np.int32), "top_rows": np.array(TOP_ROWS, np.int32)}
    for s, shot in shots.items():
        t0 = time.time()
        img_pp, img_ps, acc = oracle_images(p, vp, vs, shot)
        for k, img in (("pp", img_pp), ("ps", img_ps)):
            lap = pyoracle.laplace(np.where(np.isfinite(img), img, 0.0), p.dx)
            arrays[f"shot{s}_img_{k}"] = img[::ddx, ::ddz].astype(np.float32)
            arrays[f"shot{s}_lap_{k}"] = lap[::ddx, ::ddz].astype(np.float32)
            arrays[f"shot{s}_top_{k}"] = img[:, :TOP_ROWS].astype(np.float32)
        arrays[f"shot{s}_den"] = acc[2][::ddx, ::ddz].astype(np.float32)
        arrays[f"shot{s}_top_den"] = acc[2][:, :TOP_ROWS].astype(np.float32)
        arrays[f"shot{s}_sx"] = np.array(shot["sx"], np.int32)
        print(f"[fixture] {config} shot {s} (x = {shot['sx']}): {time.time() - t0:.0f} s", flush=True)
    path = os.path.join(out_dir, f"{config}_fp64.npz")
    np.savez_compressed(path, **arrays)
    print(f"[fixture] wrote {path} ({os.path.getsize(path) / 1e6:.1f} MB)", flush=True)


def main():
    from oracle import pyoracle

    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--only", choices=["rtm1024", "rtm2048"], default=None)
    ap.add_argument("--out", default=HERE)
    a = ap.parse_args()
    pyoracle.build()
    pyoracle.set_threads(a.threads)
    for config in (a.only,) if a.only else ("rtm1024", "rtm2048"):
        make(config, a.out)


if __name__ == "__main__":
    main()
