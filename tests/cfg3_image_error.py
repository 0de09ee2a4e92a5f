"""Image error of the batched 2048 x 1024 RTM against the FP64 oracle on two
shots (s = 0 and s = 32 of the 64-shot layout x = 16 + 32 s).

A checker run by hand, in three phases, because the FP64 oracle takes about
15 minutes per shot on 8 host cores and the GPU box has no copy of its output:

  python tests/cfg3_image_error.py oracle  --out DIR   # CPU: FP64 oracle images (float64 .npy)
  python tests/cfg3_image_error.py gpu     --out DIR [--strips fp16|fp32]
                                                       # GPU: fused engine images (float32 .npy)
  python tests/cfg3_image_error.py compare --ref DIR --gpu DIR

Workload: 2048 x 1024 interior at 5 m, N = 5, 30-cell CPML, six layers
1800-4000 m/s with a seeded 3 % random perturbation, Ricker 20 Hz, dt 0.4 ms,
6000 steps, 2048 receivers at the surface, FAITHFUL semantics.  The oracle
stores its strips exactly (FP64); the GPU stores them in FP16 (scaled, with
the encoder's error feedback) or FP32.  Errors are relative L2 over the whole
interior and below row 8 (the receiver-row rows carry the reference's own FP32
conditioning, which the FP32 oracle misses by more than the engine does), for
the raw PP/PS images and their 5-point Laplacians (computed here the same way
for both sides, P-S-seperating.cu:1457-1469).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

SHOTS = (0, 32)
KEYS = ("img_pp", "img_ps")


def workload():
    from ouc_amd import velocity_models as vm

    return vm.cfg_rtm_2048()


def run_oracle(out):
    from oracle import pyoracle

    pyoracle.set_threads(os.cpu_count() or 1)
    p, vp, vs, shots = workload()
    for s in SHOTS:
        t0 = time.time()
        r = pyoracle.rtm_shot(p, vp, vs, shots[s], 64)
        for k in KEYS:
            np.save(os.path.join(out, f"shot{s}_{k}.npy"), r[k])
        print(f"[oracle] shot {s}: {time.time() - t0:.0f} s", flush=True)


def run_gpu(out, strips):
    from ouc_amd import abi
    from ouc_amd.engine import Engine

    p, vp, vs, shots = workload()
    p.engine = abi.ENGINE_FUSED
    p.strip_precision = abi.STRIP_FP16 if strips == "fp16" else abi.STRIP_FP32
    with Engine(p, vp, vs) as e:
        for s in SHOTS:
            print(f"[gpu] shot {s}, {strips} strips", flush=True)
            r = e.rtm_shot(shots[s])
            for k in KEYS:
                np.save(os.path.join(out, f"shot{s}_{k}.npy"), r[k].astype(np.float32))


def laplace5(img, dh=5.0):
    out = np.zeros_like(img)
    c = img[2:-2, 2:-2]
    out[2:-2, 2:-2] = (img[3:-1, 2:-2] + img[1:-3, 2:-2] + img[2:-2, 3:-1] + img[2:-2, 1:-3] - 4 * c) / dh ** 2
    return out


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def compare(ref, gpu):
    res = {}
    for s in SHOTS:
        for k in KEYS:
            a = np.load(os.path.join(gpu, f"shot{s}_{k}.npy")).astype(np.float64)
            b = np.load(os.path.join(ref, f"shot{s}_{k}.npy"))
            la, lb = laplace5(a), laplace5(b)
            res[f"shot{s}_{k}"] = {"whole": rel(a, b), "below_row8": rel(a[:, 8:], b[:, 8:]),
                                   "lap_whole": rel(la, lb), "lap_below_row8": rel(la[:, 8:], lb[:, 8:])}
    print(json.dumps(res, indent=1))
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("phase", choices=["oracle", "gpu", "compare"])
    ap.add_argument("--out", default=".")
    ap.add_argument("--ref", default="")
    ap.add_argument("--gpu", default="")
    ap.add_argument("--strips", choices=["fp16", "fp32"], default="fp16")
    a = ap.parse_args()
    if a.phase == "oracle":
        os.makedirs(a.out, exist_ok=True)
        run_oracle(a.out)
    elif a.phase == "gpu":
        os.makedirs(a.out, exist_ok=True)
        run_gpu(a.out, a.strips)
    else:
        compare(a.ref, a.gpu)


if __name__ == "__main__":
    main()
