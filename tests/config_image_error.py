"""Image error of full-size RTM configurations against the FP64 oracle.

A checker run by hand, in three phases, because the FP64 oracle takes minutes
(1024 x 512, one shot) to about half an hour (2048 x 1024, per shot) on 8 host
cores and the GPU box has no copy of its output:

  python tests/config_image_error.py oracle  --config C --out DIR [--precision 64|32]
                                              # CPU: oracle images (.npy)
  python tests/config_image_error.py gpu     --config C --out DIR [--engine fused|plain] [--strips fp16|fp32]
                                              # GPU: engine images (float32 .npy)
  python tests/config_image_error.py compare --config C --ref DIR --gpu DIR

Configurations (seeded layered models with a 3 % random perturbation, Ricker
20 Hz, dt 0.4 ms, N = 5, 30-cell CPML, surface sources and receivers, FAITHFUL
semantics):
  rtm1024: 1024 x 512 at 5 m, 4000 steps, one shot at x = 512;
  rtm2048: 2048 x 1024 at 5 m, 6000 steps, shots s = 0 and s = 32 of the
           64-shot layout x = 16 + 32 s.
The oracle stores its strips exactly (FP64, or FP32 at --precision 32: the
reference's own arithmetic in FP32).  Errors are relative L2 over the
interior below row 8 (the rows next to the receiver row carry the reference's
FP32 conditioning, which the FP32 oracle misses by more than the engine does),
for the raw PP/PS images and their 5-point Laplacians (P-S-seperating.cu:
1457-1469, computed here the same way for both sides), in bands of horizontal
distance from the source.  Far from the source the image num/den divides two
weakly illuminated correlations and its relative error grows for any FP32
arithmetic; the bands show where.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

KEYS = ("img_pp", "img_ps")
BANDS = ((0, 256), (256, 512), (512, 768), (768, 1024), (1024, 4096))


def workload(config):
    from ouc_amd import velocity_models as vm

    if config == "rtm1024":
        p, vp, vs, shot = vm.cfg_rtm_1024()
        return p, vp, vs, {0: shot}
    p, vp, vs, shots = vm.cfg_rtm_2048()
    return p, vp, vs, {0: shots[0], 32: shots[32]}


def run_oracle(config, out, precision):
    from oracle import pyoracle

    pyoracle.set_threads(os.cpu_count() or 1)
    p, vp, vs, shots = workload(config)
    for s, shot in shots.items():
        t0 = time.time()
        r = pyoracle.rtm_shot(p, vp, vs, shot, precision)
        for k in KEYS:
            np.save(os.path.join(out, f"shot{s}_{k}.npy"), r[k])
        print(f"[oracle FP{precision}] {config} shot {s}: {time.time() - t0:.0f} s", flush=True)


def run_gpu(config, out, engine, strips):
    from ouc_amd import abi
    from ouc_amd.engine import Engine

    p, vp, vs, shots = workload(config)
    p.engine = abi.ENGINE_FUSED if engine == "fused" else abi.ENGINE_PLAIN
    p.strip_precision = abi.STRIP_FP16 if strips == "fp16" else abi.STRIP_FP32
    with Engine(p, vp, vs) as e:
        for s, shot in shots.items():
            print(f"[gpu] {config} shot {s}, {engine} engine, {strips} strips", flush=True)
            r = e.rtm_shot(shot)
            for k in KEYS:
                np.save(os.path.join(out, f"shot{s}_{k}.npy"), r[k].astype(np.float32))


def laplace5(img, dh=5.0):
    out = np.zeros_like(img)
    c = img[2:-2, 2:-2]
    out[2:-2, 2:-2] = (img[3:-1, 2:-2] + img[1:-3, 2:-2] + img[2:-2, 3:-1] + img[2:-2, 1:-3] - 4 * c) / dh ** 2
    return out


def rel(a, b, m):
    return float(np.linalg.norm((a - b)[m]) / np.linalg.norm(b[m]))


def compare(config, ref, gpu):
    _, _, _, shots = workload(config)
    res = {}
    for s, shot in shots.items():
        for k in KEYS:
            a = np.load(os.path.join(gpu, f"shot{s}_{k}.npy")).astype(np.float64)
            b = np.load(os.path.join(ref, f"shot{s}_{k}.npy")).astype(np.float64)
            fin = np.isfinite(a) & np.isfinite(b)
            la, lb = laplace5(np.where(fin, a, 0.0)), laplace5(np.where(fin, b, 0.0))
            dist = np.abs(np.arange(a.shape[0]) - shot["sx"])[:, None] * np.ones((1, a.shape[1]))
            row = {"nonfinite_gpu": int((~np.isfinite(a)).sum()), "nonfinite_ref": int((~np.isfinite(b)).sum())}
            below = fin.copy()
            below[:, :8] = False
            row["below_row8"] = rel(a, b, below)
            for d0, d1 in BANDS:
                m = below & (dist >= d0) & (dist < d1)
                if m.any():
                    row[f"dist_{d0}_{d1}"] = rel(a, b, m)
                    row[f"lap_dist_{d0}_{d1}"] = rel(la, lb, m)
            res[f"shot{s}_{k}"] = row
    print(json.dumps(res, indent=1))
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("phase", choices=["oracle", "gpu", "compare"])
    ap.add_argument("--config", choices=["rtm1024", "rtm2048"], default="rtm2048")
    ap.add_argument("--out", default=".")
    ap.add_argument("--ref", default="")
    ap.add_argument("--gpu", default="")
    ap.add_argument("--precision", type=int, choices=[32, 64], default=64)
    ap.add_argument("--engine", choices=["fused", "plain"], default="fused")
    ap.add_argument("--strips", choices=["fp16", "fp32"], default="fp16")
    a = ap.parse_args()
    if a.phase == "oracle":
        os.makedirs(a.out, exist_ok=True)
        run_oracle(a.config, a.out, a.precision)
    elif a.phase == "gpu":
        os.makedirs(a.out, exist_ok=True)
        run_gpu(a.config, a.out, a.engine, a.strips)
    else:
        compare(a.config, a.ref, a.gpu)


if __name__ == "__main__":
    main()
