"""FWI gradient error at the 1024 x 512 configuration against the FP64 oracle.

A checker run by hand, in three phases (the FP64 oracle takes minutes on the
host cores and the GPU box has no copy of its output):

  python tests/config_fwi_error.py oracle  --out DIR [--config C]   # CPU: observed records and FP64 gradient (.npy)
  python tests/config_fwi_error.py gpu     --out DIR --ref DIR [--config C] [--strips fp16|fp32]
                                                                    # GPU: gradient from the same observed records
  python tests/config_fwi_error.py compare --ref DIR --gpu DIR [--config C]

Workloads (seeded layered models with a 3 % perturbation, N = 5, 30-cell
CPML, Ricker 20 Hz, dt 0.4 ms, one surface shot):
  rtm1024: 1024 x 512, 4000 steps, shot at x = 512;
  rtm2048: 2048 x 1024, 6000 steps, shot 32 of the 64-shot layout (x = 1040).
The observed records come from the same model with a +5 % vp lens (radius
40 cells, below the shot at z = 300); the gradient is taken in the
unperturbed model, so the residual is the lens's scattered wavefield.  Errors are relative L2 over the interior below row 8 and
in bands of horizontal distance from the source, as for the images.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

KEYS = ("grad_vp", "grad_vs", "illum")
BANDS = ((0, 128), (128, 256), (256, 512))


def workload(config="rtm1024"):
    from ouc_amd import velocity_models as vm

    if config == "rtm1024":
        p, vp, vs, shot = vm.cfg_rtm_1024()
    else:
        p, vp, vs, shots = vm.cfg_rtm_2048()
        shot = shots[32]
    shot = dict(shot, mute=0)
    x, z = np.meshgrid(np.arange(p.nx), np.arange(p.nz), indexing="ij")
    lens = ((x - shot["sx"]) ** 2 + (z - 300) ** 2) < 40 ** 2
    vpt = np.where(lens, vp * np.float32(1.05), vp).astype(np.float32)
    return p, vp, vs, vpt, shot


def run_oracle(out, config):
    from oracle import pyoracle

    pyoracle.set_threads(os.cpu_count() or 1)
    p, vp, vs, vpt, shot = workload(config)
    t0 = time.time()
    obs = pyoracle.forward_shot(p, vpt, vs, shot)[:2]  # vx, vz records (no mute)
    np.save(os.path.join(out, "observed.npy"), obs.astype(np.float32))
    obs32 = obs.astype(np.float32).astype(np.float64)  # the GPU receives float32 records
    r = pyoracle.fwi_gradient(p, vp, vs, shot, obs32)
    for k in KEYS:
        np.save(os.path.join(out, f"{k}.npy"), r[k])
    json.dump({"misfit": r["misfit"]}, open(os.path.join(out, "misfit.json"), "w"))
    print(f"[oracle FP64] fwi {config}: {time.time() - t0:.0f} s", flush=True)


def run_gpu(out, ref, strips, config):
    from ouc_amd import abi
    from ouc_amd.engine import Engine

    p, vp, vs, vpt, shot = workload(config)
    p.engine = abi.ENGINE_FUSED
    p.strip_precision = abi.STRIP_FP16 if strips == "fp16" else abi.STRIP_FP32
    obs = np.load(os.path.join(ref, "observed.npy"))
    print(f"[gpu] fwi {config}, {strips} strips", flush=True)
    with Engine(p, vp, vs) as e:
        r = e.fwi_gradient([shot], obs[None])
    for k in KEYS:
        np.save(os.path.join(out, f"{k}.npy"), r[k])
    json.dump({"misfit": r["misfit"]}, open(os.path.join(out, "misfit.json"), "w"))


def rel(a, b, m):
    return float(np.linalg.norm((a - b)[m]) / np.linalg.norm(b[m]))


def compare(ref, gpu, config):
    _, _, _, _, shot = workload(config)
    res = {}
    for k in KEYS:
        a = np.load(os.path.join(gpu, f"{k}.npy")).astype(np.float64)
        b = np.load(os.path.join(ref, f"{k}.npy"))
        dist = np.abs(np.arange(a.shape[0]) - shot["sx"])[:, None] * np.ones((1, a.shape[1]))
        below = np.ones(a.shape, bool)
        below[:, :8] = False
        row = {"whole": rel(a, b, np.ones(a.shape, bool)), "below_row8": rel(a, b, below)}
        for d0, d1 in BANDS:
            row[f"dist_{d0}_{d1}"] = rel(a, b, below & (dist >= d0) & (dist < d1))
        res[k] = row
    ma = json.load(open(os.path.join(gpu, "misfit.json")))["misfit"]
    mb = json.load(open(os.path.join(ref, "misfit.json")))["misfit"]
    res["misfit"] = abs(ma - mb) / mb
    print(json.dumps(res, indent=1))
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("phase", choices=["oracle", "gpu", "compare"])
    ap.add_argument("--out", default=".")
    ap.add_argument("--ref", default="")
    ap.add_argument("--gpu", default="")
    ap.add_argument("--strips", choices=["fp16", "fp32"], default="fp16")
    ap.add_argument("--config", choices=["rtm1024", "rtm2048"], default="rtm1024")
    a = ap.parse_args()
    if a.phase == "oracle":
        os.makedirs(a.out, exist_ok=True)
        run_oracle(a.out, a.config)
    elif a.phase == "gpu":
        os.makedirs(a.out, exist_ok=True)
        run_gpu(a.out, a.ref, a.strips, a.config)
    else:
        compare(a.ref, a.gpu, a.config)


if __name__ == "__main__":
    main()
