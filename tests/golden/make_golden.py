"""Writes tests/golden/path_40x40.npz: FP64 oracle outputs of one small case
through the whole path (forward seismograms, RTM images in both semantics).

Case: 40 x 40 interior, L = 12, N = 6 (the reference's literal coefficients,
P-S-seperating.cu:221), dt = 0.5 ms, dx = dz = 5 m, 250 steps, R = 1e-3, the
reference's three-layer model scaled to 40 cells (vp 3400/3800/4500 split at
z = 10, 20; vs 2900/3800 split at z = 20; P-S-seperating.cu:1999-2040), Ricker
30 Hz at interior (20, 0), receivers on row 0, direct-wave mute on.

Outputs are stored as float32 (the GPU bars are rel-L2 1e-5 / 1e-4; float32
storage adds < 1e-7).  Run from the repository root:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import pyoracle  # noqa: E402
from ouc_amd import abi  # noqa: E402

NX = NZ = 40


def case(semantics=abi.SEM_FAITHFUL, engine=abi.ENGINE_PLAIN):
    p = abi.make_params(NX, NZ, 12, 6, 250, 5.0, 5.0, 0.0005, R=1e-3, semantics=semantics, engine=engine)
    vp = np.empty((NX, NZ), np.float32)
    vs = np.empty((NX, NZ), np.float32)
    for j in range(NZ):
        vp[:, j] = 3400.0 if j < 10 else (3800.0 if j < 20 else 4500.0)
        vs[:, j] = 2900.0 if j < 20 else 3800.0
    shot = dict(sx=20, sz=0, rec_z=0, fm=30.0, mute=1)
    return p, vp, vs, shot


def main():
    out = {}
    p, vp, vs, shot = case()
    out["seis"] = pyoracle.forward_shot(p, vp, vs, shot, 64).astype(np.float32)
    for sem, tag in ((abi.SEM_FAITHFUL, "faithful"), (abi.SEM_EXACT, "exact")):
        p, vp, vs, shot = case(sem)
        r = pyoracle.rtm_shot(p, vp, vs, shot, 64)
        for k in ("img_pp", "img_ps", "lap_pp", "lap_ps"):
            out[f"{tag}_{k}"] = r[k].astype(np.float32)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "path_40x40.npz")
    np.savez_compressed(path, **out)
    print(path, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
