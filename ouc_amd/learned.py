"""Learned absorbing boundary: training data from the engine, training on the
GPU, and the reflected-energy measure.

Method (no counterpart in the reference, whose boundary is the CPML of
P-S-seperating.cu:2067-2105):

* The outer `band` cells of the padded grid run second-order staggered
  differences, so only the outermost ring of velocity cells lacks the data of a
  full update (the liveness rules are mirror-symmetric on the staggered grid;
  see ouc_amd/csrc/lpad_kernels.hpp).
* A three-layer 3x3 network (5 -> 16 -> 16 -> 2 channels, linear, BF16
  weights) sets that ring.  Inputs, in the frame of each side (a: cells from
  the outer edge, b: along it): the velocity increment D = V^n - V^{n-1} (the
  time derivative of the field the second-order leapfrog integrates) and the
  stresses txx = theta - 2 q gzz, tzz = theta - 2 q gxx, txz = q (gxz + gzx)
  formed from V^n (scaled by dt^2/dx).  Output: the increment correction of
  the ring cells (normal and tangential component): V^{n+1} = 2V^n - V^{n-1} +
  delta.
* Targets come from runs of the engine itself on a domain extended 3x (the
  wave simply continues there): delta = V_ext^{n+1} - (2 V^n - V^{n-1}) at the
  ring, with the inputs formed from the extended run's V cropped to the small
  grid.  Gaussian noise added to the pad's V^n and V^{n-1} before forming
  inputs and targets teaches the network to damp what it did not see in the
  training wavefields; without it the one-step fit is accurate and the rollout
  still grows.
* The measure of quality is the energy of the difference between the small-grid run and the
  extended run over the interior, relative to the extended run's energy, in dB.
"""
import os

import numpy as np

from . import abi
from .engine import Engine, lb_loss_grad, lb_train

N_WEIGHTS = 16 * 5 * 9 + 16 * 16 * 9 + 2 * 16 * 9
WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "learned_boundary.npz")

# channel staggering (x, z), mirror signs, and the x <-> z transpose map of the
# network's five channels: D_x, D_z, txx, tzz, txz (the kernels use the same)
SX = (1, 0, 0, 0, 1)
SZ = (0, 1, 0, 0, 1)
XS = (-1.0, 1.0, 1.0, 1.0, -1.0)
ZS = (1.0, -1.0, 1.0, 1.0, -1.0)
TSWAP = (1, 0, 3, 2, 4)
WIN_H = 7          # rows a = -3..3 of a window; the output row is a = 0 (row 3)
WIN_OUT = 64       # output cells along b per window


def init_weights(seed=0):
    """Random initial FP32 weights (scaled so the composed map starts small)."""
    rng = np.random.default_rng(seed)
    w1 = rng.standard_normal(16 * 5 * 9) / np.sqrt(5 * 9)
    w2 = rng.standard_normal(16 * 16 * 9) / np.sqrt(16 * 9)
    w3 = 0.1 * rng.standard_normal(2 * 16 * 9) / np.sqrt(16 * 9)
    return np.concatenate([w1, w2, w3]).astype(np.float32)


def save_weights(w, path=WEIGHTS_PATH, **meta):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, weights=np.asarray(w, np.float32), **{k: np.asarray(v) for k, v in meta.items()})


def load_weights(path=WEIGHTS_PATH):
    with np.load(path, allow_pickle=False) as f:
        return np.array(f["weights"], np.float32)


# ---------------------------------------------------------------------------
# Inputs on a padded grid (second-order pad formulas; liveness as the kernels)
# ---------------------------------------------------------------------------
def pad_channels(V, Vm, p2, q2, dx, dz, dt):
    """The five network channels over a whole padded grid from V^n, V^{n-1}
    ([2][X][Z] each); only cells within the pad are meaningful."""
    vx, vz = V
    X, Z = vx.shape
    gxx = np.zeros((X, Z)); gzz = np.zeros((X, Z)); gxz = np.zeros((X, Z)); gzx = np.zeros((X, Z))
    gxx[1:X - 1, 1:Z - 1] = (vx[1:X - 1, 1:Z - 1] - vx[0:X - 2, 1:Z - 1]) / dx
    gzz[1:X - 1, 1:Z - 1] = (vz[1:X - 1, 1:Z - 1] - vz[1:X - 1, 0:Z - 2]) / dz
    gxz[0:X - 1, 0:Z - 1] = (vx[0:X - 1, 1:Z] - vx[0:X - 1, 0:Z - 1]) / dz
    gzx[0:X - 1, 0:Z - 1] = (vz[1:X, 0:Z - 1] - vz[0:X - 1, 0:Z - 1]) / dx
    th = p2 * (gxx + gzz)
    sc = dt * dt / dx
    return [V[0] - Vm[0], V[1] - Vm[1], sc * (th - 2 * q2 * gzz), sc * (th - 2 * q2 * gxx), sc * q2 * (gxz + gzx)]


def to_side(ch, side):
    """Channel arrays [5][X][Z] -> the frame of `side` (0: x-low, 1: x-high,
    2: z-low, 3: z-high): [5][A][B] with the outward normal at a = 0."""
    X, Z = ch[0].shape
    out = []
    for c in range(5):
        if side in (0, 1):
            src, sx, sg = ch[c], SX[c], XS[c]
            if side == 0:
                out.append(np.array(src))
            else:
                a = np.zeros_like(src)
                n = X - sx
                a[:n] = sg * src[X - 1 - sx - np.arange(n)]
                out.append(a)
        else:
            o = TSWAP[c]
            src = ch[o].T
            if side == 2:
                out.append(np.array(src))
            else:
                a = np.zeros_like(src)
                n = Z - SZ[o]
                a[:n] = ZS[o] * src[Z - 1 - SZ[o] - np.arange(n)]
                out.append(a)
    return out


def ring_side(f, i, j, X, Z):
    """Side a ring cell of field f (0: vx, 1: vz) belongs to: smallest
    canonical depth, ties to the lower side index (as the ring kernel)."""
    d = [i, X - 1 - SX[f] - i, j, Z - 1 - SZ[f] - j]
    return int(np.argmin(d))


def side_cell(side, f_canon, b, X, Z):
    """Original (field, i, j, sign) of canonical output channel f_canon at
    (a = 0, b) on `side`."""
    if side < 2:
        f = f_canon
        i = 0 if side == 0 else X - 1 - SX[f] - 0
        return f, i, b, (1.0 if side == 0 else XS[f])
    f = TSWAP[f_canon]
    j = 0 if side == 2 else Z - 1 - SZ[f]
    return f, b, j, (1.0 if side == 2 else ZS[f])


def windows_from_state(Vm, V, Vn1, p2, q2, dx, dz, dt, L, noise, rng):
    """Training windows of one time level.  Vm, V, Vn1: V^{n-1}, V^n,
    V^{n+1} of the extended run cropped to the small padded grid [2][X][Z].
    Returns x [M][5][7][W], y [M][2][7][W], mask [M][7][W]."""
    X, Z = V.shape[1:]
    ii, jj = np.meshgrid(np.arange(X), np.arange(Z), indexing="ij")
    pad = (np.minimum(np.minimum(ii, X - 1 - ii), np.minimum(jj, Z - 1 - jj)) < L).astype(np.float64)
    scale = max(np.abs(V).max(), 1e-30)
    Vt = V.astype(np.float64)
    Vmt = Vm.astype(np.float64)
    if noise > 0:
        Vt = Vt + noise * scale * rng.standard_normal(Vt.shape) * pad
        Vmt = Vmt + noise * scale * rng.standard_normal(Vmt.shape) * pad
    ch = pad_channels(Vt, Vmt, p2, q2, dx, dz, dt)
    delta = Vn1 - (2 * Vt - Vmt)  # the ring's target increment correction
    W = WIN_OUT + 6
    xs, ys, ms = [], [], []
    for side in range(4):
        cs = to_side(ch, side)
        A, Bn = cs[0].shape
        # targets and mask along b at a = 0
        tgt = np.zeros((2, Bn))
        ok = np.ones(Bn, bool)
        for fc in range(2):
            for b in range(Bn):
                f, i, j, sg = side_cell(side, fc, b, X, Z)
                ghost = (f == 0 and i == X - 1) or (f == 1 and j == Z - 1)
                if ghost or ring_side(f, i, j, X, Z) != side:
                    ok[b] = False
                    continue
                tgt[fc, b] = sg * delta[f, i, j]
        for b0 in range(0, Bn, WIN_OUT):
            x = np.zeros((5, WIN_H, W))
            lo, hi = b0 - 3, b0 + WIN_OUT + 3
            blo, bhi = max(lo, 0), min(hi, Bn)
            for c in range(5):
                x[c, 3:7, blo - lo:bhi - lo] = cs[c][0:4, blo:bhi]
            y = np.zeros((2, WIN_H, W))
            m = np.zeros((WIN_H, W))
            e = min(b0 + WIN_OUT, Bn)
            y[:, 3, 3:3 + e - b0] = tgt[:, b0:e]
            m[3, 3:3 + e - b0] = ok[b0:e]
            if m.sum() == 0:
                continue
            xs.append(x / scale); ys.append(y / scale); ms.append(m)
    return np.array(xs, np.float32), np.array(ys, np.float32), np.array(ms, np.float32)


# ---------------------------------------------------------------------------
# Models and runs
# ---------------------------------------------------------------------------
def random_layered(nx, nz, rng, vmin=1500.0, vmax=4500.0):
    """Flat layers along z with vp uniform in [vmin, vmax], vs = vp / r, r in [1.6, 2.0]."""
    nl = int(rng.integers(3, 7))
    bounds = np.sort(rng.integers(0, nz, nl - 1))
    vals = rng.uniform(vmin, vmax, nl)
    col = np.empty(nz)
    k = 0
    for j in range(nz):
        while k < nl - 1 and j >= bounds[k]:
            k += 1
        col[j] = vals[k]
    vp = np.repeat(col[None, :], nx, axis=0).astype(np.float32)
    vs = (vp / rng.uniform(1.6, 2.0)).astype(np.float32)
    return vp, vs


def edge_pad(a, L):
    return np.pad(a, L, mode="edge")


class Case:
    """One random case: an n x n interior inside a 3n x 3n extended domain."""

    def __init__(self, n, seed, nt=None, dx=5.0, dt=4e-4, N=5, band=8, ext_pml=40, fm=None):
        rng = np.random.default_rng(seed)
        self.n, self.N, self.band, self.ext_pml = n, N, band, ext_pml
        self.dx, self.dt = dx, dt
        self.nt = nt or 3 * n + 300
        self.vp_ext, self.vs_ext = random_layered(3 * n, 3 * n, rng)
        self.vp, self.vs = self.vp_ext[n:2 * n, n:2 * n].copy(), self.vs_ext[n:2 * n, n:2 * n].copy()
        self.fm = float(fm or rng.uniform(10.0, 40.0))
        self.sx = int(rng.integers(n // 8, n - n // 8))
        self.sz = int(rng.integers(n // 8, n - n // 8))
        self.rng = rng

    def shot(self, offset=0):
        return {"sx": self.sx + offset, "sz": self.sz + offset, "rec_z": self.sz + offset, "fm": self.fm}

    def ext_params(self):
        return abi.make_params(3 * self.n, 3 * self.n, self.ext_pml, self.N, self.nt, self.dx, self.dx, self.dt,
                               R=1e-3, engine=abi.ENGINE_FUSED)

    def small_params(self, boundary=abi.BOUNDARY_LEARNED, npml=None):
        if boundary == abi.BOUNDARY_CPML:
            return abi.make_params(self.n, self.n, npml or self.band, self.N, self.nt, self.dx, self.dx, self.dt,
                                   R=1e-3, engine=abi.ENGINE_FUSED)
        return abi.make_params(self.n, self.n, npml or 40, self.N, self.nt, self.dx, self.dx, self.dt, R=1e-3,
                               engine=abi.ENGINE_FUSED, boundary=boundary, band=self.band)

    def crop_origin(self):
        """Extended padded cell of the small padded grid's cell (0, 0)."""
        return self.ext_pml + self.n - self.band


def case_windows(case, n_times, noise, device=0):
    """Training windows of one case from the engine's extended-domain run."""
    L = case.band
    X = case.n + 2 * L
    t_lo = case.nt // 10
    times = np.unique(np.linspace(t_lo, case.nt - 2, n_times).astype(int))
    steps = sorted(set(int(s) for t in times for s in (t - 2, t - 1, t)))
    idx = {s: k for k, s in enumerate(steps)}
    o = case.crop_origin()
    with Engine(case.ext_params(), case.vp_ext, case.vs_ext, device=device) as e:
        snaps = e.forward_snapshots(case.shot(case.n), steps, o, o, X, X)
    p2 = edge_pad(case.vp.astype(np.float64), L) ** 2
    q2 = edge_pad(case.vs.astype(np.float64), L) ** 2
    xs, ys, ms = [], [], []
    for t in times:
        # snapshot after step s holds V^{s+1}: V^{t-1}, V^t, V^{t+1}
        Vm, V, Vn1 = snaps[idx[t - 2]], snaps[idx[t - 1]], snaps[idx[t]]
        x, y, m = windows_from_state(Vm, V, Vn1, p2, q2, case.dx, case.dx, case.dt, L, noise, case.rng)
        if len(x):
            xs.append(x); ys.append(y); ms.append(m)
    return np.concatenate(xs), np.concatenate(ys), np.concatenate(ms)


def training_set(n=512, cases=4, n_times=16, noise=2e-3, seed=1000, device=0, log=None):
    xs, ys, ms = [], [], []
    for k in range(cases):
        c = Case(n, seed + k)
        x, y, m = case_windows(c, n_times, noise, device)
        xs.append(x); ys.append(y); ms.append(m)
        if log:
            log(f"case {k}: vp {c.vp_ext.min():.0f}-{c.vp_ext.max():.0f} m/s, fm {c.fm:.1f} Hz, {len(x)} windows")
    return np.concatenate(xs), np.concatenate(ys), np.concatenate(ms)


def train(x, y, m, epochs=3000, lr=2e-3, seed=0, device=0, w0=None):
    """Adam on the GPU (the project's own forward/backward kernels)."""
    w = init_weights(seed) if w0 is None else np.asarray(w0, np.float32)
    return lb_train(x, y, m, w, epochs, lr, device)


# ---------------------------------------------------------------------------
# Quality measure: reflected energy against the extended run
# ---------------------------------------------------------------------------
def interior_snapshots(e, case, steps, origin, pad):
    return e.forward_snapshots(case.shot(origin - pad), steps, origin, origin, case.n, case.n)


def reflection_db(case, weights, variants=("learned", "no_network", "cpml_band", "cpml_40"), stride=8, device=0):
    """10 log10( sum |V - V_ext|^2 / sum |V_ext|^2 ) over the interior and the
    steps [0, nt) every `stride`, for each variant on the small grid."""
    steps = list(range(0, case.nt, stride))
    with Engine(case.ext_params(), case.vp_ext, case.vs_ext, device=device) as e:
        ref = e.forward_snapshots(case.shot(case.n), steps, case.ext_pml + case.n, case.ext_pml + case.n,
                                  case.n, case.n).astype(np.float64)
    eref = float(np.sum(ref ** 2))
    out = {}
    for v in variants:
        if v in ("learned", "no_network"):
            p = case.small_params(abi.BOUNDARY_LEARNED)
            pad = case.band
        else:
            pad = case.band if v == "cpml_band" else 40
            p = case.small_params(abi.BOUNDARY_CPML, npml=pad)
        with Engine(p, case.vp, case.vs, device=device) as e:
            if v == "learned":
                e.set_learned_boundary(weights)
            snap = e.forward_snapshots(case.shot(), steps, pad, pad, case.n, case.n).astype(np.float64)
        err = float(np.sum((snap - ref) ** 2))
        out[v] = 10.0 * np.log10(err / eref) if np.isfinite(err) else float("inf")
    return out


def main():
    import argparse
    import json
    import sys
    import time

    ap = argparse.ArgumentParser(description="Train the learned absorbing boundary on the GPU")
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--cases", type=int, default=4)
    ap.add_argument("--times", type=int, default=16)
    ap.add_argument("--noise", type=float, default=2e-3)
    ap.add_argument("--epochs", type=int, default=3000)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--eval-n", type=int, default=256)
    ap.add_argument("--eval-cases", type=int, default=2)
    ap.add_argument("--out", default=WEIGHTS_PATH)
    a = ap.parse_args()
    log = lambda s: print(f"[learned] {s}", file=sys.stderr, flush=True)
    t0 = time.time()
    x, y, m = training_set(a.n, a.cases, a.times, a.noise, log=log)
    log(f"{len(x)} windows of {x.shape[1:]} in {time.time() - t0:.1f} s")
    t1 = time.time()
    w, hist = train(x, y, m, a.epochs, a.lr)
    log(f"trained {a.epochs} epochs in {time.time() - t1:.1f} s: loss {hist[0]:.3e} -> {hist[-1]:.3e}")
    save_weights(w, a.out, loss=hist, n=a.n, cases=a.cases, times=a.times, noise=a.noise, epochs=a.epochs,
                 lr=a.lr, windows=len(x))
    res = []
    for k in range(a.eval_cases):
        c = Case(a.eval_n, 9000 + k)
        r = reflection_db(c, w)
        res.append(r)
        log(f"eval case {k} (vp {c.vp_ext.min():.0f}-{c.vp_ext.max():.0f}, fm {c.fm:.1f}): " +
            ", ".join(f"{kk} {vv:.1f} dB" for kk, vv in r.items()))
    print(json.dumps({"windows": len(x), "loss_first": float(hist[0]), "loss_last": float(hist[-1]),
                      "eval": res}))


if __name__ == "__main__":
    main()
