"""Serving-style RTM on one GPU: many shots on uneven grids with uneven step
counts.

Jobs are grouped by grid (one engine context per grid and model).  Inside a
group the shots run as batches of the fused engine: the context holds the
group's largest step count, and a shorter shot starts late in the batch (its
wavelet is zero before its start and it images only its own steps), so every
shot's image equals the image of its own run.  Each group's batches replay
their captured time-loop graphs with a tile from the autotuner's cache.  The
groups run one after another ("sjf"), concurrently from one host thread and
stream each ("concurrent"), or together in shared launches ("grouped",
ouc_rtm_grouped): every kernel of a time step is one launch over the shots of
all grids, so small grids fill the GPU beside large ones without competing
streams.

Latency of a shot = time from the common start to the end of the batch that
holds it; throughput = shots / wall time of the whole job.
"""
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import abi, tuning
from .engine import Engine, rtm_grouped
from .velocity_models import layered_random

GRIDS = ((512, 256), (1024, 512), (2048, 1024), (4096, 2048))
# Small grids and short runs: launches are short and one grid alone does not
# fill the GPU, where shared launches over all grids pay most.
SMALL_GRIDS = ((96, 48), (112, 56), (128, 64), (160, 80))
SMALL_NT = (200, 420)
@dataclass
class Job:
    grid: tuple
    nt: int
    shot: dict
    id: int = 0


@dataclass
class GroupResult:
    grid: tuple
    nshots: int
    nt_max: int
    tile: tuple
    wall_s: float
    fallback_shots: int
    stack_pp: np.ndarray = field(repr=False, default=None)


def make_jobs(n=32, seed=5, grids=GRIDS, nt_range=(2000, 8000), fm=20.0):
    """n shots, grid drawn uniformly from `grids`, step count uniform in
    nt_range, source anywhere on the surface row (seeded)."""
    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(n):
        g = tuple(grids[int(rng.integers(0, len(grids)))])
        nt = int(rng.integers(nt_range[0], nt_range[1] + 1))
        sx = int(rng.integers(0, g[0]))
        jobs.append(Job(g, nt, dict(sx=sx, sz=0, rec_z=0, fm=fm, mute=1, nt=nt), k))
    return jobs


def group_params(grid, nt, learned=True, strips=abi.STRIP_FP16, dx=5.0, dt=4e-4, npml=30, band=8):
    boundary = abi.BOUNDARY_LEARNED_FALLBACK if learned else abi.BOUNDARY_CPML
    return abi.make_params(grid[0], grid[1], npml, 5, nt, dx, dx, dt, R=1e-3, engine=abi.ENGINE_FUSED,
                           strip_precision=strips, boundary=boundary, band=band if learned else 0)


def model_for(grid, seed):
    vp, vs, _ = layered_random(grid[0], grid[1], seed)
    return vp, vs


def serve(jobs, learned=True, strips=abi.STRIP_FP16, weights=None, tune=True, schedule="sjf", warm=True,
          model_seed=11, mem_budget_gb=200.0, log=None):
    """Run all jobs; returns (per-job latency ms [len(jobs)], wall s, [GroupResult]).

    schedule "sjf": groups run one after another, least work (cells x steps x
    shots) first -- small grids finish early instead of queueing behind large
    kernels; "concurrent": all groups at once, one host thread and stream each;
    "grouped": all groups in shared launches (one launch per kernel and time
    step over every grid's shots), a shot's latency being the end of the call.
    The groups' device buffers share `mem_budget_gb`: when all groups' shots do
    not fit at once, every group's batch shrinks by the same factor (the rest
    of its shots run as later batches)."""
    concurrent = schedule == "concurrent"
    groups = {}
    for j in jobs:
        groups.setdefault(j.grid, []).append(j)
    if learned and weights is None:
        from .learned import load_weights

        weights = load_weights()
    engines = {}
    for gi, (grid, js) in enumerate(sorted(groups.items())):
        js.sort(key=lambda j: -j.nt)  # longest first: each batch wastes the least
        p = group_params(grid, js[0].nt, learned, strips)
        vp, vs = model_for(grid, model_seed + gi)
        e = Engine(p, vp, vs)
        if learned:
            e.set_learned_boundary(weights)
        per_shot = e.set_batch_capacity(0)
        engines[grid] = [e, js, None, per_shot]
    need = sum(len(v[1]) * v[3] for v in engines.values())
    scale = min(1.0, mem_budget_gb * 1e9 / need) if need > 0 else 1.0
    for grid, v in engines.items():
        e, js, _, per_shot = v
        cap = max(1, int(len(js) * scale))
        e.set_batch_capacity(cap)
        tile = tuning.tuned(e, cap)[:2] if tune else e.tile()
        v[2] = tile
        if log:
            log(f"group {grid[0]}x{grid[1]}: {len(js)} shots, nt {js[-1].nt}..{js[0].nt}, batch {cap}, "
                f"{per_shot / 1e9:.1f} GB per shot, tile {tile}")
    lat = np.zeros(len(jobs))
    results = {}
    errors = []

    def run(grid, t_start):
        try:
            run1(grid, t_start)
        except Exception as ex:  # re-raised by the caller's thread
            errors.append(ex)

    def run1(grid, t_start):
        e, js, tile, _ = engines[grid]
        t0 = time.perf_counter()
        out = e.rtm_batch([j.shot for j in js], per_shot=True)
        wall = time.perf_counter() - t0
        off = (t0 - t_start) * 1e3
        for j, ms in zip(js, out["per_shot_ms"]):
            lat[j.id] = off + float(ms)
        results[grid] = GroupResult(grid, len(js), js[0].nt, tile, wall, e.timings()["fallback_shots"],
                                    out["stack_pp"])

    def run_grouped(t_start):
        order = sorted(engines)
        es = [engines[g][0] for g in order]
        outs = rtm_grouped(es, [[j.shot for j in engines[g][1]] for g in order])
        wall = time.perf_counter() - t_start
        for g, e, out in zip(order, es, outs):
            js = engines[g][1]
            for j in js:
                lat[j.id] = wall * 1e3
            results[g] = GroupResult(g, len(js), js[0].nt, engines[g][2], wall, e.timings()["fallback_shots"],
                                     out["stack_pp"])

    def run_all():
        t_start = time.perf_counter()
        if schedule == "grouped":
            run_grouped(t_start)
        elif concurrent:
            th = [threading.Thread(target=run, args=(g, t_start)) for g in engines]
            for t in th:
                t.start()
            for t in th:
                t.join()
        else:
            work = lambda g: g[0] * g[1] * sum(j.nt for j in engines[g][1])
            for g in sorted(engines, key=work):
                run(g, t_start)
                if errors:
                    break
        if errors:
            raise errors[0]
        return time.perf_counter() - t_start

    try:
        if warm:
            run_all()  # captures the graphs; the timed pass replays them
            if log:
                log("warm pass done")
        wall = run_all()
    finally:
        for v in engines.values():
            v[0].close()
    if len(results) != len(engines):
        raise RuntimeError("a serving group failed")
    return lat, wall, [results[g] for g in sorted(results)]


def summary(jobs, lat, wall, groups):
    return {
        "shots": len(jobs),
        "wall_s": wall,
        "shots_per_hour": len(jobs) / wall * 3600.0,
        "latency_ms_p50": float(np.percentile(lat, 50)),
        "latency_ms_p95": float(np.percentile(lat, 95)),
        "groups": [{"grid": f"{g.grid[0]}x{g.grid[1]}", "shots": g.nshots, "nt_max": g.nt_max,
                    "tile": list(g.tile), "wall_s": g.wall_s, "fallback_shots": g.fallback_shots} for g in groups],
    }
