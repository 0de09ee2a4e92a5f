"""Benchmark: batched elastic RTM shots/hour at 2048 x 1024 on one MI355X.

Workload: 2048 x 1024 interior at 5 m (N = 5, 30-cell CPML, 2108 x 1084 padded),
six layers 1800-4000 m/s with a seeded 3 % random perturbation, Ricker 20 Hz,
dt = 0.4 ms, 6000 steps per shot, 2048 surface receivers, sources at
x = 16 + 32 s, boundary strips in scaled FP16 (FP32 compute).  A "step" is one
batch of --batch shots (8 by default: the 64-shot layout runs as eight such
batches) through the full RTM path (forward with strip saving, mute, backward
with reconstruction, receiver propagation and imaging, Laplacian, stack).
With --contexts C (3 by default) the job's batches run C at a time, each on
its own context and stream (ouc_rtm_concurrent), and the stack is still summed
in shot order.  Inputs are resident on the GPU before the timed region; value
= shots per hour of the whole job over the timed region's wall time.  A
wall-clock guard (--budget-s) shortens the warmup and the timed region when
they and the legs after them would not fit, and the line reports the steps
that ran.

Also reported: Gcell-updates/s, canonical and own-bytes HBM rates, the plain
(reference kernel structure, one shot at a time, FP32 strips) shots/hour on a
bounded sample, the roofline of the dominant step kernel from HIP events, and
the FP32 CPU baseline (the oracle) on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
T_START = time.perf_counter()


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=0,
                    help="shots per step (0: 8 fused, 1 plain)")
    ap.add_argument("--contexts", type=int, default=3,
                    help="fused engine: contexts (one stream and host thread each) that run the job's batches at "
                         "once through ouc_rtm_concurrent; 1 runs them one after another")
    ap.add_argument("--nt", type=int, default=6000)
    ap.add_argument("--engine", choices=["fused", "plain"], default="fused")
    ap.add_argument("--strips", choices=["fp16", "fp32"], default="fp16")
    ap.add_argument("--plain-sample-shots", type=int, default=1)
    ap.add_argument("--cpu-sample-s", type=float, default=10.0,
                    help="seconds of CPU work per sampled shot of the CPU baseline (two shots)")
    ap.add_argument("--budget-s", type=float, default=540.0,
                    help="wall-clock budget of the whole command: steps stop early (and the line says how many "
                         "ran) so that the line prints within it")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--tile", default="", help="core tile COLSxPREFETCH (e.g. 256x1) instead of the autotuner's")
    ap.add_argument("--serve", action="store_true",
                    help="uneven serving batch instead: 32 shots on grids 512x256..4096x2048, nt 2000-8000, "
                         "learned boundary with CPML fallback, FP16 strips, autotuned tiles")
    ap.add_argument("--serve-shots", type=int, default=32)
    ap.add_argument("--serve-schedule", choices=["sjf", "concurrent", "grouped"], default="sjf")
    ap.add_argument("--serve-boundary", choices=["learned", "cpml"], default="learned",
                    help="learned: 8-cell learned band with CPML fallback; cpml: 30-cell CPML")
    ap.add_argument("--serve-grids", choices=["full", "small"], default="full",
                    help="full: 512x256..4096x2048, nt 2000-8000; small: 96x48..160x80, nt 200-420")
    ap.add_argument("--fwi", action="store_true",
                    help="FWI gradient throughput (shots/hour) on the same workload instead of RTM")
    ap.add_argument("--split-pml", action="store_true",
                    help="forward modelling with the first-order split-PML engine (one shot per step)")
    ap.add_argument("--stream-probe", default="",
                    help="NINxNOUT[sS][wW]: streaming rate of a kernel reading NIN and writing NOUT arrays "
                         "(4.76 GB per launch in total, the forward core launch's compulsory bytes); with "
                         "sS the last S inputs are one 2048x1024 model plane read by every shot; wW moves W "
                         "consecutive floats per thread (1, 2, 4; default 1; -2: float2 memory ops transposed "
                         "through LDS, one value per thread for the arithmetic)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (counter calibration) and exit")
    return ap.parse_args()


def copy_probe(mib, reps=10):
    """A plain device-to-device copy of known size: under rocprofv3 --pmc its
    FETCH_SIZE / WRITE_SIZE per dispatch calibrate the counters (expected: mib
    MiB read and mib MiB written per launch)."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    print(json.dumps({"copy_probe_mib": mib, "reps": reps, "bytes_read_per_launch": n * 4,
                      "bytes_written_per_launch": n * 4, "gbs": 2 * n * 4 * reps / dt / 1e9}))


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    if a.copy_probe_mib:
        copy_probe(a.copy_probe_mib)
        return
    if a.stream_probe:
        from ouc_amd.engine import stream_probe

        import re

        spec = a.stream_probe.lower()
        mt = re.fullmatch(r"(\d+)x(\d+)(?:s(\d+))?(?:w(-?\d+))?", spec)
        if not mt:
            raise SystemExit(f"bad --stream-probe {spec}")
        nin, nout = int(mt.group(1)), int(mt.group(2))
        nshared, width = int(mt.group(3) or 0), int(mt.group(4) or 1)
        n = int(4.7613e9 / 4 / (nin + nout)) // 256 * 256
        plane = 2108 * 1088  # one pitched 2048x1024 plane with its 30-cell pad, floats
        ms = stream_probe(nin, nout, n, reps=20, nshared=nshared, m=plane if nshared else 0, width=width)
        print(json.dumps({"stream_probe": spec, "width": width, "floats_per_array": n, "shared_inputs": nshared,
                          "shared_floats": plane if nshared else 0, "bytes_per_launch": (nin + nout) * n * 4,
                          "ms_per_launch": ms, "gbs": (nin + nout) * n * 4 / ms / 1e6}))
        return
    if a.serve:
        serve_bench(a)
        return
    if a.fwi:
        fwi_bench(a)
        return
    if a.split_pml:
        splitpml_bench(a)
        return
    rtm_bench(a)


def _elapsed():
    return time.perf_counter() - T_START


def rtm_bench(a):
    from ouc_amd import abi, bytes_model, velocity_models as vm
    from ouc_amd.engine import Engine, rtm_concurrent

    engine = abi.ENGINE_FUSED if a.engine == "fused" else abi.ENGINE_PLAIN
    strips = abi.STRIP_FP16 if a.strips == "fp16" else abi.STRIP_FP32
    p, vp, vs, shots = vm.cfg_rtm_2048(nt=a.nt, engine=engine, strip_precision=strips)
    batch = a.batch or (8 if engine == abi.ENGINE_FUSED else 1)
    nctx = max(1, a.contexts) if engine == abi.ENGINE_FUSED else 1
    nsteps_total = a.warmup + a.steps + 1  # + the solo leg's batch
    if nsteps_total * batch > len(shots):
        shots = shots * ((nsteps_total * batch) // len(shots) + 1)
    log = lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True)
    log(f"{a.engine} engine, batch {batch}, {nctx} context(s), {a.steps} timed + {a.warmup} warmup steps, "
        f"grid {p.nx}x{p.nz} nt={p.nt}, budget {a.budget_s:.0f} s ({_elapsed():.1f} s used)")
    es = [Engine(p, vp, vs) for _ in range(nctx)]
    e = es[0]
    tile = None
    if engine == abi.ENGINE_FUSED:
        from ouc_amd import tuning

        for x in es:
            x.set_batch_capacity(batch)
            if a.tile:
                cols, pf = (int(v) for v in a.tile.lower().split("x"))
                x.set_tile(cols, pf)
                src = "given"
            else:
                cols, pf, cached = tuning.tuned(x, batch)
                src = "cache" if cached else "autotuned"
        tile = {"core_cols": cols, "prefetch": pf, "source": src}
        log(f"tile {cols} columns, prefetch {pf} ({src})")
    # The dominant kernel is timed with HIP events on context 0's stream: in the
    # timed region with one context; with several, whose kernels overlap there
    # (a launch's duration then includes the other contexts' work), in a solo
    # leg of one batch on context 0 after the timed region.
    e.set_profiling(nctx == 1)

    def run(sl):
        """Steps over shots[sl]: one batch each; with several contexts the
        batches run at once, batch j on context j mod nctx.  Returns the
        timings of all of them."""
        if engine != abi.ENGINE_FUSED:
            agg = {}
            for k0 in range(sl.start, sl.stop, batch):
                e.rtm_batch(shots[k0:k0 + batch])
                for key, v in e.timings().items():
                    agg[key] = agg.get(key, 0) + v
            return agg
        rtm_concurrent(es, shots[sl])
        return e.timings()

    # The wall-clock guard keeps the whole command inside its budget: the
    # timed region runs only as many steps as fit after the warmup's measured
    # rate, with the solo, plain and CPU legs still to come and a margin.
    reserve = (0.0 if a.no_plain or engine != abi.ENGINE_FUSED else 25.0) + \
              (0.0 if a.no_cpu else 2 * a.cpu_sample_s + 15.0) + 20.0
    k = 0
    warm_done, per_step = 0, None
    if a.warmup:
        # first round: every context captures its loops
        w1 = min(a.warmup, nctx)
        t1 = time.perf_counter()
        run(slice(k, k + w1 * batch))
        k += w1 * batch
        warm_done = w1
        per_step = (time.perf_counter() - t1) / w1
        log(f"warmup: {w1} step(s) in {per_step * w1:.1f} s ({_elapsed():.0f} s used)")
        rest = a.warmup - w1
        if rest and _elapsed() + per_step * (rest + a.steps) + reserve <= a.budget_s:
            t1 = time.perf_counter()
            run(slice(k, k + rest * batch))
            k += rest * batch
            warm_done += rest
            per_step = (time.perf_counter() - t1) / rest
            log(f"warmup: {rest} more step(s) in {per_step * rest:.1f} s ({_elapsed():.0f} s used)")
        elif rest:
            log(f"budget: warmup stopped after {w1} steps")
    steps_run = a.steps
    solo_reserve = 0.0 if nctx == 1 else (per_step or 0.0) * nctx * 1.2
    if per_step:
        fit = int((a.budget_s - _elapsed() - reserve - solo_reserve) / per_step)
        if fit < steps_run:
            steps_run = max(1, fit)
            log(f"budget: timed region runs {steps_run} of {a.steps} steps")
    t0 = time.perf_counter()
    tim = run(slice(k, k + steps_run * batch))
    wall = time.perf_counter() - t0
    k += steps_run * batch
    log(f"timed: {steps_run} steps in {wall:.1f} s ({_elapsed():.0f} s used)")
    nshots = steps_run * batch
    shots_per_hour = nshots / wall * 3600.0
    cell_updates, canonical, own = tim["cell_updates"], tim["canonical_bytes"], tim["own_bytes"]

    # dominant kernel roofline: algorithmic bytes per launch / mean launch time
    # (HIP events on context 0's stream)
    per_launch_bytes, kernel_name = bytes_model.dominant_kernel_bytes(p, engine, batch)

    def kernel_roofline(t):
        kms, klaunch = t["step_kernel_ms"], t["step_kernel_launches"]
        mean_ms = kms / max(klaunch, 1)
        achieved = per_launch_bytes / (mean_ms * 1e-3) / 1e9 if klaunch else None
        return {"bound": "hbm", "kernel": kernel_name, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": None,
                "bytes_per_launch": per_launch_bytes, "mean_launch_ms": mean_ms, "launches_timed": klaunch}

    roofline = kernel_roofline(tim) if nctx == 1 else None
    traffic = bytes_model.measured_traffic(kernel_name, batch if engine == abi.ENGINE_FUSED else 1)
    loops = solo = None
    if engine == abi.ENGINE_FUSED:
        fwd_b, bwd_b = bytes_model.fused_loop_bytes(p, batch, strips == abi.STRIP_FP16)
        loops = {"path": {"bytes": (fwd_b + bwd_b) * steps_run, "ms": wall * 1e3,
                          "achieved": (fwd_b + bwd_b) * steps_run / wall / 1e9,
                          "frac": (fwd_b + bwd_b) * steps_run / wall / 1e9 / HBM_PEAK_GBS},
                 "unit": "GB/s",
                 "note": "algorithmic bytes of every kernel of every timed batch over the timed region's wall time"}
        if nctx > 1:
            # one batch on context 0 alone: the kernel and the loops without the other contexts
            e.set_profiling(True)
            t1 = time.perf_counter()
            e.rtm_batch(shots[k:k + batch])
            tim_solo = e.timings()
            roofline = kernel_roofline(tim_solo)
            roofline["measured_on"] = (f"context 0 alone, one batch of {batch} shots after the timed region "
                                       f"(in the timed region {nctx} contexts overlap on the GPU)")
            solo = {"batch_wall_s": time.perf_counter() - t1,
                    "shots_per_hour": batch / (tim_solo["total_ms"] / 1e3) * 3600.0}
            log(f"solo leg: one batch in {solo['batch_wall_s']:.1f} s ({_elapsed():.0f} s used)")
        else:
            tim_solo = tim
        per = 1 if solo else steps_run
        fms = tim_solo["forward_ms"] / per
        bms = tim_solo["backward_ms"] / per
        loops["forward"] = {"bytes": fwd_b, "ms": fms, "achieved": fwd_b / (fms * 1e-3) / 1e9,
                            "frac": fwd_b / (fms * 1e-3) / 1e9 / HBM_PEAK_GBS}
        loops["backward"] = {"bytes": bwd_b, "ms": bms, "achieved": bwd_b / (bms * 1e-3) / 1e9,
                             "frac": bwd_b / (bms * 1e-3) / 1e9 / HBM_PEAK_GBS}
        loops["loops_note"] = "forward and backward: one batch's loops" + (" on context 0 alone" if solo else "")

    if roofline and traffic is not None:
        roofline["traffic"] = traffic
    plain = None
    for x in es:
        x.close()
    if not a.no_plain and engine == abi.ENGINE_FUSED:
        pp, _, _, _ = vm.cfg_rtm_2048(nt=a.nt, engine=abi.ENGINE_PLAIN, strip_precision=abi.STRIP_FP32)
        ep = Engine(pp, vp, vs)
        ep.rtm_shot(shots[0])  # warm
        ms = 0.0
        for s_ in range(a.plain_sample_shots):
            ep.rtm_shot(shots[s_])
            ms += ep.timings()["total_ms"]
        plain = {"shots_per_hour": a.plain_sample_shots / (ms / 1e3) * 3600.0,
                 "sample": f"{a.plain_sample_shots} shot(s), reference kernel structure, FP32 strips"}
        ep.close()
        log(f"plain leg done ({_elapsed():.0f} s used)")

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(p, vp, vs, [shots[0], shots[32 % len(shots)]], a.cpu_sample_s)
        log(f"CPU leg done ({_elapsed():.0f} s used)")

    conc = f", {nctx} contexts at once" if nctx > 1 else ""
    line = {
        "metric": "RTM shots/hour at 2048x1024 (batched); Gcell-updates/s; HBM GB/s",
        "value": shots_per_hour,
        "unit": "shots/hour",
        "n_gpus": 1,
        "steps": steps_run,
        "warmup": warm_done,
        "ms_per_step": wall * 1e3 / steps_run,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded layered + 3% random model)",
        "config": {"workload": "batched elastic RTM, 2048x1024 interior, 5 m, N=5, 30-cell CPML, "
                               f"{p.nt} steps, {batch} shots per step{conc}, {a.strips} strips, {a.engine} engine",
                   "nx": p.nx, "nz": p.nz, "nt": p.nt, "batch": batch, "contexts": nctx, "tile": tile},
        "steps_requested": a.steps,
        "warmup_requested": a.warmup,
        "gcell_updates_per_s": cell_updates / wall / 1e9,
        "canonical_gbs": canonical / wall / 1e9,
        "own_bytes_gbs": own / wall / 1e9 if own else None,
        "wall_s": wall,
        "device_ms_per_step": tim["total_ms"] / steps_run,
        "plain": plain,
        "batched_over_plain": (shots_per_hour / plain["shots_per_hour"]) if plain else None,
        "roofline": roofline,
        "solo_batch": solo,
        "roofline_loops": loops,
        "cpu_baseline": cpu,
        "command_s": _elapsed(),
    }
    print(json.dumps(line), flush=True)


def fwi_bench(a):
    """FWI gradient (residual-driven backward loop with the gradient kernel)
    over batches of shots; observed records are zero, so the residual is the
    modelled record (the work does not depend on the values)."""
    import numpy as np

    from ouc_amd import abi, tuning, velocity_models as vm
    from ouc_amd.engine import Engine

    strips = abi.STRIP_FP16 if a.strips == "fp16" else abi.STRIP_FP32
    p, vp, vs, shots = vm.cfg_rtm_2048(nt=a.nt, engine=abi.ENGINE_FUSED, strip_precision=strips)
    batch = a.batch or 32
    for sh in shots:
        sh["mute"] = 0
    log = lambda m: print(f"[fwi] {m}", file=sys.stderr, flush=True)
    log(f"batch {batch}, {a.steps} timed + {a.warmup} warmup steps, grid {p.nx}x{p.nz} nt={p.nt}")
    obs = np.zeros((batch, 2, p.nx, p.nt), np.float32)
    with Engine(p, vp, vs) as e:
        cols, pf, _ = tuning.tuned(e, batch)
        for w in range(a.warmup):
            e.fwi_gradient(shots[:batch], obs)
            log(f"warmup step {w} done")
        dev_ms = 0.0
        for s_ in range(a.steps):
            e.fwi_gradient(shots[:batch], obs)
            dev_ms += e.timings()["total_ms"]
            log(f"step {s_}: {e.timings()['total_ms']:.1f} ms device")
    nshots = a.steps * batch
    line = {
        "metric": "FWI gradient shots/hour at 2048x1024 (one GPU)", "value": nshots / (dev_ms / 1e3) * 3600.0,
        "unit": "shots/hour", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": dev_ms / a.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded layered model, zero observed records)",
        "config": {"workload": f"FWI gradient, 2048x1024 interior, nt {p.nt}, {batch} shots per step, "
                               f"{a.strips} strips, tile {cols}x{pf}"},
    }
    print(json.dumps(line))


def splitpml_bench(a):
    """Forward modelling of one shot per step with the first-order split-PML
    engine (stress and velocity launches per time step) on the same grid."""
    from ouc_amd import abi, velocity_models as vm
    from ouc_amd.engine import Engine

    p, vp, vs, shots = vm.cfg_rtm_2048(nt=a.nt, engine=abi.ENGINE_SPLIT_PML)
    log = lambda m: print(f"[split-pml] {m}", file=sys.stderr, flush=True)
    with Engine(p, vp, vs) as e:
        for w in range(a.warmup):
            e.forward_shot(dict(shots[w % len(shots)], mute=0))
            log(f"warmup step {w} done")
        dev_ms = 0.0
        for s_ in range(a.steps):
            e.forward_shot(dict(shots[(a.warmup + s_) % len(shots)], mute=0))
            dev_ms += e.timings()["total_ms"]
            log(f"step {s_}: {e.timings()['total_ms']:.1f} ms device")
    upd = (p.nx + 2 * p.npml - 2 * p.half_order) * (p.nz + 2 * p.npml - 2 * p.half_order)
    line = {
        "metric": "split-PML forward modelling, Gcell-updates/s at 2048x1024 (one GPU)",
        "value": upd * p.nt * a.steps / (dev_ms / 1e3) / 1e9, "unit": "Gcell-updates/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": dev_ms / a.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded layered model)",
        "config": {"workload": f"split-PML first-order forward, 2048x1024 interior, nt {p.nt}, one shot per step"},
        "shots_per_hour": a.steps / (dev_ms / 1e3) * 3600.0,
    }
    print(json.dumps(line))


def serve_bench(a):
    """Serving-style batch: shots/hour of the whole job and per-shot p50/p95
    latency (a warm pass captures the graphs and fills the tile cache first)."""
    from ouc_amd import serve

    small = a.serve_grids == "small"
    grids, nt_range = (serve.SMALL_GRIDS, serve.SMALL_NT) if small else (serve.GRIDS, (2000, 8000))
    jobs = serve.make_jobs(n=a.serve_shots, grids=grids, nt_range=nt_range)
    log = lambda m: print(f"[serve] {m}", file=sys.stderr, flush=True)
    learned = a.serve_boundary == "learned"
    lat, wall, groups = serve.serve(jobs, learned=learned, schedule=a.serve_schedule, log=log)
    s = serve.summary(jobs, lat, wall, groups)
    line = {
        "metric": "uneven serving batch: RTM shots/hour and per-shot latency p50/p95",
        "value": s["shots_per_hour"], "unit": "shots/hour", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": wall * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded layered models per grid, random shots)",
        "config": {"workload": f"{len(jobs)} RTM shots, grids {'/'.join(f'{g[0]}x{g[1]}' for g in grids)}, "
                               f"nt {nt_range[0]}-{nt_range[1]}, "
                               + ("learned 8-cell boundary with CPML fallback" if learned else "30-cell CPML")
                               + ", FP16 strips, autotuned tiles, "
                               + ("all grids in shared launches per time step" if a.serve_schedule == "grouped"
                                  else f"per-grid batches, schedule {a.serve_schedule}")},
        "latency_ms_p50": s["latency_ms_p50"], "latency_ms_p95": s["latency_ms_p95"], "groups": s["groups"],
    }
    print(json.dumps(line))


def cpu_threads():
    """Threads for the CPU baseline: the cores this process may run on, capped
    by OMP_NUM_THREADS where the environment sets the process's CPU share."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_baseline(p, vp, vs, sample_shots, seconds):
    """FP32 oracle (OpenMP over x columns, OMP_PROC_BIND=close) on bounded
    samples: each listed shot at the full grid for a number of time steps sized
    to about `seconds` of CPU time, extrapolated linearly in steps to the full
    step count; shots/hour from the mean extrapolated shot time."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    from oracle import pyoracle
    from ouc_amd import abi

    threads = cpu_threads()
    pyoracle.set_threads(threads)
    # probe: a short run sizes the samples
    probe_steps = 40
    ps = abi.make_params(p.nx, p.nz, p.npml, p.half_order, probe_steps, p.dx, p.dz, p.dt, R=p.R)
    t0 = time.perf_counter()
    pyoracle.rtm_shot(ps, vp, vs, sample_shots[0], precision=32)
    per_step = (time.perf_counter() - t0) / probe_steps
    nsteps = int(min(p.nt, max(50, seconds / max(per_step, 1e-6))))
    ps = abi.make_params(p.nx, p.nz, p.npml, p.half_order, nsteps, p.dx, p.dz, p.dt, R=p.R)
    shot_s = []
    for sh in sample_shots:
        t0 = time.perf_counter()
        pyoracle.rtm_shot(ps, vp, vs, sh, precision=32)
        shot_s.append((time.perf_counter() - t0) * p.nt / nsteps)
    per_shot_s = sum(shot_s) / len(shot_s)
    return {"value": 3600.0 / per_shot_s, "unit": "shots/hour", "cores": threads, "kind": "port",
            "host_cpus": os.cpu_count(), "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
            "extrapolated": True,
            "sample": f"shots at x={','.join(str(sh['sx']) for sh in sample_shots)} of {p.nx}x{p.nz}, "
                      f"{nsteps} of {p.nt} steps each (forward with strips, backward with imaging), FP32 oracle "
                      f"on {threads} threads; per-shot times {', '.join(f'{v:.0f}' for v in shot_s)} s "
                      f"extrapolated linearly in steps to {p.nt}"}


if __name__ == "__main__":
    main()
