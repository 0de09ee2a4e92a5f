"""The learned boundary's CPML fallback on the batched paths.

A network scaled by 300 makes every shot's pad diverge; each shot then runs
again on the context's CPML fallback context, so every stack must equal, bit
for bit, the same shots run on a CPML context with the fallback's parameters.
Covered: one batched context (ouc_rtm_batch), two contexts on concurrent host
threads (the serving driver's concurrent schedule), and shared launches over
two grids (ouc_rtm_grouped)."""
import ctypes
import threading

import numpy as np
import pytest

from ouc_amd import abi, serve
from ouc_amd.engine import Engine, rtm_grouped

pytestmark = pytest.mark.gpu

GRIDS = [((96, 48), 300), ((128, 64), 360)]


@pytest.fixture(scope="module")
def bad_weights():
    from ouc_amd.learned import load_weights

    return 300.0 * load_weights()


def _params(grid, nt, boundary):
    p = serve.group_params(grid, nt, True, abi.STRIP_FP16)
    q = abi.Params()
    ctypes.pointer(q)[0] = p  # a copy: the fallback context keeps every field but the boundary
    q.boundary = boundary
    return q


def _shots(grid, nt):
    nx = grid[0]
    sh = [dict(sx=x, sz=0, rec_z=0, fm=20.0, mute=1) for x in (nx // 4, nx // 2, 3 * nx // 4)]
    sh[1]["nt"] = nt - 50
    return sh


def _cpml_stack(grid, nt, gi):
    vp, vs = serve.model_for(grid, 11 + gi)
    with Engine(_params(grid, nt, abi.BOUNDARY_CPML), vp, vs) as ref:
        return ref.rtm_batch(_shots(grid, nt))


def _learned(grid, nt, gi, weights):
    vp, vs = serve.model_for(grid, 11 + gi)
    e = Engine(_params(grid, nt, abi.BOUNDARY_LEARNED_FALLBACK), vp, vs)
    e.set_learned_boundary(weights)
    return e


def _same(a, b, tag):
    for key in ("stack_pp", "stack_ps"):
        assert a[key].tobytes() == b[key].tobytes(), f"{tag}: {key} differs from the CPML context's"


def test_batched_rtm_falls_back_to_cpml(bad_weights):
    (grid, nt) = GRIDS[0]
    print(f"[case] batched fallback {grid} nt={nt}", flush=True)
    ref = _cpml_stack(grid, nt, 0)
    with _learned(grid, nt, 0, bad_weights) as e:
        with pytest.warns(RuntimeWarning):
            out = e.rtm_batch(_shots(grid, nt))
        assert e.fell_back and e.timings()["fallback_shots"] == 3
    _same(out, ref, "batched")


def test_concurrent_contexts_fall_back_to_cpml(bad_weights):
    print(f"[case] concurrent fallback {GRIDS}", flush=True)
    refs = [_cpml_stack(g, nt, gi) for gi, (g, nt) in enumerate(GRIDS)]
    engines = [_learned(g, nt, gi, bad_weights) for gi, (g, nt) in enumerate(GRIDS)]
    outs, errors = [None] * len(engines), []

    def run(k):
        try:
            g, nt = GRIDS[k]
            with pytest.warns(RuntimeWarning):
                outs[k] = engines[k].rtm_batch(_shots(g, nt))
        except Exception as ex:  # re-raised below
            errors.append(ex)

    try:
        th = [threading.Thread(target=run, args=(k,)) for k in range(len(engines))]
        for t in th:
            t.start()
        for t in th:
            t.join()
        if errors:
            raise errors[0]
        for k in range(len(engines)):
            _same(outs[k], refs[k], f"concurrent {GRIDS[k]}")
    finally:
        for e in engines:
            e.close()


def test_grouped_run_falls_back_to_cpml(bad_weights):
    print(f"[case] grouped fallback {GRIDS}", flush=True)
    refs = [_cpml_stack(g, nt, gi) for gi, (g, nt) in enumerate(GRIDS)]
    engines = [_learned(g, nt, gi, bad_weights) for gi, (g, nt) in enumerate(GRIDS)]
    try:
        with pytest.warns(RuntimeWarning):
            outs = rtm_grouped(engines, [_shots(g, nt) for g, nt in GRIDS])
        for k in range(len(engines)):
            _same(outs[k], refs[k], f"grouped {GRIDS[k]}")
    finally:
        for e in engines:
            e.close()

