"""One job's shots on several contexts at once (ouc_rtm_concurrent).

Each batch runs on its own context and stream; the shots' Laplacian images are
summed in shot order on the host, so the stacks must equal, bit for bit, one
context's ouc_rtm_batch of the same shots (the same FP32 additions in the same
order), whatever the batch size and however the batches are spread."""
import ctypes

import numpy as np
import pytest

from ouc_amd import abi, serve
from ouc_amd.engine import Engine, OucError, rtm_concurrent

pytestmark = pytest.mark.gpu

GRID, NT = (128, 64), 360


def _shots(n):
    nx = GRID[0]
    sh = [dict(sx=8 + (k * 37) % (nx - 16), sz=0, rec_z=0, fm=20.0 + k, mute=1) for k in range(n)]
    sh[1]["nt"] = NT - 60
    sh[4]["nt"] = NT - 25
    return sh


def _engines(n, cap, learned=False, nt=NT):
    vp, vs = serve.model_for(GRID, 11)
    es = []
    for _ in range(n):
        e = Engine(serve.group_params(GRID, nt, learned, abi.STRIP_FP16), vp, vs)
        if cap:
            e.set_batch_capacity(cap)
        es.append(e)
    return es


def _same(a, b, tag):
    for key in ("stack_pp", "stack_ps"):
        assert a[key].tobytes() == b[key].tobytes(), f"{tag}: {key} differs"


def test_concurrent_contexts_equal_one_context():
    """Seven shots of uneven length in batches of two over three contexts
    (context 0 runs batches 0 and 3), called twice (the second call replays
    the captured loops), against one context's single batch of all seven."""
    print(f"[case] concurrent: 3 contexts x batches of 2, 7 shots, {GRID}, nt {NT}", flush=True)
    shots = _shots(7)
    vp, vs = serve.model_for(GRID, 11)
    with Engine(serve.group_params(GRID, NT, False, abi.STRIP_FP16), vp, vs) as one:
        ref = one.rtm_batch(shots)
    es = _engines(3, 2)
    try:
        first = rtm_concurrent(es, shots, per_shot=True)
        _same(first, ref, "first call")
        assert np.all(first["per_shot_ms"] > 0)
        _same(rtm_concurrent(es, shots), ref, "replayed call")
    finally:
        for e in es:
            e.close()


def test_one_context_and_more_contexts_than_batches():
    """One context (a plain sequence of batches) and four contexts for three
    batches (one context idle) give the same stacks."""
    print("[case] concurrent: 1 context and 4 contexts, 5 shots", flush=True)
    shots = _shots(5)
    es1 = _engines(1, 2)
    es4 = _engines(4, 2)
    try:
        a = rtm_concurrent(es1, shots)
        b = rtm_concurrent(es4, shots)
        _same(a, b, "1 vs 4 contexts")
    finally:
        for e in es1 + es4:
            e.close()


def test_concurrent_fallback_equals_cpml():
    """Learned boundary scaled by 300 on two contexts: every shot falls back
    to CPML, the call warns, and the stacks equal a CPML context's batch."""
    print("[case] concurrent: learned fallback, 2 contexts, 5 shots", flush=True)
    from ouc_amd.learned import load_weights

    shots = _shots(5)
    bad = 300.0 * load_weights()
    vp, vs = serve.model_for(GRID, 11)
    p = serve.group_params(GRID, NT, True, abi.STRIP_FP16)
    q = abi.Params()
    ctypes.pointer(q)[0] = p
    q.boundary = abi.BOUNDARY_CPML
    with Engine(q, vp, vs) as cpml:
        ref = cpml.rtm_batch(shots)
    es = _engines(2, 2, learned=True)
    try:
        for e in es:
            e.set_learned_boundary(bad)
        with pytest.warns(RuntimeWarning, match="ran again with CPML"):
            out = rtm_concurrent(es, shots)
        assert es[0].fell_back
        _same(out, ref, "fallback")
    finally:
        for e in es:
            e.close()


def test_concurrent_rejects_contexts_that_differ():
    print("[case] concurrent: mismatched step counts are refused", flush=True)
    a = _engines(1, 2)[0]
    b = _engines(1, 2, nt=NT + 10)[0]
    try:
        with pytest.raises(OucError) as ei:
            rtm_concurrent([a, b], _shots(5))
        assert ei.value.status == abi.OUC_INVALID_ARG
        with pytest.raises(OucError):
            rtm_concurrent([a, a], _shots(5))
    finally:
        a.close()
        b.close()


def test_concurrent_at_the_bench_grid():
    """The bench's grid and tile (2048 x 1024, 30-cell CPML, FP16 strips, core
    tile 108 x 1) cut to 60 steps: six shots in batches of two over three
    contexts against one context's batch of six, bit for bit."""
    from ouc_amd import velocity_models as vm

    print("[case] concurrent: 2048x1024, nt 60, 3 contexts x batches of 2, 6 shots", flush=True)
    p, vp, vs, shots = vm.cfg_rtm_2048(nt=60, engine=abi.ENGINE_FUSED, strip_precision=abi.STRIP_FP16)
    shots = shots[:6]
    with Engine(p, vp, vs) as one:
        one.set_tile(108, 1)
        ref = one.rtm_batch(shots)
    es = [Engine(p, vp, vs) for _ in range(3)]
    try:
        for e in es:
            e.set_batch_capacity(2)
            e.set_tile(108, 1)
        _same(rtm_concurrent(es, shots), ref, "2048x1024")
    finally:
        for e in es:
            e.close()
