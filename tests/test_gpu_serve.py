"""Uneven shots: per-shot step counts inside a fused batch, and the serving
driver (uneven grids, learned boundary, concurrent groups)."""
import numpy as np
import pytest

from ouc_amd import abi

pytestmark = pytest.mark.gpu


def test_shorter_shots_in_a_batch_equal_their_own_runs():
    """A shot of nt_s < nt starts late in the batch and images only its own
    steps: bit for bit the image of a context with nt_s steps."""
    from ouc_amd import velocity_models as vm
    from ouc_amd.engine import Engine

    vp, vs, _ = vm.layered_random(160, 80, seed=9)
    nts = [300, 240, 170]
    shots = [dict(sx=s, sz=0, rec_z=0, fm=22.0, mute=1, nt=n) for s, n in zip((30, 80, 130), nts)]
    p = abi.make_params(160, 80, 20, 5, max(nts), 5.0, 5.0, 4e-4, engine=abi.ENGINE_FUSED)
    with Engine(p, vp, vs) as e:
        stack = e.rtm_batch(shots)["stack_pp"]
    acc = np.zeros_like(stack)
    for s, n in zip(shots, nts):
        q = abi.make_params(160, 80, 20, 5, n, 5.0, 5.0, 4e-4, engine=abi.ENGINE_FUSED)
        with Engine(q, vp, vs) as e:
            one = dict(s)
            one.pop("nt")
            acc += e.rtm_shot(one)["lap_pp"]
    np.testing.assert_array_equal(stack, acc)


def test_serving_groups_equal_standalone_runs(tmp_path, monkeypatch):
    from ouc_amd import serve
    from ouc_amd.engine import Engine
    from ouc_amd.learned import load_weights

    monkeypatch.setenv("OUC_TILE_CACHE", str(tmp_path / "tiles.json"))
    jobs = serve.make_jobs(n=6, seed=3, grids=((96, 48), (128, 64)), nt_range=(200, 320))
    lat, wall, groups = serve.serve(jobs, learned=True, strips=abi.STRIP_FP32, warm=True, schedule="concurrent")
    lat2, wall2, groups2 = serve.serve(jobs, learned=True, strips=abi.STRIP_FP32, warm=False, schedule="sjf")
    for a, b in zip(groups, groups2):
        np.testing.assert_array_equal(a.stack_pp, b.stack_pp)
    s = serve.summary(jobs, lat, wall, groups)
    print(f"[serve] {s['shots']} shots in {s['wall_s']:.2f} s, p50 {s['latency_ms_p50']:.1f} ms, "
          f"p95 {s['latency_ms_p95']:.1f} ms, groups {s['groups']}", flush=True)
    assert np.all(lat > 0)
    assert all(g.fallback_shots == 0 for g in groups)
    w = load_weights()
    for gi, g in enumerate(groups):
        js = sorted([j for j in jobs if j.grid == g.grid], key=lambda j: -j.nt)
        vp, vs = serve.model_for(g.grid, 11 + gi)
        acc = np.zeros_like(g.stack_pp)
        for j in js:
            p = serve.group_params(g.grid, j.nt, True, abi.STRIP_FP32)
            with Engine(p, vp, vs) as e:
                e.set_learned_boundary(w)
                one = dict(j.shot)
                one.pop("nt")
                acc += e.rtm_shot(one)["lap_pp"]
        np.testing.assert_array_equal(g.stack_pp, acc)
