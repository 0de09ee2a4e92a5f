"""Uneven shots: per-shot step counts inside a fused batch, and the serving
driver (uneven grids, learned boundary, concurrent groups)."""
import numpy as np
import pytest

from ouc_amd import abi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("strips", [abi.STRIP_FP32, abi.STRIP_FP16])
def test_shorter_shots_in_a_batch_equal_their_own_runs(strips):
    """A shot of nt_s < nt starts late in the batch and images only its own
    steps: bit for bit the image of a context with nt_s steps.  With FP16
    strips the codec's keyframes count from each shot's own first slot (the
    starts here, 60 and 130, are off the 32-slot keyframe period)."""
    from ouc_amd import velocity_models as vm
    from ouc_amd.engine import Engine

    vp, vs, _ = vm.layered_random(160, 80, seed=9)
    nts = [300, 240, 170]
    shots = [dict(sx=s, sz=0, rec_z=0, fm=22.0, mute=1, nt=n) for s, n in zip((30, 80, 130), nts)]
    p = abi.make_params(160, 80, 20, 5, max(nts), 5.0, 5.0, 4e-4, engine=abi.ENGINE_FUSED, strip_precision=strips)
    with Engine(p, vp, vs) as e:
        stack = e.rtm_batch(shots)["stack_pp"]
    acc = np.zeros_like(stack)
    for s, n in zip(shots, nts):
        q = abi.make_params(160, 80, 20, 5, n, 5.0, 5.0, 4e-4, engine=abi.ENGINE_FUSED, strip_precision=strips)
        with Engine(q, vp, vs) as e:
            one = dict(s)
            one.pop("nt")
            acc += e.rtm_shot(one)["lap_pp"]
    np.testing.assert_array_equal(stack, acc)


def test_a_shot_longer_than_the_context_is_refused():
    from ouc_amd import velocity_models as vm
    from ouc_amd.engine import Engine, OucError

    vp, vs, _ = vm.layered_random(96, 48, seed=9)
    p = abi.make_params(96, 48, 20, 5, 200, 5.0, 5.0, 4e-4, engine=abi.ENGINE_FUSED)
    with Engine(p, vp, vs) as e:
        with pytest.raises(OucError) as err:
            e.rtm_batch([dict(sx=40, sz=0, rec_z=0, fm=22.0, mute=1, nt=201)])
    assert err.value.status == abi.OUC_INVALID_ARG


def test_serving_groups_equal_standalone_runs(tmp_path, monkeypatch):
    from ouc_amd import serve
    from ouc_amd.engine import Engine
    from ouc_amd.learned import load_weights

    monkeypatch.setenv("OUC_TILE_CACHE", str(tmp_path / "tiles.json"))
    jobs = serve.make_jobs(n=6, seed=3, grids=((96, 48), (128, 64)), nt_range=(200, 320))
    lat, wall, groups = serve.serve(jobs, learned=True, strips=abi.STRIP_FP32, warm=True, schedule="concurrent")
    lat2, wall2, groups2 = serve.serve(jobs, learned=True, strips=abi.STRIP_FP32, warm=False, schedule="sjf")
    lat3, wall3, groups3 = serve.serve(jobs, learned=True, strips=abi.STRIP_FP32, warm=True, schedule="grouped")
    for a, b, c in zip(groups, groups2, groups3):
        np.testing.assert_array_equal(a.stack_pp, b.stack_pp)
        np.testing.assert_array_equal(a.stack_pp, c.stack_pp)  # all grids in shared launches
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


def test_serving_at_the_configured_grid_sizes(tmp_path, monkeypatch):
    """The serving configuration at its own sizes: two shots on each grid of
    512 x 256, 1024 x 512, 2048 x 1024 and 4096 x 2048 (5 m), 2000 and 2600
    steps, learned 8-cell boundary with CPML fallback, scaled-FP16 strips,
    autotuned tiles, groups on concurrent streams.  Each group's stack equals,
    bit for bit, the sum of its shots run alone on a context of their own step
    count."""
    from ouc_amd import serve
    from ouc_amd.engine import Engine
    from ouc_amd.learned import load_weights

    monkeypatch.setenv("OUC_TILE_CACHE", str(tmp_path / "tiles.json"))
    jobs = []
    for gi, g in enumerate(serve.GRIDS):
        for k, nt in enumerate((2000, 2600)):
            sx = g[0] // 3 + k * g[0] // 4
            jobs.append(serve.Job(tuple(g), nt, dict(sx=sx, sz=0, rec_z=0, fm=20.0, mute=1, nt=nt), len(jobs)))
    print(f"[case] serving {len(jobs)} shots on grids {serve.GRIDS}, nt 2000/2600, learned, fp16 strips", flush=True)
    lat, wall, groups = serve.serve(jobs, learned=True, strips=abi.STRIP_FP16, warm=False, schedule="concurrent")
    s = serve.summary(jobs, lat, wall, groups)
    print(f"[serve full] {s['shots']} shots in {s['wall_s']:.2f} s, p50 {s['latency_ms_p50']:.0f} ms, "
          f"p95 {s['latency_ms_p95']:.0f} ms", flush=True)
    assert all(g.fallback_shots == 0 for g in groups)
    w = load_weights()
    for gi, g in enumerate(groups):
        js = sorted([j for j in jobs if j.grid == g.grid], key=lambda j: -j.nt)
        vp, vs = serve.model_for(g.grid, 11 + gi)
        acc = np.zeros_like(g.stack_pp)
        for j in js:
            print(f"[case] standalone {g.grid} nt={j.nt}", flush=True)
            p = serve.group_params(g.grid, j.nt, True, abi.STRIP_FP16)
            with Engine(p, vp, vs) as e:
                e.set_learned_boundary(w)
                one = dict(j.shot)
                one.pop("nt")
                acc += e.rtm_shot(one)["lap_pp"]
        np.testing.assert_array_equal(g.stack_pp, acc)
