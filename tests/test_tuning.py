"""Tile autotuner: the results cache (CPU) and, on the GPU, that every tile
gives bit-identical results and the tuner picks one of its candidates."""
import numpy as np
import pytest

from ouc_amd import abi


def test_cache_roundtrip(tmp_path):
    from ouc_amd import tuning

    path = str(tmp_path / "tiles.json")
    p = abi.make_params(512, 256, 30, 5, 100, 5.0, 5.0, 4e-4, engine=abi.ENGINE_FUSED)
    key = tuning.shape_key("gfx950", p, 8)
    assert key == "gfx950|572x316|N5|B8|cpml"
    q = abi.make_params(512, 256, 30, 5, 100, 5.0, 5.0, 4e-4, engine=abi.ENGINE_FUSED,
                        boundary=abi.BOUNDARY_LEARNED_FALLBACK, band=8)
    assert tuning.shape_key("gfx950", q, 8) == "gfx950|528x272|N5|B8|learned"
    tuning.store(key, {"core_cols": 128, "prefetch": 3}, path)
    tuning.store("other", {"core_cols": 32, "prefetch": 1}, path)
    c = tuning.load_cache(path)
    assert c[key] == {"core_cols": 128, "prefetch": 3} and "other" in c
    assert tuning.load_cache(str(tmp_path / "missing.json")) == {}


@pytest.mark.gpu
def test_tiles_give_identical_results_and_tuner_picks_a_candidate(tmp_path, monkeypatch):
    from ouc_amd import tuning, velocity_models as vm
    from ouc_amd.engine import Engine

    vp, vs, _ = vm.layered_random(300, 120, seed=5)
    p = abi.make_params(300, 120, 20, 5, 300, 5.0, 5.0, 4e-4, engine=abi.ENGINE_FUSED)
    shots = [dict(sx=s, sz=0, rec_z=0, fm=22.0, mute=1) for s in (40, 150, 260)]
    outs = {}
    with Engine(p, vp, vs) as e:
        assert e.tile() == (64, 2)
        for tile in ((64, 2), (8, 1), (16, 3), (32, 1), (128, 3), (256, 2)):
            e.set_tile(*tile)
            outs[tile] = e.rtm_batch(shots)["stack_pp"]
        monkeypatch.setenv("OUC_TILE_CACHE", str(tmp_path / "tiles.json"))
        cols, pf, cached = tuning.tuned(e, len(shots), reps=3)
        assert not cached and (cols, pf) == e.tile()
        assert 8 <= cols <= 1024 and pf in (1, 2, 3)
        cols2, pf2, cached2 = tuning.tuned(e, len(shots))
        assert cached2 and (cols2, pf2) == (cols, pf)
        table = tuning.load_cache()[tuning.shape_key(tuning.device_arch(), p, len(shots))]["ms_per_launch"]
        print("[autotune] ms per launch: " + ", ".join(f"{k} {v:.3f}" for k, v in table.items()), flush=True)
        assert len(table) == 18 and all(v > 0 for v in table.values())
        outs["tuned"] = e.rtm_batch(shots)["stack_pp"]
    ref = outs[(64, 2)]
    for k, v in outs.items():
        np.testing.assert_array_equal(v, ref, err_msg=str(k))
