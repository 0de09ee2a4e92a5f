"""Time loops longer than one capture chunk (2048 steps) are captured as
several graphs replayed in order; the replay must equal eager launches bit for
bit, on CPML and learned contexts, for the batched RTM (forward loop with
strips, backward loop on two streams) and forward modelling, and a grouped run
over two grids must equal each grid's own batch."""
import numpy as np
import pytest

from ouc_amd import abi, serve, velocity_models
from ouc_amd.engine import Engine, rtm_grouped

pytestmark = pytest.mark.gpu

NT = 2100  # two chunks


def _engine(grid, learned, seed):
    vp, vs = serve.model_for(grid, seed)
    p = serve.group_params(grid, NT, learned, abi.STRIP_FP16)
    e = Engine(p, vp, vs)
    if learned:
        from ouc_amd.learned import load_weights

        e.set_learned_boundary(load_weights())
    return e


@pytest.mark.parametrize("learned", [False, True], ids=["cpml", "learned"])
def test_chunked_replay_equals_eager_launches(learned):
    grid = (96, 48)
    print(f"[case] chunked loops {grid} nt={NT} learned={learned}", flush=True)
    shots = [dict(sx=s, sz=0, rec_z=0, fm=20.0, mute=1) for s in (20, 48, 70)]
    shots[1]["nt"] = NT - 300
    with _engine(grid, learned, 11) as e:
        e.set_graphs(False)
        ref = e.rtm_batch(shots)
        fref = e.forward_shot(dict(sx=30, sz=0, rec_z=0, fm=20.0))
        e.set_graphs(True)
        for k in range(2):  # capture, then replay
            out = e.rtm_batch(shots)
            for key in ("stack_pp", "stack_ps"):
                assert out[key].tobytes() == ref[key].tobytes(), (k, key)
            f = e.forward_shot(dict(sx=30, sz=0, rec_z=0, fm=20.0))
            np.testing.assert_array_equal(f, fref)
        assert not getattr(e, "fell_back", False)


def test_chunked_grouped_run_equals_each_grid_alone():
    grids = [(96, 48), (128, 64)]
    print(f"[case] chunked grouped CPML {grids} nt={NT}", flush=True)
    es = [_engine(g, False, 20 + k) for k, g in enumerate(grids)]
    try:
        lists = [[dict(sx=g[0] // 3, sz=0, rec_z=0, fm=20.0, mute=1),
                  dict(sx=2 * g[0] // 3, sz=0, rec_z=0, fm=20.0, mute=1, nt=NT - 200 * (k + 1))]
                 for k, g in enumerate(grids)]
        grouped = rtm_grouped(es, lists)
        for k, (e, shots) in enumerate(zip(es, lists)):
            own = e.rtm_batch(shots)
            for key in ("stack_pp", "stack_ps"):
                assert grouped[k][key].tobytes() == own[key].tobytes(), (grids[k], key)
    finally:
        for e in es:
            e.close()
