"""The fused engine (batched single-pass kernels): forward seismograms against
the FP64 oracle, batches against single shots, graph replays against eager
launches.  Its images are held to the FP64 oracle in test_gpu_parity.py."""
import numpy as np
import pytest

from conftest import rel_l2
from image_metrics import FIELD_BAR, bar_param
from ouc_amd import abi, velocity_models as vm
from ouc_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _case(semantics=abi.SEM_FAITHFUL, engine=abi.ENGINE_FUSED, nx=160, nz=96, nt=900, **kw):
    vp, vs, _ = vm.layered_random(nx, nz, seed=3)
    p = abi.make_params(nx, nz, 24, 5, nt, 5.0, 5.0, 0.0004, R=1e-3, semantics=semantics, engine=engine, **kw)
    return p, vp, vs


def test_fused_batch_is_sum_of_single_shots(oracle):
    """Shots batched in one launch give, bit for bit, the images they give
    alone, and the stack is their sum in shot order."""
    p, vp, vs = _case(nt=500)
    shots = [dict(sx=s, sz=0, rec_z=0, fm=22.0, mute=1) for s in (30, 80, 130)]
    with Engine(p, vp, vs) as e:
        singles = [e.rtm_shot(s) for s in shots]
        out = e.rtm_batch(shots)
    spp = np.zeros_like(singles[0]["lap_pp"])
    sps = np.zeros_like(spp)
    for s in singles:
        spp += s["lap_pp"]
        sps += s["lap_ps"]
    print(f"[batch] max |stack - sum| pp {np.nanmax(np.abs(out['stack_pp'] - spp)):.3e}", flush=True)
    assert np.array_equal(out["stack_pp"], spp, equal_nan=True)
    assert np.array_equal(out["stack_ps"], sps, equal_nan=True)


FWD_MISSES = {
    'vpz': 1.243e-05,
    'vsx': 6.990e-05,
    'vsz': 1.250e-04,
}
FWD_COMP = ["vx", "vz", "vpx", "vpz", "vsx", "vsz"]
_fwd = {}


@pytest.mark.parametrize("comp", [bar_param(c, FWD_MISSES, FIELD_BAR, c, id=c) for c in FWD_COMP])
def test_fused_forward_seismograms_vs_oracle(oracle, comp):
    """The fused engine's forward path (ouc_forward_shot with ENGINE_FUSED),
    160 x 96, 700 steps, source and receivers at depth 5: each of the six
    seismogram components against the FP64 oracle at the 1e-5 field bar."""
    p, vp, vs = _case(nt=700)
    shot = dict(sx=80, sz=5, rec_z=5, fm=25.0)
    if not _fwd:
        with Engine(p, vp, vs) as e:
            _fwd["gpu"] = e.forward_shot(shot)
        _fwd["ref"] = oracle.forward_shot(p, vp, vs, shot, 64)
    q = FWD_COMP.index(comp)
    err = rel_l2(_fwd["gpu"][q], _fwd["ref"][q])
    print(f"[fused forward] {comp}: {err:.2e}", flush=True)
    assert err < FIELD_BAR, err


@pytest.mark.parametrize("learned", [False, True])
def test_graph_replays_equal_eager_launches(learned):
    """Time loops captured as HIP graphs and replayed give, bit for bit, what
    eager launches give -- on the first replay and on later ones, with eager
    work of another batch size in between."""
    if learned:
        from ouc_amd import learned as LB

        c = LB.Case(64, 31, nt=300)
        p, vp, vs = c.small_params(abi.BOUNDARY_LEARNED), c.vp, c.vs
        w = LB.load_weights()
    else:
        p, vp, vs = _case(nt=300)
    shots = [dict(sx=s, sz=0, rec_z=0, fm=22.0, mute=1) for s in (10, 30, 50)]
    with Engine(p, vp, vs) as e:
        if learned:
            e.set_learned_boundary(w)
        e.set_graphs(False)
        ref = e.rtm_batch(shots)
        fref = e.forward_shot(shots[0])
        e.set_graphs(True)
        outs = []
        for k in range(3):
            outs.append(e.rtm_batch(shots))
            f = e.forward_shot(shots[0])      # another captured loop (forward only, one shot)
            np.testing.assert_array_equal(f, fref)
            e.set_graphs(False)
            e.rtm_batch(shots[:2])            # eager work between replays
            e.set_graphs(True)
    for o in outs:
        np.testing.assert_array_equal(o["stack_pp"], ref["stack_pp"])
        np.testing.assert_array_equal(o["stack_ps"], ref["stack_ps"])


def test_graph_key_follows_the_receiver_row():
    """Two batches of the same size on one context with different receiver
    rows: the second replays no loop captured for the first.  Its images and
    seismograms equal, bit for bit, the same calls made with eager launches."""
    p, vp, vs = _case(nt=300)
    mk = lambda rz: [dict(sx=s, sz=0, rec_z=rz, fm=22.0, mute=1) for s in (30, 90)]
    with Engine(p, vp, vs) as e:
        e.set_graphs(True)
        e.rtm_batch(mk(0))
        e.forward_shot(mk(0)[0])
        got = e.rtm_batch(mk(5))
        fgot = e.forward_shot(mk(5)[0])
        e.set_graphs(False)
        ref = e.rtm_batch(mk(5))
        fref = e.forward_shot(mk(5)[0])
    print(f"[graph key] rec_z 0 then 5: max |stack - eager| "
          f"{np.nanmax(np.abs(got['stack_pp'] - ref['stack_pp'])):.3e}", flush=True)
    np.testing.assert_array_equal(got["stack_pp"], ref["stack_pp"])
    np.testing.assert_array_equal(got["stack_ps"], ref["stack_ps"])
    np.testing.assert_array_equal(fgot, fref)
