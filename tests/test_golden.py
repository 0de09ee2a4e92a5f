"""Golden fixtures (tests/golden/path_40x40.npz, written by
tests/golden/make_golden.py from the FP64 oracle): the oracle still
reproduces them (CPU), and both GPU engines match them (GPU)."""
import os
import sys

import numpy as np
import pytest

from conftest import rel_l2

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden  # noqa: E402

GOLD = os.path.join(HERE, "golden", "path_40x40.npz")


@pytest.fixture(scope="module")
def gold():
    with np.load(GOLD, allow_pickle=False) as f:
        return {k: np.array(f[k]) for k in f.files}


def test_oracle_reproduces_golden_seismograms(oracle, gold):
    p, vp, vs, shot = make_golden.case()
    seis = oracle.forward_shot(p, vp, vs, shot, 64)
    assert rel_l2(seis, gold["seis"]) < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("engine", ["plain", "fused"])
def test_engines_match_golden(gold, engine):
    from ouc_amd import abi
    from ouc_amd.engine import Engine

    eng = abi.ENGINE_FUSED if engine == "fused" else abi.ENGINE_PLAIN
    p, vp, vs, shot = make_golden.case(engine=eng)
    print(f"[case] golden 40x40 N=6 engine={engine}", flush=True)
    with Engine(p, vp, vs) as e:
        seis = e.forward_shot(shot)
    names = ["vx", "vz", "vpx", "vpz", "vsx", "vsz"]
    errs = {n: rel_l2(seis[k], gold["seis"][k]) for k, n in enumerate(names)}
    for tag, sem in (("faithful", abi.SEM_FAITHFUL), ("exact", abi.SEM_EXACT)):
        p, vp, vs, shot = make_golden.case(sem, eng)
        with Engine(p, vp, vs) as e:
            img = e.rtm_shot(shot)
        for k in ("img_pp", "img_ps", "lap_pp", "lap_ps"):
            errs[f"{tag}_{k}"] = rel_l2(img[k], gold[f"{tag}_{k}"])
            errs[f"{tag}_{k}_deep"] = rel_l2(img[k][:, 8:], gold[f"{tag}_{k}"][:, 8:])
    print(f"[golden {engine}] " + " ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)
    # 1e-5 is the target.  Totals meet it in the fused engine (it carries total V);
    # the plain engine's potential form measures up to 1.7e-5 on these muted
    # records (mostly reflected energy, 12-cell CPML).  The fused engine forms
    # omega from total V, whose P part cancels in FP32, so its split parts carry
    # ~1e-4 (measured vpz 1.1e-4, vsz 2.0e-4).
    tol_total = 3e-5 if engine == "plain" else 1e-5
    tol_split = 3e-5 if engine == "plain" else 5e-4
    for n in ("vx", "vz"):
        assert errs[n] < tol_total, (n, errs[n])
    for n in ("vpx", "vpz", "vsx", "vsz"):
        assert errs[n] < tol_split, (n, errs[n])
    for tag in ("faithful", "exact"):
        for k in ("img_pp", "img_ps"):
            assert errs[f"{tag}_{k}_deep"] < 1e-3, (tag, k, errs[f"{tag}_{k}_deep"])
            assert errs[f"{tag}_{k}"] < 1e-2, (tag, k, errs[f"{tag}_{k}"])
