"""GPU FWI gradient (fused engine) through the C ABI against the FP64 oracle.

Tolerances: relative L2 against the FP64 oracle, the image bar 1e-4 below row
8 (the gradient, like the images, is a correlation of the source wavefield
with the receiver wavefield, and the rows next to the receiver row carry the
reference's FP32 conditioning); the whole-interior error is printed beside it.
The misfit and the illumination are sums of squares of one wavefield each and
meet the field bar 1e-5.
"""
import numpy as np
import pytest

from conftest import rel_l2
from ouc_amd import abi
from ouc_amd.engine import Engine, model_update
from test_fwi_oracle import _anomaly_case, model_update_ref

pytestmark = pytest.mark.gpu

KEYS = ("grad_vp", "grad_vs", "illum")


@pytest.fixture(scope="module")
def observed(oracle):
    out = {}
    for sem in (abi.SEM_FAITHFUL, abi.SEM_EXACT):
        p, vp0, vs0, vpt, shot = _anomaly_case(sem, nt=500)
        out[sem] = oracle.rtm_shot(p, vpt, vs0, shot)["rec"]
    return out


@pytest.mark.parametrize("strips", [abi.STRIP_FP32, abi.STRIP_FP16])
@pytest.mark.parametrize("semantics", [abi.SEM_FAITHFUL, abi.SEM_EXACT])
def test_fwi_gradient_vs_oracle(oracle, observed, semantics, strips):
    p, vp0, vs0, vpt, shot = _anomaly_case(semantics, nt=500)
    obs = observed[semantics]
    ref = oracle.fwi_gradient(p, vp0, vs0, shot, obs)
    p.engine = abi.ENGINE_FUSED
    p.strip_precision = strips
    print(f"[case] fwi 80x60 nt={p.nt} sem={semantics} strips={strips}", flush=True)
    with Engine(p, vp0, vs0) as e:
        got = e.fwi_gradient([shot], obs[None])
    errs = {}
    for k in KEYS:
        errs[k] = (rel_l2(got[k], ref[k]), rel_l2(got[k][:, 8:], ref[k][:, 8:]))
    mis = abs(got["misfit"] - ref["misfit"]) / ref["misfit"]
    print("[fwi] rel-L2 vs FP64 (whole, rows >= 8):",
          " ".join(f"{k}={a:.2e}/{b:.2e}" for k, (a, b) in errs.items()), f"misfit={mis:.2e}", flush=True)
    assert mis < 1e-5
    assert errs["illum"][0] < 1e-5
    for k in ("grad_vp", "grad_vs"):
        assert errs[k][0] < 1e-4 and errs[k][1] < 1e-4, (k, errs[k])


def test_fwi_batch_is_sum_of_shots_and_replays(oracle, observed):
    """Two shots in one call equal the sum of the two single-shot calls, also
    when the batch capacity splits them into two batches; a second call on the
    same context (the captured backward loop replayed) returns the same bits."""
    p, vp0, vs0, vpt, shot = _anomaly_case(abi.SEM_FAITHFUL, nt=500)
    shot2 = dict(shot, sx=25)
    obs2 = oracle.rtm_shot(p, vpt, vs0, shot2)["rec"]
    obs = np.stack([observed[abi.SEM_FAITHFUL], obs2])
    p.engine = abi.ENGINE_FUSED
    with Engine(p, vp0, vs0) as e:
        both = e.fwi_gradient([shot, shot2], obs)
        again = e.fwi_gradient([shot, shot2], obs)
        one = e.fwi_gradient([shot], obs[:1])
        two = e.fwi_gradient([shot2], obs[1:])
        e.set_batch_capacity(1)  # the two shots now run as two batches of one
        split = e.fwi_gradient([shot, shot2], obs)
    for k in KEYS:
        np.testing.assert_array_equal(both[k], again[k])
        s = one[k].astype(np.float64) + two[k]
        assert rel_l2(both[k], s) < 1e-6, k
        assert rel_l2(split[k], s) < 1e-6, k
    assert both["misfit"] == pytest.approx(one["misfit"] + two["misfit"], rel=1e-12)
    assert split["misfit"] == pytest.approx(both["misfit"], rel=1e-12)


def test_model_update_bit_exact():
    rng = np.random.default_rng(11)
    v = rng.uniform(1500, 4500, (64, 48)).astype(np.float32)
    g = (rng.standard_normal((64, 48)) * 3).astype(np.float32)
    for it in (0, 1, 7):
        got = model_update(v, g, 30.0, it)
        np.testing.assert_array_equal(got, model_update_ref(v, g, 30.0, it))


def test_fwi_refuses_what_it_does_not_run(observed):
    """The gradient runs on the fused engine with CPML and full-length shots;
    other requests return OUC_UNSUPPORTED instead of a wrong answer."""
    from ouc_amd.engine import OucError

    p, vp0, vs0, vpt, shot = _anomaly_case(abi.SEM_FAITHFUL, nt=500)
    obs = observed[abi.SEM_FAITHFUL][None]
    p.engine = abi.ENGINE_PLAIN
    with Engine(p, vp0, vs0) as e:
        with pytest.raises(OucError) as err:
            e.fwi_gradient([shot], obs)
        assert err.value.status == abi.OUC_UNSUPPORTED
    p.engine = abi.ENGINE_FUSED
    with Engine(p, vp0, vs0) as e:
        with pytest.raises(OucError) as err:
            e.fwi_gradient([dict(shot, nt=400)], obs)
        assert err.value.status == abi.OUC_UNSUPPORTED
