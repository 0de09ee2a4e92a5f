"""The streaming-rate probe runs and reports a launch time."""
import pytest

from ouc_amd.engine import stream_probe

pytestmark = pytest.mark.gpu


def test_stream_probe_times_launches():
    ms = stream_probe(10, 8, 1 << 20, reps=3)
    print(f"[probe] 10x8 x 1M floats: {ms:.4f} ms per launch", flush=True)
    assert 0 < ms < 100


def test_stream_probe_with_shared_inputs():
    ms = stream_probe(10, 8, 1 << 20, reps=3, nshared=2, m=1 << 16)
    assert 0 < ms < 100
    for w in (2, 4):
        assert 0 < stream_probe(10, 8, 1 << 20, reps=3, width=w) < 100
