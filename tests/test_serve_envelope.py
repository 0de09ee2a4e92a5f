"""The serving driver refuses jobs outside its validated envelope before any
work (CPU: no engine is created)."""
import pytest

from ouc_amd import serve


def test_full_size_jobs_are_served_up_to_the_validated_step_count():
    jobs = serve.make_jobs(n=16, grids=serve.GRIDS, nt_range=(2000, serve.FULL_MAX_NT))
    serve.check_validated(jobs)
    long_job = serve.Job((512, 256), serve.FULL_MAX_NT + 1, dict(sx=1, sz=0, rec_z=0, fm=20.0), 0)
    with pytest.raises(ValueError, match="outside the validated envelope"):
        serve.check_validated(jobs + [long_job])


def test_small_grids_are_not_limited_by_the_full_size_step_count():
    jobs = serve.make_jobs(n=16, grids=serve.SMALL_GRIDS, nt_range=(200, 8000))
    serve.check_validated(jobs)
