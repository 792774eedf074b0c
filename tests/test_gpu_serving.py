"""Continuous batching (sph07_amd.serving): requests arriving while others run
are admitted into the live batch; every result equals the same scene run
alone (bit-identical: scenes share launches, never data)."""
import numpy as np
import pytest

from sph07_amd.native import ARM_FP32C, Scene, F_V
from sph07_amd.serving import serve_batched, serve_one_at_a_time, serving_stream

pytestmark = pytest.mark.gpu


def small_stream():
    return serving_stream(seed=11, n=10, rate=400.0, size_range=(800, 5000),
                          steps_range=(20, 60))


def test_served_results_equal_alone():
    reqs = small_stream()
    st = serve_batched(reqs, ARM_FP32C, max_live=6, chunk=8, keep_results=True)
    assert st.requests == len(reqs)
    assert all(r.t_done >= r.arrival for r in reqs)
    # some requests were admitted while others were still running
    overlap = sum(1 for a in reqs for b in reqs
                  if a is not b and a.t_admit < b.t_done and b.t_admit < a.t_done)
    assert overlap > 0
    for r in reqs:
        alone = Scene(r.deck, ARM_FP32C)
        assert alone.Solution(r.steps) == r.steps
        assert np.array_equal(r.result[F_V], alone.field(F_V)), (r.rid, r.kind)
        alone.close()


def test_one_at_a_time_baseline_runs():
    reqs = small_stream()[:4]
    st = serve_one_at_a_time(reqs, ARM_FP32C)
    s = st.summary()
    assert s["requests"] == 4 and s["particle_updates_per_s"] > 0
