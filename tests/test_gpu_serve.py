"""Serving on the GPU: each job's result (final fields and the probe frames
streamed back) is the same run alone and served in a mix with continuous
batching, jobs joining and leaving between chunks.  The tolerance is zero:
a batch member is bit-identical to the job alone."""
import numpy as np
import pytest

from phonomena_amd import fdtd, serve, sweep

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp32c", "fp32"])
def test_served_results_equal_alone(precision):
    jobs = serve.stream_jobs(7, seed=5, nxy=(24, 56), nz=(20, 44), steps=(30, 90))
    arrivals = np.array([0.0, 0.0, 0.01, 0.02, 0.03, 0.5, 0.55])
    srv = serve.Server(precision=precision, chunk=9, record_every=5, probe={"count": 6}, keep_fields=True)
    results, stats = srv.run(jobs, arrivals)
    assert stats["cell_updates"] == sum(r.nodes * j["steps"] for r, j in zip(results, jobs))
    assert stats["p99_latency_s"] >= stats["p50_latency_s"] > 0
    for r, job in zip(results, jobs):
        # alone, with a probe ring that holds every frame of the run
        spec = sweep.job_spec(job, probe=srv.probe, every=5, capacity=100)
        ctx = fdtd.Context(spec, precision=precision)
        ctx.step(job["steps"])
        ctx.sync()
        for f in ("ux", "uy", "uz"):
            assert np.array_equal(ctx.get(f), r.fields[f]), (r.index, f)
        alone = ctx.probes()
        assert len(r.frames) == job["steps"] // 5 == len(alone)
        assert np.array_equal(alone, r.probes), r.index
        ctx.close()


def test_budget_limits_resident_jobs():
    jobs = serve.stream_jobs(5, seed=7, nxy=(40, 48), nz=(30, 34), steps=(40, 40))
    srv = serve.Server(precision="fp32", chunk=10, probe={"count": 4})
    one = serve.resident_bytes(srv.spec_fn(jobs[0]), "fp32")
    srv.hbm_budget = 2.5 * one          # two jobs at a time
    results, stats = srv.run(jobs, np.zeros(5))
    assert stats["peak_resident_bytes"] <= srv.hbm_budget
    assert all(r.finished is not None for r in results)
    assert sorted(r.admitted for r in results)[2] >= min(r.finished for r in results)
