"""Host logic of the serving loop, on stand-in contexts (no GPU): every
job is admitted, run to its step count and retired; admission keeps the
resident set inside the HBM budget; probe frames stream back complete and
in order; latencies and throughput are accounted."""
import numpy as np

from phonomena_amd import serve


class FakeSpec:
    def __init__(self, cells, steps, every):
        self.cells = cells
        self.wave = np.zeros(steps)
        self.probe = (0, 1, 2, 0, 1, 2)
        self.record_every = every
        self.probe_capacity = 1000


class FakeCtx:
    def __init__(self, spec):
        self.spec = spec
        self.steps_done = 0
        self.frames = []
        self.closed = False

    def probes(self):
        return np.array(self.frames[-self.spec.probe_capacity:])

    def get(self, f):
        return np.full((2, 2, 2), float(self.steps_done))

    def close(self):
        self.closed = True


class FakeBatch:
    def __init__(self, log):
        self.members = []
        self.log = log

    def add(self, c):
        self.members.append(c)
        self.log.append(("add", sum(m.spec.cells for m in self.members)))

    def remove(self, c):
        self.members.remove(c)

    def step(self, n):
        for c in self.members:
            for _ in range(min(n, c.spec.wave.size - c.steps_done)):
                c.steps_done += 1
                if c.steps_done % c.spec.record_every == 0:
                    c.frames.append(np.full((2, 2), float(c.steps_done)))

    def sync(self):
        pass

    def set_tiles(self, *a):
        pass

    def close(self):
        pass


def _server(budget, log, chunk=25):
    bytes_per = serve.NODE_BYTES["fp32c"]
    return serve.Server(precision="fp32c", hbm_budget=budget * bytes_per, chunk=chunk, keep_fields=True,
                        spec_fn=lambda job: FakeSpec(job["cells"], job["steps"], 5),
                        context_factory=FakeCtx, batch_factory=lambda: FakeBatch(log))


def test_all_jobs_run_to_completion_with_streamed_frames():
    rng = np.random.default_rng(1)
    jobs = [{"cells": int(rng.integers(100, 1000)), "steps": int(rng.integers(10, 120))} for _ in range(20)]
    arrivals = serve.poisson_arrivals(len(jobs), 2000.0, 3)
    log = []
    results, stats = _server(1e9, log).run(jobs, arrivals)
    assert stats["jobs"] == 20 and stats["cell_updates"] == sum(j["cells"] * j["steps"] for j in jobs)
    for r, j in zip(results, jobs):
        assert r.finished is not None and r.latency >= 0 and r.admitted >= 0
        got = [f[0, 0] for f in r.frames]
        assert got == [float(s) for s in range(5, j["steps"] + 1, 5)]   # complete, in order
        assert r.fields["ux"][0, 0, 0] == j["steps"]
    assert stats["p50_latency_s"] <= stats["p99_latency_s"]


def test_admission_keeps_resident_set_in_budget():
    jobs = [{"cells": 400, "steps": 60} for _ in range(12)]
    arrivals = np.zeros(12)            # all at once
    log = []
    budget_nodes = 1000                # room for two jobs of 400 (plus tables)
    results, stats = _server(budget_nodes + 2 * 70000 / serve.NODE_BYTES["fp32c"], log).run(jobs, arrivals)
    assert all(r.finished is not None for r in results)
    assert max(n for _, n in log) <= 1200
    assert stats["peak_resident_bytes"] <= stats["hbm_budget_bytes"]
    # jobs waited for memory: later jobs were admitted after earlier ones finished
    assert max(r.admitted for r in results) >= min(r.finished for r in results)


def test_oversized_job_still_runs_alone():
    jobs = [{"cells": 10 ** 6, "steps": 20}]
    log = []
    results, stats = _server(10, log).run(jobs, np.zeros(1))
    assert results[0].finished is not None
