"""Serving: a design-sweep request loop with continuous batching.

Jobs (seeded crystals with their own grid sizes and step counts) arrive
over time.  The server keeps one running batch (``fdtd.Batch``) resident on
the GPU and advances it in chunks of steps.  Between chunks it
  * streams back each member's new probe frames;
  * retires members that reached their step count (their last frames
    delivered, their device memory freed);
  * admits queued jobs as memory frees up: a job joins when its fields fit
    in the HBM budget beside the members already resident, and from the
    next chunk on its work overlaps with theirs in the same launches.
Host-side preparation of arriving jobs (lattice, inclusion list, source
table) runs while the GPU executes the current chunk; the material map is
rasterized on the device when the job's context is created.

Every member's fields and probes are bit-identical to running the job
alone (the batch's guarantee), so a served result does not depend on what
else was in the mix.

Latency of a job = time from its arrival to the delivery of its last probe
frame.  Throughput = cell-updates of all jobs / wall time of the whole run.
"""
import heapq
import time

import numpy as np

from . import fdtd, sweep

# bytes per node of a resident job: two time levels of three components
# (FP32 hi + 16-bit lo when compensated) and the material index
NODE_BYTES = {"fp64": 2 * 3 * 8 + 1, "fp32": 2 * 3 * 4 + 1, "fp32c": 2 * 3 * (4 + 2) + 1}


def resident_bytes(spec, precision):
    """Device memory a job holds while it is a batch member."""
    per = NODE_BYTES[precision] * spec.cells
    probe = 0
    if spec.probe is not None:
        probe = spec.probe[2] * spec.probe[5] * spec.probe_capacity * 8
    return per + probe + 64 * 1024  # + spacing/material/source tables


def poisson_arrivals(n, rate, seed):
    """n arrival times (s) of a Poisson process of the given rate (jobs/s)."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


def stream_jobs(n, seed, nxy=(32, 256), nz=(32, 128), steps=(500, 8000)):
    """The served stream's jobs: grids U[32^3 .. 256x256x128] cells, step
    counts U{500..8000}, lattice parameters as in the batched mix."""
    rng = np.random.default_rng(seed)
    return [sweep.random_job(rng, nxy=nxy, nz=nz, steps=steps) for _ in range(n)]


class Result:
    """What the server hands back for one job."""

    def __init__(self, index, job, nodes, arrival):
        self.index = index
        self.job = job
        self.nodes = nodes
        self.arrival = arrival
        self.admitted = None
        self.finished = None
        self.frames = []        # probe frames, in order, as they streamed back
        self.fields = None      # final fields when the server keeps them

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival

    @property
    def probes(self):
        return np.array(self.frames) if self.frames else np.zeros((0, 0, 0))


class Server:
    """Continuous-batching request loop over one GPU.

    precision: 'fp32c' (default; compensated FP32 fields), 'fp32' or 'fp64'.
    hbm_budget: bytes of device memory the resident jobs may hold.
    chunk: steps per batched chunk (admission / retirement granularity).
    probe, record_every: the probe grid every job records (surface u_z).
    keep_fields: keep each job's final fields in its Result.
    device_ids: jobs' material maps are rasterized on the device from their
    inclusion lists (no host map, no map upload).
    """

    def __init__(self, precision="fp32c", hbm_budget=200e9, chunk=200, probe=None, record_every=10,
                 keep_fields=False, tile=None, spec_fn=None, context_factory=None, batch_factory=None,
                 clock=time.perf_counter, device_ids=True):
        self.precision = precision
        self.hbm_budget = float(hbm_budget)
        self.chunk = int(chunk)
        self.probe = probe if probe is not None else {"count": 32}
        self.record_every = int(record_every)
        self.keep_fields = keep_fields
        self.device_ids = device_ids
        self.tile = tile
        self.spec_fn = spec_fn or self._spec
        # served contexts only step inside the batch: no shape measurement
        self.context_factory = context_factory or (lambda spec: fdtd.Context(spec, precision=self.precision,
                                                                             tune=False))
        self.batch_factory = batch_factory or (lambda: fdtd.Batch(self.precision))
        self.clock = clock

    def _spec(self, job):
        # the ring holds two chunks of frames, so nothing is overwritten
        # before it is pulled
        cap = 2 * (self.chunk // self.record_every + 1)
        return sweep.job_spec(job, probe=self.probe, every=self.record_every, capacity=cap,
                              device_ids=self.device_ids)

    def run(self, jobs, arrivals):
        """Serve ``jobs`` arriving at ``arrivals`` (seconds from the start).
        Returns (results, stats)."""
        order = sorted(range(len(jobs)), key=lambda i: arrivals[i])
        t_start = self.clock()
        results = [None] * len(jobs)
        pending = []          # (arrival, index) arrived, waiting for memory
        prepared = {}         # index -> spec, meshed on the host
        ready = {}            # index -> (ctx, bytes): uploaded, not yet a member
        nxt = 0               # next job (in arrival order) not yet arrived
        live = {}             # index -> [ctx, spec, delivered frames, bytes]
        used = 0.0
        batch = self.batch_factory()
        if self.tile:
            batch.set_tiles(*self.tile)
        cells_done = 0
        peak = 0.0
        chunks = 0
        spent = {"prepare_s": 0.0, "admit_s": 0.0, "issue_s": 0.0, "wait_s": 0.0, "stream_s": 0.0,
                 "retire_s": 0.0, "idle_s": 0.0}

        def prep(i):
            t = self.clock()
            sp = self.spec_fn(jobs[i])
            spent["prepare_s"] += self.clock() - t
            return sp
        try:
            while nxt < len(order) or pending or live:
                now = self.clock() - t_start
                while nxt < len(order) and arrivals[order[nxt]] <= now:
                    i = order[nxt]
                    heapq.heappush(pending, (arrivals[i], i))
                    nxt += 1
                # admit in arrival order while memory allows (contexts
                # uploaded during the last chunk already hold their memory)
                while pending:
                    _, i = pending[0]
                    t = self.clock()
                    if i in ready:
                        ctx, need = ready.pop(i)
                        spec = ctx.spec
                    else:
                        spec = prepared.pop(i, None) or prep(i)
                        need = resident_bytes(spec, self.precision)
                        if live and used + need > self.hbm_budget:
                            prepared[i] = spec
                            break
                        ctx = self.context_factory(spec)
                        used += need
                    heapq.heappop(pending)
                    batch.add(ctx)
                    spent["admit_s"] += self.clock() - t
                    r = Result(i, jobs[i], spec.cells, arrivals[i])
                    r.admitted = self.clock() - t_start
                    results[i] = r
                    live[i] = [ctx, spec, 0, need]
                peak = max(peak, used)
                if not live:
                    # idle until the next arrival
                    if nxt < len(order):
                        wait = arrivals[order[nxt]] - (self.clock() - t_start)
                        if wait > 0:
                            time.sleep(min(wait, 0.05))
                            spent["idle_s"] += min(wait, 0.05)
                    continue
                t = self.clock()
                batch.step(self.chunk)
                spent["issue_s"] += self.clock() - t
                chunks += 1
                # host work overlapping the chunk: mesh the jobs that arrived
                # and upload those that fit in the budget
                now = self.clock() - t_start
                while nxt < len(order) and arrivals[order[nxt]] <= now and len(prepared) + len(ready) < 8:
                    i = order[nxt]
                    heapq.heappush(pending, (arrivals[i], i))
                    nxt += 1
                for _, i in sorted(pending):
                    if i in ready:
                        continue
                    if i not in prepared:
                        prepared[i] = prep(i)
                    need = resident_bytes(prepared[i], self.precision)
                    if used + need > self.hbm_budget:
                        break
                    t = self.clock()
                    ready[i] = (self.context_factory(prepared.pop(i)), need)
                    used += need
                    spent["admit_s"] += self.clock() - t
                t = self.clock()
                batch.sync()
                spent["wait_s"] += self.clock() - t
                # stream probes, retire finished jobs
                for i in list(live):
                    ctx, spec, delivered, need = live[i]
                    done = ctx.steps_done
                    total = done // spec.record_every if spec.probe is not None else 0
                    if total > delivered:
                        t = self.clock()
                        frames = ctx.probes()
                        new = total - delivered
                        results[i].frames.extend(list(frames[len(frames) - new:]))
                        live[i][2] = total
                        spent["stream_s"] += self.clock() - t
                    if done >= spec.wave.size:
                        t = self.clock()
                        if self.keep_fields:
                            results[i].fields = {f: ctx.get(f) for f in ("ux", "uy", "uz")}
                        results[i].finished = self.clock() - t_start
                        batch.remove(ctx)
                        ctx.close()
                        cells_done += spec.cells * spec.wave.size
                        used -= need
                        del live[i]
                        spent["retire_s"] += self.clock() - t
        finally:
            for ctx, _, _, _ in live.values():
                try:
                    batch.remove(ctx)
                    ctx.close()
                except Exception:
                    pass
            for ctx, _ in ready.values():
                try:
                    ctx.close()
                except Exception:
                    pass
            batch.close()
        wall = self.clock() - t_start
        lat = np.array([r.latency for r in results if r is not None and r.latency is not None])
        stats = {"jobs": len(jobs), "wall_s": wall, "cell_updates": cells_done,
                 "value": cells_done / wall if wall > 0 else 0.0, "unit": "cell-updates/s",
                 "p50_latency_s": float(np.percentile(lat, 50)) if lat.size else None,
                 "p99_latency_s": float(np.percentile(lat, 99)) if lat.size else None,
                 "peak_resident_bytes": peak, "hbm_budget_bytes": self.hbm_budget, "chunks": chunks,
                 "chunk_steps": self.chunk, "precision": self.precision,
                 "host_seconds": {k: round(v, 4) for k, v in spent.items()}}
        return results, stats
