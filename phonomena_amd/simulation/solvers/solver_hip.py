"""GPU solver plugin: phonomena's step loop on an MI355X.

A drop-in ``Solver`` for the reference's plugin discovery
(phonomena/common.py:88-110): no-argument constructor, ``name``,
``description``, ``cfg``; ``init(grid, material, steps)``;
``run(*args, signals=...)`` blocking, emitting ``status``/``progress`` and
leaving the final fields in ``self.g`` (as base_solver.py:224-280 does);
``cancel``; ``test``.  The stencil runs in libphonofdtd through
``phonomena_amd.fdtd``; there is no CPU path.

cfg keys added to the reference's {wave, wave_args, write_mode}:
  precision     'fp64' (reference arithmetic) | 'fp32' | 'fp32c' (compensated)
  kernel        'auto' | 'naive' | 'tiled' | 'strip' | 'tiled2' (two steps per pass)
  chunk         steps issued between cancel checks / progress updates
  probe         None | [i0, si, ni, j0, sj, nj] | {'stride': s} | {'count': n}
  record_every  probe frame every k steps (frames land in ``self.probes``)
  sponge        None | {'width': cells, 'strength': s}  (extension absorber)
  free_surface  'reference' (default, bit for bit) | 'fixed': the surface u_z
                update with its precedence corrected (base_solver.py:516), for
                physical spacings (extension; equals 'reference' at unit spacing)
  record_slots  snapshot slots of the per-step recording (write_mode != 'off')
write_mode adds 'npy' to the reference's 'off' / 'thread' / 'process': the
Writer's datasets and attributes as a .npy directory at ``file + '.d'``
(no h5py needed).  Every recorded frame is a device snapshot copied to the
host while the next steps run (simulation.recording).
"""
import logging
from time import time

import numpy as np

from .. import base_solver, recording
from ...gui.worker import WorkerSignals
from ... import fdtd, problem

logger = logging.getLogger(__name__)

cfg = {
    "precision": "fp64",
    "kernel": "auto",
    "chunk": 50,
    "probe": None,
    "record_every": 10,
    "sponge": None,
}


class Solver(base_solver.BaseSolver):

    def __init__(self):
        super().__init__(logger)
        self.name = "hip"
        self.description = ("<p>MI355X (gfx950) HIP solver: fused stress/displacement sweeps, "
                            "FP64 reference arithmetic or FP32 production modes.</p>")
        self.cfg = {**self.cfg, **cfg}
        self.ctx = None
        self.probes = None
        self.elapsed = None

    def init(self, grid, material, steps):
        super().init(grid, material, steps)
        if self.ctx is not None:
            self.ctx.close()
            self.ctx = None
        spec = problem.mesh_spec(self.g, self.m, self.source_table(self.t), self.cfg)
        self.ctx = fdtd.Context(spec, precision=self.cfg["precision"], kernel=self.cfg["kernel"])
        self.probes = None

    def _recorder(self):
        """Per-step frames from device snapshots into the Writer (HDF,
        'thread'/'process') or a .npy directory ('npy', at self.file + '.d')."""
        if self.cfg["write_mode"] == "npy":
            nodes = (self.g.x.size, self.g.y.size, self.g.z.size)
            big = int(np.prod(nodes)) > 2_000_000  # dense 6x6 stiffness: 288 B per node
            sink = recording.NpySink(self.file + ".d", nodes, self.t,
                                     recording.schema_attrs(self.g, self.m, self.t, self.cfg),
                                     density=self.m.P, elasticity=None if big else self.m.C)
        else:
            sink = recording.WriterSink(self.writer)
        return recording.AsyncRecorder(self.ctx, sink, slots=int(self.cfg.get("record_slots", 4)))

    def run(self, *args, **kwargs):
        signals = kwargs.get("signals") or WorkerSignals()
        if self.ctx is None:
            raise RuntimeError("Solver.run() before init()")
        writing = self.cfg["write_mode"] != "off"
        chunk = 1 if writing else max(1, int(self.cfg.get("chunk", 50)))
        rec = self._recorder() if writing else None
        signals.status.emit("Solver starting..")
        self.running.set()
        signals.status.emit("Running simulation.")
        stime = time()
        progress = 0
        signals.progress.emit(0)
        done = self.ctx.steps_done
        while done < self.t:
            if not self.running.is_set():
                self.logger.warning("Simulation cancelled.")
                break
            n = min(chunk, self.t - done)
            self.ctx.step(n)
            done += n
            if writing:
                rec.record()  # frame done-1, copied while the next steps run
            else:
                self.ctx.sync()
            if progress < 99:
                progress = min(99, int((done / self.t) * 100))
                signals.progress.emit(progress)
        if writing:
            rec.close()
        self.ctx.sync()
        self.running.clear()
        self._pull_fields(old=True)
        self.probes = self.ctx.probes()
        if writing and self.cfg["write_mode"] != "npy":
            self.writer.notify_finished()
            self.writer.join(timeout=5 * 60)
        self.elapsed = time() - stime
        signals.status.emit("Simulation finished in {:.2f}s.".format(self.elapsed))
        signals.progress.emit(100)

    def _pull_fields(self, old=False):
        for c in ("ux", "uy", "uz"):
            getattr(self.g, c)[...] = self.ctx.get(c)
            if old:
                # after the reference's time_step, u_new == u and u_old is
                # the previous level
                getattr(self.g, c + "_new")[...] = getattr(self.g, c)
                getattr(self.g, c + "_old")[...] = self.ctx.get(c + "_old")

    def close(self):
        if self.ctx is not None:
            self.ctx.close()
            self.ctx = None
