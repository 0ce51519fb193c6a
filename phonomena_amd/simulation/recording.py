"""Per-step field recording from device snapshots, asynchronously (the
output path of the reference's Writer, base_solver.py:72-169 and the
``writer.put`` of every step, :256-260).

The reference copies the three displacement arrays after every step and
hands them to a writer process.  Here every frame is a device snapshot
(``Context.snapshot``): the current level is packed to FP64 in the
reference's shapes on the context's stream and copied to pinned host memory
on a second stream, so the next steps run while it travels.  A writer thread
puts the frames into a sink in step order; the step loop waits only when all
snapshot slots are still in flight or being written.

Sinks keep the Writer's schema: datasets ``ux`` (nx-1, ny, nz, T), ``uy``,
``uz`` with frame ``tt`` at ``[..., tt]``, ``density`` and ``elasticity``,
and the attributes x, y, z, sdx..fdz, steps, dt, prim_material,
sec_material, solver_cfg.
  * ``WriterSink``: the reference's HDF Writer (h5py);
  * ``NpySink``: a directory of .npy datasets (memory-mapped) plus
    attrs.json, for machines without h5py.
"""
import collections
import json
import os
import queue
import threading

import numpy as np

ATTR_KEYS = ("x", "y", "z", "sdx", "sdy", "sdz", "fdx", "fdy", "fdz")


def schema_attrs(grid, material, steps, cfg):
    """The Writer's attributes (base_solver.py:105-133)."""
    a = {k: np.asarray(getattr(grid, k)) for k in ATTR_KEYS}
    a["steps"] = int(steps)
    a["dt"] = float(material.dt)
    a["prim_material"] = material.primary['name']
    a["sec_material"] = material.secondary['name']
    a["solver_cfg"] = json.dumps(cfg, default=str)
    return a


def frame_shapes(nodes):
    nx, ny, nz = nodes
    return {"ux": (nx - 1, ny, nz), "uy": (nx, ny - 1, nz), "uz": (nx, ny, nz - 1)}


class NpySink:
    """The Writer schema as a directory: ux.npy, uy.npy, uz.npy of shape
    (..., steps), density.npy, elasticity.npy (when given) and attrs.json."""

    def __init__(self, path, nodes, steps, attrs, density=None, elasticity=None):
        os.makedirs(path, exist_ok=True)
        self.path = path
        self.data = {}
        for name, shape in frame_shapes(nodes).items():
            self.data[name] = np.lib.format.open_memmap(os.path.join(path, name + ".npy"), mode="w+",
                                                        dtype=np.float64, shape=shape + (int(steps),))
        if density is not None:
            np.save(os.path.join(path, "density.npy"), np.asarray(density, dtype=np.float64))
        if elasticity is not None:
            np.save(os.path.join(path, "elasticity.npy"), np.asarray(elasticity, dtype=np.float64))
        with open(os.path.join(path, "attrs.json"), "w") as fh:
            json.dump({k: (v.tolist() if isinstance(v, np.ndarray) else v) for k, v in attrs.items()}, fh)

    def put(self, frame, tt):
        for name, arr in self.data.items():
            arr[..., tt] = frame[name]

    def close(self):
        for arr in self.data.values():
            arr.flush()
        self.data = {}


def load_npy(path):
    """(datasets, attrs) of an NpySink directory, datasets memory-mapped."""
    data = {}
    for name in ("ux", "uy", "uz", "density", "elasticity"):
        p = os.path.join(path, name + ".npy")
        if os.path.exists(p):
            data[name] = np.load(p, mmap_mode="r")
    with open(os.path.join(path, "attrs.json")) as fh:
        attrs = json.load(fh)
    return data, attrs


class _Frame:
    """What the reference's writer queue carries (Grid.freezeData)."""

    def __init__(self, frame):
        self.ux, self.uy, self.uz = frame["ux"], frame["uy"], frame["uz"]


class WriterSink:
    """Frames into the reference's HDF Writer (its queue and process)."""

    def __init__(self, writer):
        self.writer = writer

    def put(self, frame, tt):
        self.writer.put((_Frame({k: np.copy(v) for k, v in frame.items()}), tt))

    def close(self):
        pass


class AsyncRecorder:
    """Frames of a context's steps into a sink, from device snapshots.

    Call ``record()`` after each step (or group of steps) whose frame is
    wanted; frame ``tt`` is the state after ``tt + 1`` steps, as the
    reference's ``writer.put((..., tt))``.  ``close()`` waits for every
    frame to be written.
    """

    def __init__(self, ctx, sink, slots=4):
        self.ctx = ctx
        self.sink = sink
        self.slots = int(slots)
        ctx.snapshot_open(self.slots)
        self.free = collections.deque(range(self.slots))
        self.inflight = collections.deque()   # slots queued on the device, in order
        self.written = {s: threading.Event() for s in range(self.slots)}
        self.held = collections.deque()       # slots handed to the writer thread
        self.q = queue.Queue()
        self.error = None
        self.thread = threading.Thread(target=self._write, daemon=True)
        self.thread.start()

    def _write(self):
        while True:
            item = self.q.get()
            if item is None:
                return
            slot, frame, tt = item
            try:
                if self.error is None:
                    self.sink.put(frame, tt)
            except Exception as e:  # surfaced by record()/close()
                self.error = e
            finally:
                self.written[slot].set()

    def _hand_over_oldest(self):
        # the oldest in-flight snapshot: wait for its copy, give it to the
        # writer thread
        slot = self.inflight.popleft()
        frame, steps = self.ctx.snapshot_wait(slot)
        self.written[slot].clear()
        self.held.append(slot)
        self.q.put((slot, frame, steps - 1))

    def _reclaim_oldest(self):
        slot = self.held.popleft()
        self.written[slot].wait()
        self.ctx.snapshot_release(slot)
        self.free.append(slot)

    def record(self):
        if self.error is not None:
            raise self.error
        if not self.free:
            if self.inflight:
                self._hand_over_oldest()
            if not self.free:
                self._reclaim_oldest()
        slot = self.free.popleft()
        self.ctx.snapshot(slot)
        self.inflight.append(slot)
        # keep the writer busy: hand over copies that have had a step to land
        while len(self.inflight) > 1:
            self._hand_over_oldest()
        while self.held and self.written[self.held[0]].is_set():
            self._reclaim_oldest()

    def close(self):
        while self.inflight:
            self._hand_over_oldest()
        while self.held:
            self._reclaim_oldest()
        self.q.put(None)
        self.thread.join()
        self.sink.close()
        if self.error is not None:
            raise self.error
