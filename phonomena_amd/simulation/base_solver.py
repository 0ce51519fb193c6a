"""Solver base class: the host side of phonomena's ``BaseSolver``.

Mirrors ``phonomena/simulation/base_solver.py``: the ``BaseSolver``
life cycle (``__init__(logger)``, ``init(grid, material, steps)``,
``run(*args, signals=...)``, ``cancel``, ``test``), the per-solver ``cfg``
dict, the ``Writer`` HDF schema, ``TestDefaults`` and the source functions
``update_sin``/``update_ricker``.

The per-step stencil hooks of the reference (``update_T``, ``apply_T_tfbc``,
``update_u``, ``apply_u_tfbc``, ``apply_u_abc``, ``time_step``,
base_solver.py:323-571) are part of the interface but have no host
implementation here: the GPU solver (``solvers/solver_hip.py``) fuses them
into device kernels and overrides ``run``.  Calling a hook on this base class
raises, so nothing can silently fall back to a CPU stencil.
"""
import copy
import json
import logging
import multiprocessing as mp
import os
import queue
import sys
import tempfile
import threading
from pathlib import Path
from time import time

import numpy as np

from . import grid, material
from ..gui.worker import WorkerSignals

logger = logging.getLogger(__name__)

DTYPE = np.float64


class TestDefaults:
    """Minimal set-up for solver self-tests (base_solver.py:28-70).

    As in the reference, the GaAs stiffness here is already in Pa and
    ``Material.init`` multiplies it by 1e10 again, so C ~ 1e21 Pa; the run
    stays self-consistent because dt follows from C.
    """
    properties = {"GaAs": {
        "name": "Gallium Arsenide",
        "p": 5307,
        "c": np.array([
            [11.88, 5.87, 5.38, 0, 0, 0],
            [5.87, 11.88, 5.38, 0, 0, 0],
            [5.87, 5.38, 11.88, 0, 0, 0],
            [0, 0, 0, 5.94, 0, 0],
            [0, 0, 0, 0, 5.94, 0],
            [0, 0, 0, 0, 0, 5.94]]) * 1e10,
    }}

    g = grid.Grid()
    g.init(size_x=4, size_y=5, size_z=6)
    m = material.Material()
    m.init(grid=g, properties=properties)
    g.min_d = 1
    g.max_dx = 1
    g.max_dy = 1
    g.max_dz = 1
    g.slope = 0
    g.buildMesh()
    m.setPrimary("GaAs")
    m.setSecondary("GaAs")


class Writer:
    """HDF output in the reference schema (base_solver.py:72-169).

    Datasets ``ux`` (nx-1, ny, nz, T), ``uy``, ``uz`` chunked one frame per
    chunk, ``density``, ``elasticity``, and the grid/material attributes.
    h5py is imported on first use; without it, ``init`` raises.
    """

    def __init__(self, mode='thread'):
        assert mode in ['thread', 'process']
        self.queue = {'process': mp.JoinableQueue, 'thread': queue.Queue}[mode]()
        self.daemon = True
        self.mode = {'process': mp.Process, 'thread': threading.Thread}[mode]
        self.process = self.mode()
        self.file = None

    def join(self, *args, **kwargs):
        self.process.join(*args, **kwargs)
        if self.process.is_alive():
            raise mp.TimeoutError()

    def is_alive(self, *args, **kwargs):
        return self.process.is_alive(*args, **kwargs)

    def put(self, *args, **kwargs):
        if not self.process.is_alive():
            raise Exception("Process not started")
        self.queue.put(*args, **kwargs)

    def notify_finished(self):
        self.queue.put(None)

    def init(self, steps, material, grid, cfg, file):
        try:
            import h5py as h5
        except ImportError as e:
            raise RuntimeError("write_mode %r needs h5py, which is not installed; "
                               "use write_mode='off'" % cfg.get('write_mode')) from e
        x, y, z, t = grid.x.size, grid.y.size, grid.z.size, steps
        logger.info("Writing HDF to file {}".format(file))
        self.file = file
        with h5.File(file, mode='w') as hdf:
            hdf.create_dataset("ux", (x - 1, y, z, t), chunks=(x - 1, y, z, 1), dtype=DTYPE)
            hdf.create_dataset("uy", (x, y - 1, z, t), chunks=(x, y - 1, z, 1), dtype=DTYPE)
            hdf.create_dataset("uz", (x, y, z - 1, t), chunks=(x, y, z - 1, 1), dtype=DTYPE)
            hdf.create_dataset("density", data=material.P, dtype=DTYPE)
            hdf.create_dataset("elasticity", data=material.C, dtype=DTYPE)
            for k in ("x", "y", "z", "sdx", "sdy", "sdz", "fdx", "fdy", "fdz"):
                hdf.attrs[k] = getattr(grid, k)
            hdf.attrs["steps"] = steps
            hdf.attrs["dt"] = material.dt
            hdf.attrs["prim_material"] = material.primary['name']
            hdf.attrs["sec_material"] = material.secondary['name']
            hdf.attrs["solver_cfg"] = json.dumps(cfg)

    @staticmethod
    def run(file, queue_obj):
        import h5py as h5
        hdf = h5.File(file, mode='r+')
        ux, uy, uz = hdf.get('ux'), hdf.get('uy'), hdf.get('uz')
        while True:
            obj = queue_obj.get(timeout=2 * 60)
            if obj is None:
                break
            g, tt = obj
            ux[:, :, :, tt] = g.ux
            uy[:, :, :, tt] = g.uy
            uz[:, :, :, tt] = g.uz
        hdf.close()

    def start(self):
        if self.file is None:
            raise Exception("Must run init before start.")
        assert Path(self.file).exists()
        self.process = self.mode(target=self.run, args=(self.file, self.queue))
        self.process.daemon = self.daemon
        self.process.start()


class BaseSolver:
    """Life cycle shared by every solver plugin (base_solver.py:172-292)."""

    def __init__(self, logger):
        self.name = "default"
        self.description = '''
            <p>Base Solver. This string should be overwritten by the child class</p>'''
        with tempfile.NamedTemporaryFile() as f:
            self.file = os.path.realpath(f.name)
        self.running = threading.Event()
        self.writer = Writer()
        self.logger = logger
        self.cfg = {'wave': 'ricker',
                    'wave_args': {'f': 100},
                    'write_mode': 'process'}

    def init(self, grid, material, steps):
        """Own deep copies of grid/material, rebuild mesh, start the writer."""
        self.g = copy.deepcopy(grid)
        self.m = copy.deepcopy(material)
        self.t = copy.deepcopy(steps)
        logger.info("Initializing {} with settings: {}".format(self.logger.name, self.cfg))
        self.g.buildMesh()
        self.g.update()
        self.m.grid = self.g
        self.m.update()
        if self.writer.is_alive():
            self.writer.notify_finished()
            self.writer.join(timeout=1)
        # 'npy' (extension): the solver records into a directory sink itself
        if self.cfg['write_mode'] not in ('off', 'npy'):
            self.writer = Writer(self.cfg['write_mode'])
            self.writer.init(steps=self.t, file=self.file, grid=self.g, cfg=self.cfg, material=self.m)
            self.writer.start()

    def run(self, *args, **kwargs):
        """Step loop of base_solver.py:224-280 over the per-step hooks."""
        signals = kwargs.get('signals') or WorkerSignals()
        wave_fn = self.wave_function()
        signals.status.emit("Solver starting..")
        self.running.set()
        signals.status.emit("Running simulation.")
        stime = time()
        progress = 0
        signals.progress.emit(0)
        for tt in range(self.t):
            if not self.running.is_set():
                self.logger.warning("Simulation cancelled.")
                break
            self.g.uz[0, :, 0] = wave_fn(tt=tt, **self.cfg['wave_args'])
            self.update_T()
            self.update_T_BC()
            self.update_u()
            self.update_u_BC()
            self.time_step()
            if self.cfg['write_mode'] != 'off':
                self.writer.put((self.g.freezeData(), tt))
            if progress < 99:
                progress = min(99, int(((tt + 1) / self.t) * 100))
                signals.progress.emit(progress)
        self.running.clear()
        if self.cfg['write_mode'] != 'off':
            self.writer.notify_finished()
            self.writer.join(timeout=5 * 60)
        signals.status.emit("Simulation finished in {:.2f}s.".format(time() - stime))
        signals.progress.emit(100)

    def cancel(self):
        if self.running.is_set():
            self.running.clear()

    def test(self):
        self.init(grid=TestDefaults.g, material=TestDefaults.m, steps=10)
        self.run()

    # ------------------------------------------------------------ source --
    def wave_function(self):
        return {'sin': self.update_sin, 'ricker': self.update_ricker,
                'gauss': self.update_gauss}[self.cfg['wave']]

    def update_sin(self, tt, f, **kwargs):
        """sin(2 pi f tt dt) (base_solver.py:294-299)."""
        return np.sin(2 * np.pi * f * tt * self.m.dt)

    def update_ricker(self, tt, f, source_delay=0, **kwargs):
        """Ricker wavelet (1 - 2a) exp(-a), a = (pi f (dt tt - delay))^2 (:301-312)."""
        arg = (np.pi * f * (self.m.dt * tt - source_delay)) ** 2
        return (1 - 2 * arg) * np.exp(-arg)

    def update_gauss(self, tt, f, source_delay=0, width=None, **kwargs):
        """Extension: Gaussian-modulated sinusoid
        sin(2 pi f (t - t0)) exp(-((t - t0)/s)^2 / 2), s = 1/f, t0 = 4 s."""
        s = (1.0 / f) if width is None else width
        t0 = 4.0 * s if not source_delay else source_delay
        t = tt * self.m.dt - t0
        return np.sin(2 * np.pi * f * t) * np.exp(-0.5 * (t / s) ** 2)

    def source_table(self, steps):
        """w(tt) for every step, FP64, one scalar evaluation per step exactly
        as the reference's loop evaluates it (base_solver.py:251)."""
        fn = self.wave_function()
        args = self.cfg['wave_args']
        return np.array([fn(tt=tt, **args) for tt in range(steps)], dtype=np.float64)

    # ------------------------------------------------------ stencil hooks --
    def _no_host_stencil(self, hook):
        raise NotImplementedError(
            "%s.%s: the stencil runs on the GPU inside run(); there is no host "
            "implementation of the per-step hooks" % (type(self).__name__, hook))

    def update_T(self):
        self._no_host_stencil("update_T")

    def update_T_BC(self):
        self.apply_T_tfbc()

    def apply_T_tfbc(self):
        self._no_host_stencil("apply_T_tfbc")

    def update_u(self):
        self._no_host_stencil("update_u")

    def update_u_BC(self):
        self.apply_u_tfbc()
        self.apply_u_abc()

    def apply_u_tfbc(self):
        self._no_host_stencil("apply_u_tfbc")

    def apply_u_abc(self):
        self._no_host_stencil("apply_u_abc")

    def time_step(self):
        self._no_host_stencil("time_step")


if __name__ == '__main__':
    sys.exit(0)
