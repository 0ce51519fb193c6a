"""Mesh and field store: the host side of phonomena's ``Grid``.

Mirrors the public interface of ``phonomena/simulation/grid.py`` (attributes
``x, y, z, fdx..sdz, ux, uy, uz, targets, size_*, max_d*, min_d, slope,
SI_conversion`` and methods ``init, update, freezeData, clearMesh,
clearInclusions, addInclusion, inclusionIndices, buildMesh``) so the mesher,
GUI and scripts that drive the reference drive this one unchanged.

Differences, each deliberate:

* ``update`` allocates only the three result arrays ``ux, uy, uz``; the
  reference's ``*_new/_old/_temp`` and ``T1..T6`` (grid.py:92-110) exist as
  lazily allocated zeros, because the GPU solver keeps the time levels on the
  device and never stores stresses (at 512x512x256 the reference's 18 host
  arrays would be 9.7 GB of zeros).
* ``setSpacing`` (extension): a uniform grid with a physical node spacing in
  metres; ``size_*`` then count cells and node coordinates are
  ``i * d``.  The reference can only express coordinates numerically equal to
  metres (``SI_conversion = 1``, grid.py:46), i.e. 1 m cells.
* ``buildMesh`` raises instead of looping forever when the function mesh
  cannot grow past ``max_d`` (slope 0 with ``min_d < max_d``; the
  reference's ``fineMesh`` loop at grid.py:186-194 never ends there).
"""
import math

import numpy as np

from ..gui.worker import WorkerSignals

DTYPE = np.float64

_LAZY = ("ux_new", "uy_new", "uz_new", "ux_old", "uy_old", "uz_old",
         "ux_temp", "uy_temp", "uz_temp", "T1", "T2", "T3", "T4", "T5", "T6")


class FrozenGrid:
    """Copies of the three displacement arrays (grid.py:23-27)."""
    ux = None
    uy = None
    uz = None


def _merge(base, extra):
    """Sorted union keeping ``base`` as is (grid.py:179-184 ``appendSorted``)."""
    new = np.setdiff1d(extra, base)
    out = np.append(base, new)
    out.sort()
    return out


class Grid:
    """Node coordinates, spacings, inclusion list and result arrays."""

    def __init__(self):
        self.trgt_dtype = np.dtype([('x', 'f'), ('y', 'f'), ('z', 'f'), ('r', 'f')])
        self.max_dx = 1
        self.max_dy = 1
        self.max_dz = 1
        self.min_d = 1
        self.slope = 1
        self.SI_conversion = 1
        self.spacing = None
        self.clearMesh()
        self.clearInclusions()

    # ------------------------------------------------------------ set-up --
    def init(self, size_x, size_y, size_z, fn='linear'):
        """grid.py:51-66: sizes (truncated to int), spacing function, mesh."""
        spacing_fns = {'linear': lambda v: abs(v * self.slope)}
        self.size_x = int(size_x)
        self.size_y = int(size_y)
        self.size_z = int(size_z)
        self.spacing_fn = spacing_fns[fn]
        self.buildMesh()
        self.update()

    def setSpacing(self, dx, dy=None, dz=None):
        """Extension: uniform grid with node spacing (dx, dy, dz) in metres."""
        dy = dx if dy is None else dy
        dz = dx if dz is None else dz
        self.spacing = (float(dx), float(dy), float(dz))

    def extent(self, axis):
        """Physical length of an axis (``size_*`` times the spacing)."""
        size = {'x': self.size_x, 'y': self.size_y, 'z': self.size_z}[axis]
        if self.spacing is None:
            return size
        return size * self.spacing['xyz'.index(axis)]

    def freezeData(self):
        fg = FrozenGrid()
        fg.ux = np.copy(self.ux)
        fg.uy = np.copy(self.uy)
        fg.uz = np.copy(self.uz)
        return fg

    def update(self):
        """Allocate result arrays and compute spacings (grid.py:79-127)."""
        nx, ny, nz = self.x.size, self.y.size, self.z.size
        self.ux = np.zeros((nx - 1, ny, nz), dtype=DTYPE)
        self.uy = np.zeros((nx, ny - 1, nz), dtype=DTYPE)
        self.uz = np.zeros((nx, ny, nz - 1), dtype=DTYPE)
        for name in _LAZY:
            self.__dict__.pop(name, None)
        si = self.SI_conversion
        xs, ys, zs = self.x * si, self.y * si, self.z * si
        self.fdx = (xs[1:] - xs[:-1]).reshape((nx - 1, 1, 1))
        self.fdy = (ys[1:] - ys[:-1]).reshape((1, ny - 1, 1))
        self.fdz = (zs[1:] - zs[:-1]).reshape((1, 1, nz - 1))
        # staggered spacing: mean of adjacent full spacings
        self.sdx = np.mean([self.fdx[1:, :, :], self.fdx[:-1, :, :]], axis=0)
        self.sdy = np.mean([self.fdy[:, 1:, :], self.fdy[:, :-1, :]], axis=0)
        self.sdz = np.mean([self.fdz[:, :, 1:], self.fdz[:, :, :-1]], axis=0)

    def __getattr__(self, name):
        # lazily allocated reference arrays (see the module docstring)
        if name in _LAZY and 'x' in self.__dict__:
            nx, ny, nz = self.x.size, self.y.size, self.z.size
            shp = {'x': (nx - 1, ny, nz), 'y': (nx, ny - 1, nz), 'z': (nx, ny, nz - 1)}
            if name.startswith('T'):
                shape = {'T1': (nx, ny, nz), 'T2': (nx, ny, nz), 'T3': (nx, ny, nz),
                         'T4': (nx, ny - 1, nz - 1), 'T5': (nx - 1, ny, nz - 1),
                         'T6': (nx - 1, ny - 1, nz)}[name]
            else:
                shape = shp[name[1]]
            arr = np.zeros(shape, dtype=DTYPE)
            self.__dict__[name] = arr
            return arr
        raise AttributeError(name)

    def clearMesh(self):
        self.x = np.array((), dtype=DTYPE)
        self.y = np.array((), dtype=DTYPE)
        self.z = np.array((), dtype=DTYPE)

    def clearInclusions(self):
        self.targets = np.empty((0, 0), self.trgt_dtype)

    def addInclusion(self, x, y, r, z=None):
        """Cylinder of radius r at (x, y), from z = 0 to depth z (grid.py:144-156)."""
        x = float(x)
        y = float(y)
        r = float(r)
        assert x + r < self.extent('x') and x - r > 0
        assert y + r < self.extent('y') and y - r > 0
        if z is None:
            z = self.size_z
        target = np.array([(x, y, z, r)], dtype=self.trgt_dtype)
        self.targets = np.append(self.targets, target)

    def inclusionIndices(self):
        """Node indices inside each inclusion (grid.py:158-175).

        Returns, per target, the (y, x) index pairs with distance < r
        (strict, compared in float64 against the float32-stored target) and
        the z indices.  As in the reference, the z test of every target after
        the first is applied to the previous target's index array rather than
        to the coordinates; with full-depth cylinders (z = size_z, the
        default) it selects every index either way.
        """
        out = []
        x, y, z = self.x, self.y, self.z
        X0, Y0 = np.meshgrid(x, y)
        for t in self.targets:
            X = X0 - t['x']
            Y = Y0 - t['y']
            R = np.sqrt(np.add(np.square(X), np.square(Y)))
            yx = np.array(np.where(R < t['r'])).transpose()
            z = np.array(np.where(z <= t['z'])).flatten()
            out.append((yx, z))
        return out

    # ------------------------------------------------------------ mesher --
    def buildMesh(self, *args, **kwargs):
        """Node coordinates per axis (grid.py:177-303), or uniform spacing."""
        signals = kwargs.get('signals') or WorkerSignals()
        signals.status.emit("Rebuilding mesh...")
        self.clearMesh()
        if self.spacing is not None:
            dx, dy, dz = self.spacing
            self.x = np.arange(self.size_x + 1, dtype=DTYPE) * dx
            self.y = np.arange(self.size_y + 1, dtype=DTYPE) * dy
            self.z = np.arange(self.size_z + 1, dtype=DTYPE) * dz
            signals.status.emit("Completed mesh.")
            return
        for axis in ('x', 'y'):
            size = getattr(self, 'size_' + axis)
            mesh = _merge(getattr(self, axis), 0)
            mesh = _merge(mesh, size)
            if len(self.targets) > 0:
                mesh = _merge(mesh, self._fill_inclusions(axis))
                mesh = _merge(mesh, self._function_mesh(axis))
                mesh = self._remove_close(mesh)
            setattr(self, axis, mesh)
            mesh = _merge(mesh, self._closest_fit(axis))
            setattr(self, axis, mesh)
        n = int(self.size_z / self.max_dz)
        self.z = np.linspace(0, self.size_z, n + 1, dtype=float)
        signals.status.emit("Completed mesh.")

    def _fine_mesh(self, max_d):
        # graded offsets 0, d1, d1+d2, ... with d = spacing_fn(prev + min_d) + min_d
        pts = np.array((0,))
        step = 0
        while step < max_d:
            step = self.spacing_fn(pts[-1] + self.min_d) + self.min_d
            if step < max_d:
                pts = np.append(pts, pts[-1] + step)
                if pts.size > 1_000_000:
                    raise ValueError("function mesh does not reach max_d (slope 0 and min_d < max_d?)")
        return pts

    def _remove_close(self, mesh):
        gaps = np.absolute(np.subtract(mesh[0:-1], mesh[1:]))
        drop = np.add(np.where(gaps < 0.9 * self.min_d), 1)
        return np.delete(mesh, drop)

    def _function_mesh(self, axis):
        max_d = {'x': self.max_dx, 'y': self.max_dy}[axis]
        size = {'x': self.size_x, 'y': self.size_y}[axis]
        fine = self._fine_mesh(max_d)
        t = np.sort(self.targets, order=axis)
        mesh = np.array(())
        for i in range(len(t)):
            c, r = t[i][axis], t[i]['r']
            if c == t[0][axis]:
                lo = 0
            elif len(t) > 1:
                lo = c - c - (c - t[i - 1][axis]) / 2
            else:
                raise Exception("Neither condition met")
            if c == t[-1][axis]:
                hi = size
            elif len(t) > 1:
                hi = (t[i + 1][axis] - c) / 2 + c + r
            else:
                raise Exception("Neither condition met")
            mesh = _merge(mesh, c + r + fine)
            mesh = _merge(mesh, c - r - fine)
            mesh = mesh[np.greater_equal(mesh, 0)]
            mesh = mesh[np.greater_equal(mesh, lo)]
            mesh = mesh[np.less_equal(mesh, hi)]
            mesh = mesh[np.less_equal(mesh, size)]
        return mesh

    def _closest_fit(self, axis):
        base = getattr(self, axis)
        max_d = {'x': self.max_dx, 'y': self.max_dy}[axis]
        mesh = np.array(())
        for i in range(1, len(base)):
            n = int(abs(base[i] - base[i - 1]) / max_d)
            if n > 0:
                mesh = _merge(mesh, np.linspace(base[i - 1], base[i], n + 1, dtype=float))
        return mesh

    def _fill_inclusions(self, axis):
        mesh = np.array(())
        for t in self.targets:
            n = int(math.floor(t['r'] * 2 / self.min_d))
            step = t['r'] * 2 / n
            pts = np.fromfunction(lambda v: step * v + t[axis] - t['r'], (n + 1,))
            mesh = _merge(mesh, pts)
        return mesh
