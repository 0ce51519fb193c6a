"""Seeded phononic-crystal generator (extension; the reference builds its
inclusion list by hand from JSON or the GUI table).

A square lattice of vertical cylinders, lattice constant ``a`` cells, radius
``r`` cells, centres at a*(m + 1/2) with a uniform +-``jitter`` cell offset
per axis drawn from ``seed``.  Sites whose cylinder would cross the domain
edge are dropped, the rule ``Grid.addInclusion`` asserts
(phonomena/simulation/grid.py:151-152).  The result is an ordinary Grid /
Material pair: inclusions go through ``addInclusion`` and the material map
through ``Material.update`` -> ``inclusionIndices``, so the lattice is
rasterised exactly as a hand-entered inclusion list would be.
"""
import numpy as np

from . import grid as _grid
from . import material as _material

MATRIX = ("matrix", 2700.0, 5.1e10, 2.6e10)
TUNGSTEN = ("tungsten-like", 19300.0, 2.0e11, 1.6e11)
VOID = ("void-like", 1.2, 1.4e5, 0.0)


def lattice_sites(cells, a, r, seed, jitter=0.5):
    """Cylinder centres (cell units) of a jittered square lattice."""
    rng = np.random.default_rng(seed)
    sites = []
    nx, ny = cells[0], cells[1]
    for cx in np.arange(a / 2.0, nx, a):
        for cy in np.arange(a / 2.0, ny, a):
            jx, jy = rng.uniform(-jitter, jitter, 2)
            px, py = cx + jx, cy + jy
            if px + r < nx and px - r > 0 and py + r < ny and py - r > 0:
                sites.append((px, py))
    return sites


def crystal(cells, d, a, r, seed, inclusion=TUNGSTEN, matrix=MATRIX, jitter=0.5, courant=0.5):
    """(Grid, Material) of a crystal with ``cells`` = (Nx, Ny, Nz) cells of
    size ``d`` metres; node counts are cells + 1 per axis."""
    g = _grid.Grid()
    g.setSpacing(d)
    g.init(*cells)
    for (px, py) in lattice_sites(cells, a, r, seed, jitter):
        g.addInclusion(px * d, py * d, r * d)
    m = _material.Material()
    m.init(g, {"matrix": _material.isotropic(*matrix), "inclusion": _material.isotropic(*inclusion)})
    m.setPrimary("matrix")
    m.setSecondary("inclusion")
    m.c_max = courant
    m.update()
    return g, m


def fill_fraction(m):
    """Fraction of nodes in the secondary material."""
    return float(m.ids.mean()) if m.ids.size else 0.0
