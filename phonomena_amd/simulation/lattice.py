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


def pairing_speed(m):
    """Fastest speed the stencil can realise: sqrt(C11 of one material /
    rho of the other).  The reference samples stiffness and density at
    different nodes without averaging (base_solver.py:353-367, 443-459), so
    next to an interface a node can combine the stiffness of one material
    with the density of the other."""
    mats = [m.primary, m.secondary]
    return max(np.sqrt(np.asarray(a['c'])[0, 0] / float(b['p'])) for a in mats for b in mats)


def reference_speed(m):
    """vmax of the reference's CFL rule (material.py:80-93)."""
    return max(max(np.sqrt(np.asarray(x['c'])[0, 0] / float(x['p'])),
                   np.sqrt(np.asarray(x['c'])[3, 3] / float(x['p']))) for x in (m.primary, m.secondary))


def stable_courant(m, courant=0.5, ratio=4.0):
    """Courant number for Material.c_max.

    The reference's rule dt = c_max * dx_min / vmax is kept.  When the
    stiffness/density pairing of the two materials is more than ``ratio``
    times faster than vmax (air-like voids in a solid: ~47x), c_max is scaled
    so that dt = courant * dx_min / pairing_speed; with the reference's
    c_max = 0.5 such crystals diverge within a few hundred steps.
    """
    v_ref, v_mix = reference_speed(m), pairing_speed(m)
    if v_mix > ratio * v_ref:
        return courant * v_ref / v_mix
    return courant


def crystal(cells, a, r, seed, inclusion=TUNGSTEN, matrix=MATRIX, jitter=0.5, courant=0.5, d=1.0, host_ids=True):
    """(Grid, Material) of a crystal of ``cells`` = (Nx, Ny, Nz) cells.

    Coordinates are in cell units by default (d = 1), the way the reference's
    own settings are written (SI_conversion = 1, one unit per cell): the
    scheme then runs exactly as the reference runs it.  A physical cell size
    dx maps onto time: t_phys = t * dx and f_phys = f / dx.  (At physical
    coordinates such as dx = 5e-5 the reference's free-surface u_z update,
    which adds T3[..,1] without dividing by dz (base_solver.py:516), makes
    the scheme diverge; at unit spacing that term is what was intended.)
    With host_ids False the material map is not built on the host (the
    device rasterizes it from the inclusion list).
    """
    g = _grid.Grid()
    g.setSpacing(d)
    g.init(*cells)
    for (px, py) in lattice_sites(cells, a, r, seed, jitter):
        g.addInclusion(px * d, py * d, r * d)
    m = _material.Material()
    m.init(g, {"matrix": _material.isotropic(*matrix), "inclusion": _material.isotropic(*inclusion)})
    m.setPrimary("matrix")
    m.setSecondary("inclusion")
    m.c_max = stable_courant(m, courant)
    m.update(rasterize=host_ids)
    return g, m


def fill_fraction(m):
    """Fraction of nodes in the secondary material."""
    return float(m.ids.mean()) if m.ids.size else 0.0
