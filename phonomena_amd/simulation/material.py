"""Material model: the host side of phonomena's ``Material``.

Mirrors ``phonomena/simulation/material.py`` (``init, update,
setConstants, setPrimary, setSecondary, setTimeStep``; attributes
``primary, secondary, *_key, c_max, dt, C, P``).  The reference stores a
dense per-node 6x6 stiffness ``C[nx,ny,nz,6,6]`` and density ``P`` (288 B per
node; 19.5 GB at 513^2 x 257).  Every node holds either the primary or the
secondary material (material.py:55-63), so this class stores one byte per node
(``ids``: 0 = primary, 1 = secondary) plus the two material rows, and builds
``C``/``P`` only when someone reads them.

Extension: a property may give ``c_Pa`` (stiffness in Pa, used verbatim)
instead of ``c`` (in units of 1e10 Pa, multiplied by 1e10 on ``init`` as in
material.py:42).
"""
import copy
import logging

import numpy as np

logger = logging.getLogger(__name__)

DYTPE = np.float64


class Material:

    def __init__(self):
        self.primary = None
        self.primary_key = None
        self.secondary = None
        self.secondary_key = None
        self.c_max = 0.5  # Courant number
        self.dt = 0
        self.ids = np.zeros((0, 0, 0), dtype=np.uint8)

    def init(self, grid, properties):
        """Copy the property table; stiffness to Pa (material.py:36-46)."""
        self.grid = grid
        self.properties = dict(copy.deepcopy(properties))
        for key, val in self.properties.items():
            if 'c_Pa' in val:
                val['c'] = np.array(val['c_Pa'], dtype=np.float64)
            else:
                val['c'] = np.array(val['c']) * 1e10
        keys = list(self.properties.keys())
        self.setPrimary(keys[0])
        self.setSecondary(keys[0])

    def update(self, rasterize=True):
        """Per-node material map and time step (material.py:48-53).  With
        rasterize False only the time step is set: the map is left to the
        device rasterizer (problem.mesh_spec(..., device_ids=True))."""
        g = self.grid
        self.setTimeStep()
        if not rasterize:
            self.ids = np.zeros((0, 0, 0), dtype=np.uint8)
            return
        self.ids = np.zeros((g.x.size, g.y.size, g.z.size), dtype=np.uint8)
        self.setConstants()

    def setConstants(self):
        """Primary everywhere, secondary inside inclusions (material.py:55-63)."""
        self.ids[...] = 0
        for yx, z in self.grid.inclusionIndices():
            if len(yx) == 0 or len(z) == 0:
                continue
            ys, xs = yx[:, 0], yx[:, 1]
            self.ids[xs[:, None], ys[:, None], z[None, :]] = 1

    def setPrimary(self, m):
        if m in self.properties.keys():
            self.primary = self.properties[m]
            self.primary_key = m
        else:
            raise Exception()

    def setSecondary(self, m):
        if m in self.properties.keys():
            self.secondary = self.properties[m]
            self.secondary_key = m
        else:
            raise Exception()

    def setTimeStep(self):
        """CFL step from both materials (material.py:80-93).

        As in the reference, the smallest spacing is multiplied by
        ``SI_conversion`` although the spacings already include it; with the
        default SI_conversion = 1 this is a no-op.
        """
        def calc(c, p):
            vl = np.sqrt(c[0][0] / p)
            vt = np.sqrt(c[3][3] / p)
            vmax = max((vl, vt))
            g = self.grid
            dxmin = min((np.amin(g.fdx), np.amin(g.fdy), np.amin(g.fdz))) * g.SI_conversion
            return self.c_max * dxmin / vmax

        dt1 = calc(self.primary['c'], self.primary['p'])
        dt2 = calc(self.secondary['c'], self.secondary['p'])
        self.dt = min((dt1, dt2))

    # ------------------------------------------------- derived arrays --
    def rows(self):
        """(2, 13) rows [primary, secondary]: C00 C01 C02 C10 C11 C12 C20 C21
        C22 C33 C44 C55 rho, the entries of the 6x6 stiffness the stencils use."""
        out = []
        for m in (self.primary, self.secondary):
            c = np.asarray(m['c'], dtype=np.float64)
            out.append([c[0, 0], c[0, 1], c[0, 2], c[1, 0], c[1, 1], c[1, 2],
                        c[2, 0], c[2, 1], c[2, 2], c[3, 3], c[4, 4], c[5, 5], float(m['p'])])
        return np.array(out, dtype=np.float64)

    @property
    def C(self):
        """Dense per-node stiffness (nx, ny, nz, 6, 6), as the reference stores it."""
        tab = np.stack([np.asarray(self.primary['c'], dtype=np.float64),
                        np.asarray(self.secondary['c'], dtype=np.float64)])
        return tab[self.ids]

    @property
    def P(self):
        """Dense per-node density (nx, ny, nz)."""
        tab = np.array([float(self.primary['p']), float(self.secondary['p'])])
        return tab[self.ids]


def isotropic(name, rho, lam, mu):
    """Property entry of an isotropic solid from (rho, lambda, mu) in SI units."""
    c = np.zeros((6, 6))
    c[:3, :3] = lam
    for a in range(3):
        c[a, a] = lam + 2 * mu
        c[3 + a, 3 + a] = mu
    return {"name": name, "p": float(rho), "c_Pa": c.tolist()}
