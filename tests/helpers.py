"""Shared test helpers: golden fixtures as oracle problems and device specs."""
import json
import os

import numpy as np

from oracle import reference_math as rm
from oracle import brute
from phonomena_amd import fdtd, problem
from phonomena_amd.fdtd import MeshSpec

FIELDS = ["ux", "uy", "uz", "ux_old", "uy_old", "uz_old"]
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["testdefaults_10", "homog16_200", "default_json_100", "nonuniform_json_60", "fine_json_20"]


def golden(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


def record_every(name):
    return 1 if name == "testdefaults_10" else 10


def oracle_problem(d, wave=None, sponge=None):
    """Oracle Problem (dense reference layout) for a fixture's inputs."""
    cfg = json.loads(str(d["cfg"]))
    C, P = rm.dense_material(d["ids"], [d["C_primary"], d["C_secondary"]],
                             [d["p_primary"], d["p_secondary"]])
    dt = d["dt"][()]
    steps = int(d["steps"])
    if wave is None:
        wave = rm.source_table(cfg["wave"], steps, dt, cfg["wave_args"])
    return rm.Problem(d["x"], d["y"], d["z"], C, P, dt, wave, sponge=sponge)


def rows_of(d):
    return brute.material_table([d["C_primary"], d["C_secondary"]], [d["p_primary"], d["p_secondary"]])


def spec_from(x, y, z, ids, rows, dt, wave, probe=None, every=1, sponge=None, cap=None):
    """Device MeshSpec built the way the product builds it (problem.py)."""
    fdx, fdy, fdz = np.diff(x), np.diff(y), np.diff(z)
    sd = [np.mean([f[1:], f[:-1]], axis=0) for f in (fdx, fdy, fdz)]
    fd = [fdx, fdy, fdz]
    dt = np.float64(dt)
    mur = problem.mur_coefficients(rows[int(ids[0, 0, 0])], dt, fd, sd)
    nodes = (x.size, y.size, z.size)
    if cap is None:
        cap = max(1, len(wave) // max(every, 1))
    return MeshSpec(nodes, fd, sd, rows, ids, dt, dt ** 2, mur, wave, sponge=sponge,
                    probe=probe, record_every=every, probe_capacity=cap)


def spec_from_golden(d, wave=None, probe="full", sponge=None):
    cfg = json.loads(str(d["cfg"]))
    dt = d["dt"][()]
    steps = int(d["steps"])
    if wave is None:
        wave = rm.source_table(cfg["wave"], steps, dt, cfg["wave_args"])
    x, y = d["x"], d["y"]
    pr = (0, 1, x.size, 0, 1, y.size) if probe == "full" else probe
    return spec_from(x, y, d["z"], d["ids"], rows_of(d), dt, wave, probe=pr,
                     every=1 if steps <= 10 else 10, sponge=sponge)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    n = np.linalg.norm(b)
    if n == 0:
        return float(np.linalg.norm(a))
    return float(np.linalg.norm(a - b) / n)


def uniform_crystal(cells, d, rho=2700.0, lam=5.1e10, mu=2.6e10, incl=None, steps=100,
                    wave_kind="sin", f=None, courant=0.5, seed=0):
    """Uniform-spacing problem (x, y, z, ids, rows, dt, wave) for tests.

    incl: None or (rho, lam, mu, a_cells, r_cells) square lattice of full-depth
    cylinders with +-0.5 cell centre jitter drawn from ``seed``.
    """
    nx, ny, nz = (c + 1 for c in cells)
    x = np.arange(nx) * d
    y = np.arange(ny) * d
    z = np.arange(nz) * d
    mats = [(rho, lam, mu)]
    ids = np.zeros((nx, ny, nz), dtype=np.uint8)
    if incl is not None:
        r2, l2, m2, a, r = incl
        mats.append((r2, l2, m2))
        rng = np.random.default_rng(seed)
        X, Y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
        for cx in np.arange(a / 2, cells[0], a):
            for cy in np.arange(a / 2, cells[1], a):
                jx, jy = rng.uniform(-0.5, 0.5, 2)
                px, py = cx + jx, cy + jy
                if px - r <= 0 or px + r >= cells[0] or py - r <= 0 or py + r >= cells[1]:
                    continue
                ids[(X - px) ** 2 + (Y - py) ** 2 < r * r, :] = 1
    rows = []
    for (p, l, m) in mats:
        c = np.zeros((6, 6))
        c[:3, :3] = l
        for a_ in range(3):
            c[a_, a_] = l + 2 * m
            c[3 + a_, 3 + a_] = m
        rows.append(c)
    table = brute.material_table(rows, [m[0] for m in mats])
    vmax = max(max(np.sqrt(t[0] / t[12]), np.sqrt(t[9] / t[12])) for t in table)
    vmix = max(np.sqrt(a_[0] / b_[12]) for a_ in table for b_ in table)
    if vmix > 4 * vmax:  # air-like voids: Courant of the fastest pairing
        courant = courant * vmax / vmix
    dt = np.float64(courant * d / vmax)
    if f is None:
        f = 1.0 / (40 * dt)
    wave = rm.source_table(wave_kind, steps, dt, {"f": f})
    return x, y, z, ids, table, dt, wave


class ReferenceStyleMaterial:
    """Stand-in with only the attributes of the reference's Material
    (phonomena/simulation/material.py:22-93): ``primary``, ``secondary``
    (property dicts, ``c`` in Pa after init), their keys, ``c_max``, ``dt``
    and the dense per-node arrays ``C`` (nx, ny, nz, 6, 6) and ``P``.  No
    ``rows()`` and no ``ids``: what the reference's loader hands a solver
    plugin (common.py:141-147)."""

    def __init__(self):
        self.primary = self.primary_key = None
        self.secondary = self.secondary_key = None
        self.c_max = 0.5
        self.dt = 0
        self.C = np.zeros((0, 0, 0, 6, 6))
        self.P = np.zeros((0, 0, 0))

    def init(self, grid, properties):
        import copy
        self.grid = grid
        self.properties = dict(copy.deepcopy(properties))
        for val in self.properties.values():
            val["c"] = np.array(val["c"]) * 1e10
        first = list(self.properties)[0]
        self.setPrimary(first)
        self.setSecondary(first)

    def setPrimary(self, key):
        self.primary, self.primary_key = self.properties[key], key

    def setSecondary(self, key):
        self.secondary, self.secondary_key = self.properties[key], key

    def setTimeStep(self):
        g = self.grid
        h = min(np.amin(g.fdx), np.amin(g.fdy), np.amin(g.fdz)) * g.SI_conversion
        steps = []
        for mat in (self.primary, self.secondary):
            v = max(np.sqrt(mat["c"][0][0] / mat["p"]), np.sqrt(mat["c"][3][3] / mat["p"]))
            steps.append(self.c_max * h / v)
        self.dt = min(steps)

    def update(self):
        g = self.grid
        shape = (g.x.size, g.y.size, g.z.size)
        self.C = np.zeros(shape + (6, 6))
        self.P = np.zeros(shape)
        self.setTimeStep()
        self.C[...] = np.array(self.primary["c"])
        self.P[...] = float(self.primary["p"])
        for yx, z in g.inclusionIndices():
            for y, x in yx:
                self.C[x, y, z] = np.array(self.secondary["c"])
                self.P[x, y, z] = float(self.secondary["p"])


def reference_style_material(grid, d):
    """ReferenceStyleMaterial of a golden fixture's two materials (stiffness
    given in units of 1e10 Pa, as the reference's JSON holds it)."""
    m = ReferenceStyleMaterial()
    props = {"M0": {"name": "m0", "p": float(d["p_primary"]), "c": (d["C_primary"] / 1e10).tolist()},
             "M1": {"name": "m1", "p": float(d["p_secondary"]), "c": (d["C_secondary"] / 1e10).tolist()}}
    m.init(grid, props)
    # the fixture's stiffness in Pa, bit for bit (no round trip through 1e10)
    m.properties["M0"]["c"] = np.array(d["C_primary"], dtype=np.float64)
    m.properties["M1"]["c"] = np.array(d["C_secondary"], dtype=np.float64)
    m.setPrimary("M0")
    m.setSecondary("M1")
    return m


def random_start(nodes, seed):
    """Random level n (and n-1 = half of it) of every component; the x = 0
    wall of uy, uz stays 0 as in the reference."""
    rng = np.random.default_rng(seed)
    start = {c: rng.standard_normal(s) for c, s in (("ux", (nodes[0] - 1,) + tuple(nodes[1:])),
                                                   ("uy", (nodes[0], nodes[1] - 1, nodes[2])),
                                                   ("uz", (nodes[0], nodes[1], nodes[2] - 1)))}
    for c in ("uy", "uz"):
        start[c][0] = 0.0
    return start


def run_from(spec, start, precision, kernel, steps, **tiles):
    ctx = fdtd.Context(spec, precision=precision, kernel=kernel, **tiles)
    for c in ("ux", "uy", "uz"):
        ctx.set(c, start[c])
        ctx.set(c + "_old", 0.5 * start[c])
    ctx.step(steps)
    ctx.sync()
    out = {f: ctx.get(f) for f in FIELDS}
    out["probes"] = ctx.probes()
    ctx.close()
    return out
