"""Design-sweep jobs: seeded crystals as device-ready mesh specs.

A job is one phononic crystal (a square lattice of vertical cylinders in a
matrix, ``simulation.lattice.crystal``) with a Gaussian-modulated 2 MHz wall
source and a step budget.  Coordinates are the reference's unit-cell ones
(one unit per cell); a physical cell of ``DX_PHYS`` metres maps the source
frequency to f0 * DX_PHYS per unit time.

``mix(...)`` draws the batched-sweep mix (uneven grids, lattice constants,
fill fractions, void-like or tungsten-like inclusions, from one seed);
``random_job(rng, ...)`` draws one job of a served stream.
"""
import numpy as np

from . import problem
from .simulation import lattice
from .simulation.base_solver import BaseSolver

F0 = 2e6           # source centre frequency, Hz
DX_PHYS = 5e-5     # physical cell size, m
INCLUSIONS = {"void": lattice.VOID, "tungsten": lattice.TUNGSTEN}


def crystal_spec(cells, a, r, seed, steps, inclusion="tungsten", probe=None, every=10, sponge=None,
                 kind="gauss", capacity=None, device_ids=False, d=1.0, f0=F0, dx_phys=DX_PHYS,
                 free_surface="reference"):
    """(MeshSpec, Grid, Material) of one crystal job.  device_ids: the
    material map is rasterized on the device (the Material's map is left
    empty).  d: the cell size in the grid's coordinates (1 = unit cells; a
    physical size such as dx_phys needs free_surface "fixed", see
    problem.free_surface); the source frequency is f0 * dx_phys / d per unit
    time, the same per step at every d.  inclusion: a name of INCLUSIONS or
    a (name, rho, lambda, mu) tuple (the matrix's own gives a homogeneous
    box)."""
    inc = INCLUSIONS[inclusion] if isinstance(inclusion, str) else inclusion
    g, m = lattice.crystal(tuple(int(c) for c in cells), float(a), float(r), int(seed), inclusion=inc,
                           host_ids=not device_ids, d=float(d))
    b = BaseSolver(None)
    b.m = m
    b.cfg["wave"], b.cfg["wave_args"] = kind, {"f": f0 * dx_phys / float(d)}
    wave = b.source_table(int(steps))
    cfg = {"probe": probe, "record_every": every, "sponge": sponge, "free_surface": free_surface}
    if capacity:
        cfg["probe_capacity"] = int(capacity)
    spec = problem.mesh_spec(g, m, wave, cfg, device_ids=device_ids)
    return spec, g, m


def random_job(rng, nxy=(48, 192), nz=(24, 96), steps=(2000, 2000), a=(8.0, 24.0), ra=(0.15, 0.45)):
    """One job's parameters: Nx, Ny ~ U{nxy}, Nz ~ U{nz} cells, steps ~
    U{steps}, lattice constant ~ U[a], r/a ~ U[ra], inclusion void-like or
    tungsten-like with p = 1/2, lattice jitter seed."""
    cells = (int(rng.integers(nxy[0], nxy[1] + 1)), int(rng.integers(nxy[0], nxy[1] + 1)),
             int(rng.integers(nz[0], nz[1] + 1)))
    lat = float(rng.uniform(*a))
    return {"cells": cells, "a": lat, "r": lat * float(rng.uniform(*ra)),
            "inclusion": "void" if rng.random() < 0.5 else "tungsten",
            "seed": int(rng.integers(0, 2 ** 31 - 1)),
            "steps": int(rng.integers(steps[0], steps[1] + 1))}


def mix(n=64, seed=64, steps=2000):
    """The batched-sweep mix: n crystals of uneven size, 48x48x24 to
    192x192x96 cells, all with the same step count."""
    rng = np.random.default_rng(seed)
    return [random_job(rng, steps=(steps, steps)) for _ in range(n)]


def job_spec(job, probe=None, every=10, sponge=None, steps=None, capacity=None, device_ids=False):
    """MeshSpec of a job dict (as drawn by ``random_job``)."""
    spec, _, _ = crystal_spec(job["cells"], job["a"], job["r"], job["seed"],
                              job["steps"] if steps is None else steps, inclusion=job["inclusion"],
                              probe=probe, every=every, sponge=sponge, capacity=capacity, device_ids=device_ids)
    return spec


def rel_l2(a, b):
    """||a - b|| / ||b|| (0 when both vanish)."""
    d = float(np.linalg.norm((np.asarray(a) - np.asarray(b)).ravel()))
    n = float(np.linalg.norm(np.asarray(b).ravel()))
    return d / n if n > 0 else d
