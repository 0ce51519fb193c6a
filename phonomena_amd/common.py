"""Solver discovery and settings I/O (phonomena/common.py:88-181).

``findSolvers`` imports every module in ``simulation/solvers`` and registers
its ``Solver()`` by ``name``; ``importSolver`` builds one by module name;
``loadSettings``/``saveSettings`` map the JSON settings schema onto Grid,
Material and the selected solver's cfg, as the reference does.  Logging,
temp-dir and multiprocessing start-up plumbing of the reference are not part
of this package.
"""
import importlib
import json
import logging
import os
from pathlib import Path

from .simulation import grid as _grid
from .simulation import material as _material

logger = logging.getLogger(__name__)

SOLVER_PKG = "phonomena_amd.simulation.solvers"
SOLVER_DIR = Path(__file__).parent.joinpath("simulation", "solvers")

cfg = None
grid = None
material = None
solver_dict = {}
solver = None


def importSolver(module):
    mod = importlib.import_module("{}.{}".format(SOLVER_PKG, module))
    return getattr(mod, "Solver")()


def findSolvers():
    global solver_dict, solver
    solver_dict = {}
    for f in sorted(os.listdir(SOLVER_DIR)):
        if not f.endswith(".py") or f.startswith("__"):
            continue
        name = f[:-3]
        try:
            s = importSolver(name)
            solver_dict[s.name] = s
        except Exception as e:  # a broken plugin is logged and skipped
            logger.error("Error importing solver {}: {}".format(name, e))
    logger.info("Found solvers: {}".format(list(solver_dict.keys())))
    solver = next(iter(solver_dict.values())) if solver_dict else None
    return solver_dict


def loadSettings(path):
    """Grid, Material and solver cfg from a settings JSON (common.py:112-156)."""
    global cfg, grid, material, solver
    with open(path) as fh:
        cfg = json.load(fh)
    grid = _grid.Grid()
    grid.init(size_x=cfg['grid']['size_x'], size_y=cfg['grid']['size_y'], size_z=cfg['grid']['size_z'])
    grid.min_d = cfg['grid']['min_d']
    grid.max_dx = cfg['grid']['max_dx']
    grid.max_dy = cfg['grid']['max_dy']
    grid.max_dz = cfg['grid']['max_dz']
    grid.slope = cfg['grid']['slope']
    for t in cfg["inclusions"]:
        grid.addInclusion(x=t['x'], y=t['y'], z=t['z'], r=t['r'])
    grid.buildMesh()
    grid.update()
    material = _material.Material()
    material.init(grid=grid, properties=cfg['material']['properties'])
    material.c_max = cfg['simulation']['courant']
    material.setPrimary(cfg['material']['primary'])
    material.setSecondary(cfg['material']['secondary'])
    material.update()
    s = cfg['simulation']['solver']
    if s in solver_dict:
        solver = solver_dict[s]
        solver.cfg = {**solver.cfg, **cfg['simulation']['cfg']}
    else:
        logger.error("Solver not found: {}".format(s))
    return cfg, grid, material


def saveSettings(path):
    """Write the current grid/material/solver back to JSON (common.py:158-181)."""
    cfg['grid'] = {"slope": grid.slope, "max_dx": grid.max_dx, "max_dy": grid.max_dy,
                   "max_dz": grid.max_dz, "min_d": grid.min_d, "size_x": grid.size_x,
                   "size_y": grid.size_y, "size_z": grid.size_z}
    cfg['inclusions'] = [{'x': float(r[0]), 'y': float(r[1]), 'z': float(r[2]), 'r': float(r[3])}
                         for r in grid.targets]
    cfg['simulation']['courant'] = material.c_max
    cfg['material']['primary'] = material.primary_key
    cfg['material']['secondary'] = material.secondary_key
    cfg['simulation']['solver'] = solver.name
    cfg['simulation']['cfg'] = solver.cfg
    with open(path, 'w') as fh:
        json.dump(cfg, fh)
