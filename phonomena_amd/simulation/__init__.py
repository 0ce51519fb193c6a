"""Solver core: mesh, material, solver base class and solver plugins."""
from . import grid, material  # noqa: F401
