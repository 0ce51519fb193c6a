"""Solver plugins, discovered by ``phonomena_amd.common.findSolvers``."""
