"""phonomena_amd: MI355X (gfx950) FDTD time-stepping solver for phonomena.

The package mirrors the reference's ``simulation`` plugin interface
(``simulation.grid``, ``simulation.material``, ``simulation.base_solver``,
``simulation.solvers.solver_*``, ``common``) and adds the GPU solver plugin
``solver_hip`` backed by the HIP library ``lib/libphonofdtd.so``.
"""
__version__ = "0.1.0"
