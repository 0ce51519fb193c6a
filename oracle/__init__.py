"""CPU oracle for the FDTD step: test infrastructure, never product code.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this package.
"""
