"""Write the golden trajectory fixtures by running the reference solver itself.

This script is run once, by hand, in the development container (the one that
holds the reference tree); its outputs are the ``*.npz`` files next to it.
It is NOT run by the test suite and nothing under ``tests/`` imports it.

Why an import is needed: the reference's own tests hold no expected outputs
(``tests/test_solvers.py:18-29`` only smoke-runs ``BaseSolver.test``), so the
complete ``BaseSolver.run`` trajectory (``phonomena/simulation/base_solver.py:
224-280``) and the non-uniform mesher (``phonomena/simulation/grid.py:177-303``)
are pinned bit-exactly only by running them.

The reference is imported unmodified.  Three things it imports are absent or
changed in this image and are bridged from the outside:

* ``gui.worker`` imports PyQt5 (``phonomena/gui/worker.py:1-2``); a module
  object exposing a no-op ``WorkerSignals`` is registered in ``sys.modules``.
* ``h5py`` (``base_solver.py:3``) is only used when ``write_mode != 'off'``;
  every run here uses ``write_mode='off'``, so an empty module is registered.
* ``np.float`` (``grid.py:258,301``) was removed in NumPy 1.24; it is aliased
  to the builtin ``float`` (what it always meant).

Per-step frames are taken by wrapping the instance's ``time_step`` hook, so
frame ``t`` is the displacement after the rotation of step ``t`` — the same
frame the reference's HDF writer stores (``base_solver.py:256-260``).

Usage:  python tests/golden/make_golden.py [REFERENCE_ROOT] [--only nonuniform_deep_40]
"""
import json
import os
import sys
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = sys.argv[1] if len(sys.argv) > 1 and not sys.argv[1].startswith("--") else "/root/reference"
# --only NAME writes that one fixture (the others are left as they are)
ONLY = sys.argv[sys.argv.index("--only") + 1] if "--only" in sys.argv else None


def _install_bridges():
    class _Sig:
        def emit(self, *a, **k):
            pass

        def connect(self, *a, **k):
            pass

    class WorkerSignals:
        def __init__(self):
            for name in ("finished", "success", "error", "warning", "result",
                         "progress", "status", "quit"):
                setattr(self, name, _Sig())

    gui = types.ModuleType("gui")
    gui.__path__ = []
    worker = types.ModuleType("gui.worker")
    worker.WorkerSignals = WorkerSignals
    gui.worker = worker
    sys.modules["gui"] = gui
    sys.modules["gui.worker"] = worker
    sys.modules["h5py"] = types.ModuleType("h5py")
    if not hasattr(np, "float"):
        np.float = float


def _frames_solver(solver_cls, record_every, full_every):
    """Instantiate a reference solver whose time_step hook records frames."""
    s = solver_cls()
    s.cfg["write_mode"] = "off"
    rec = {"uz_surface": [], "full": [], "full_steps": [], "step": 0}
    orig = s.time_step

    def time_step():
        orig()
        t = rec["step"]
        if (t + 1) % record_every == 0:
            rec["uz_surface"].append(np.copy(s.g.uz[:, :, 0]))
        if full_every and (t + 1) % full_every == 0:
            rec["full"].append((np.copy(s.g.ux), np.copy(s.g.uy), np.copy(s.g.uz)))
            rec["full_steps"].append(t)
        rec["step"] = t + 1

    s.time_step = time_step
    return s, rec


def _save(name, s, rec, extra):
    g, m = s.g, s.m
    P = m.P
    ids = np.zeros(P.shape, dtype=np.uint8)
    ids[P == float(m.secondary["p"])] = 1
    if m.secondary is m.primary:
        ids[:] = 0
    out = dict(
        x=g.x, y=g.y, z=g.z, dt=np.float64(m.dt), steps=np.int64(s.t),
        ids=ids,
        C_primary=np.array(m.primary["c"], dtype=np.float64),
        C_secondary=np.array(m.secondary["c"], dtype=np.float64),
        p_primary=np.float64(m.primary["p"]), p_secondary=np.float64(m.secondary["p"]),
        ux=g.ux, uy=g.uy, uz=g.uz,
        ux_old=g.ux_old, uy_old=g.uy_old, uz_old=g.uz_old,
        cfg=np.array(json.dumps({k: v for k, v in s.cfg.items() if k != "write_mode"})),
        uz_surface=np.array(rec["uz_surface"]),
    )
    if rec["full"]:
        out["full_steps"] = np.array(rec["full_steps"], dtype=np.int64)
        out["full_ux"] = np.array([f[0] for f in rec["full"]])
        out["full_uy"] = np.array([f[1] for f in rec["full"]])
        out["full_uz"] = np.array([f[2] for f in rec["full"]])
    # The dense reference material arrays are re-derivable from ids + tables;
    # check that here so a fixture never disagrees with its own inputs.
    C = np.where(ids[..., None, None] == 1, out["C_secondary"], out["C_primary"])
    assert np.array_equal(C, m.C), name
    assert np.array_equal(np.where(ids == 1, out["p_secondary"], out["p_primary"]), P), name
    out.update(extra)
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


def _random_start(g, seed):
    """Seeded non-zero start on every field the reference updates, so the
    run exercises every cell (faces, edges, free surface) from step 1.  The
    reference invariants hold: u_new == u at a step's start (time_step
    copies, base_solver.py:556-571), and the x = 0 wall of uy/uz, which the
    reference never updates (apply_u_abc has no x = 0 face,
    base_solver.py:519-554), holds 0.  Returns the start, for the fixture."""
    rng = np.random.default_rng(seed)
    start = {}
    for c in ("x", "y", "z"):
        a = rng.standard_normal(getattr(g, "u" + c).shape)
        o = a * 0.75 + 0.25 * rng.standard_normal(a.shape)
        if c != "x":
            a[0] = 0.0
            o[0] = 0.0
        getattr(g, "u" + c)[...] = a
        getattr(g, "u" + c + "_new")[...] = a
        getattr(g, "u" + c + "_old")[...] = o
        start["start_u" + c] = a.copy()
        start["start_u" + c + "_old"] = o.copy()
    return start


def case_nonuniform_deep(common, solver_default, steps=40, seed=11):
    """tests/data/nonuniform.json (graded x/y spacing, 0.3..1.4) with
    size_z = 10 and max_dz = 0.8 (13 z nodes at 10/12), courant 0.25: at
    least 8 nodes on every axis and no unit spacing on any, so the tiled,
    two-step and batched sweeps all run on it.  The inclusions reach
    z = 5, so they cover a leading part of the depth (7 z nodes for the
    first, 6 for the others: the reference's rebinding of z in
    inclusionIndices, grid.py:158-175).  Seeded random start (see
    _random_start), the json's ricker source, 40 steps."""
    import tempfile
    cfg = json.load(open(os.path.join(REF, "tests", "data", "nonuniform.json")))
    cfg["grid"]["size_z"] = 10
    cfg["grid"]["max_dz"] = 0.8
    cfg["simulation"]["courant"] = 0.25
    with tempfile.NamedTemporaryFile("w", suffix=".json", delete=False) as fh:
        json.dump(cfg, fh)
        path = fh.name
    try:
        cfg, g, m = common.loadSettings(path)
    finally:
        os.unlink(path)
    s, rec = _frames_solver(solver_default.Solver, 10, 0)
    s.cfg.update({k: v for k, v in cfg["simulation"]["cfg"].items() if k != "write_mode"})
    s.cfg["write_mode"] = "off"
    s.init(g, m, steps)
    start = _random_start(s.g, seed)
    s.run()
    incl = np.array([[t["x"], t["y"], t["z"], t["r"]] for t in cfg["inclusions"]])
    extra = {"courant": np.float64(cfg["simulation"]["courant"]), "inclusions": incl,
             "grid_cfg": np.array(json.dumps(cfg["grid"]))}
    extra.update(start)
    return s, rec, extra


def main():
    _install_bridges()
    sys.path.insert(0, os.path.join(REF, "phonomena"))
    from simulation import base_solver, grid, material  # noqa: E402
    from simulation.solvers import solver_default  # noqa: E402
    import common  # noqa: E402

    if ONLY == "nonuniform_deep_40":
        s, rec, extra = case_nonuniform_deep(common, solver_default)
        _save("nonuniform_deep_40", s, rec, extra)
        return

    # 1. BaseSolver.test(): TestDefaults, 10 steps, default cfg (ricker, f=100).
    s, rec = _frames_solver(solver_default.Solver, 1, 1)
    s.test()
    _save("testdefaults_10", s, rec, {})

    # 2. Homogeneous GaAs 16x16x8, sin source, courant 0.5, 200 steps.
    props = json.load(open(os.path.join(REF, "data", "default.json")))["material"]["properties"]
    g = grid.Grid()
    g.init(size_x=16, size_y=16, size_z=8)
    m = material.Material()
    m.init(grid=g, properties=props)
    m.setPrimary("GaAs")
    m.setSecondary("GaAs")
    g.buildMesh()
    g.update()
    m.update()
    s, rec = _frames_solver(solver_default.Solver, 10, 50)
    s.cfg["wave"] = "sin"
    s.cfg["wave_args"] = {"f": 250.0}
    s.init(g, m, 200)
    s.run()
    _save("homog16_200", s, rec, {})

    # 3-5. The reference's own JSON settings files, through common.loadSettings.
    for fname, steps, tag in (
        (os.path.join(REF, "data", "default.json"), 100, "default_json_100"),
        (os.path.join(REF, "tests", "data", "nonuniform.json"), 60, "nonuniform_json_60"),
        (os.path.join(REF, "tests", "data", "fine.json"), 20, "fine_json_20"),
    ):
        cfg, g, m = common.loadSettings(fname)
        s, rec = _frames_solver(solver_default.Solver, 10, 0)
        s.cfg.update({k: v for k, v in cfg["simulation"]["cfg"].items() if k != "write_mode"})
        s.cfg["write_mode"] = "off"
        s.init(g, m, steps)
        s.run()
        incl = np.array([[t["x"], t["y"], t["z"], t["r"]] for t in cfg["inclusions"]])
        _save(tag, s, rec, {"courant": np.float64(cfg["simulation"]["courant"]),
                            "inclusions": incl,
                            "grid_cfg": np.array(json.dumps(cfg["grid"]))})

    # 6. Mesher-only vectors: buildMesh on hand-picked non-uniform settings,
    #    including a lone inclusion and off-grid (float32-rounded) centres.
    mesh_cases = []
    for (sx, sy, sz, mdx, mdy, mdz, mind, slope, incs) in (
        (20, 20, 5, 1.0, 1.0, 1.0, 0.8, 1.0, [(15, 15, 1.0), (9.5, 5, 2.0)]),
        (30, 20, 5, 0.5, 0.7, 0.5, 0.25, 0.5, [(7.3, 6.1, 1.7)]),
        (24, 16, 4, 2.0, 2.0, 1.0, 0.5, 2.0, [(5, 5, 1.5), (12, 8, 2.5), (19, 11, 1.2)]),
    ):
        g = grid.Grid()
        g.init(size_x=sx, size_y=sy, size_z=sz)
        g.max_dx, g.max_dy, g.max_dz, g.min_d, g.slope = mdx, mdy, mdz, mind, slope
        for (ix, iy, r) in incs:
            g.addInclusion(ix, iy, r)
        g.buildMesh()
        g.update()
        idx = g.inclusionIndices()
        mask = np.zeros((g.x.size, g.y.size, g.z.size), dtype=np.uint8)
        for yx, zz in idx:
            for yy, xx in yx:
                mask[xx, yy, zz] = 1
        mesh_cases.append(dict(params=np.array([sx, sy, sz, mdx, mdy, mdz, mind, slope], dtype=np.float64),
                               incs=np.array(incs, dtype=np.float64), x=g.x, y=g.y, z=g.z, mask=mask))
    out = {}
    for n, c in enumerate(mesh_cases):
        for k, v in c.items():
            out["c%d_%s" % (n, k)] = v
    out["n"] = np.int64(len(mesh_cases))
    path = os.path.join(HERE, "mesh_cases.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
