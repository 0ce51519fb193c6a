"""Host-side mirror of the reference interface, checked without a GPU:
mesher, material map, dt, source tables, Mur coefficients, plugin discovery,
signals, and the solver's run/cancel logic against a stand-in context."""
import json
import os
import threading

import numpy as np
import pytest

from oracle import reference_math as rm
from helpers import CASES, GOLDEN, golden, oracle_problem, rows_of
from phonomena_amd import common, fdtd, problem
from phonomena_amd.gui.worker import WorkerSignals
from phonomena_amd.simulation import base_solver, grid, material
from phonomena_amd.simulation.solvers import solver_hip


def test_mesher_matches_reference_vectors():
    d = np.load(os.path.join(GOLDEN, "mesh_cases.npz"))
    for n in range(int(d["n"])):
        sx, sy, sz, mdx, mdy, mdz, mind, slope = d["c%d_params" % n]
        g = grid.Grid()
        g.init(size_x=sx, size_y=sy, size_z=sz)
        g.max_dx, g.max_dy, g.max_dz, g.min_d, g.slope = mdx, mdy, mdz, mind, slope
        for (ix, iy, r) in d["c%d_incs" % n]:
            g.addInclusion(ix, iy, r)
        g.buildMesh()
        g.update()
        for a in "xyz":
            assert np.array_equal(getattr(g, a), d["c%d_%s" % (n, a)]), (n, a)
        mask = np.zeros((g.x.size, g.y.size, g.z.size), dtype=np.uint8)
        for yx, zz in g.inclusionIndices():
            for yy, xx in yx:
                mask[xx, yy, zz] = 1
        assert np.array_equal(mask, d["c%d_mask" % n]), n


def _settings_from_fixture(d):
    props = {"M0": {"name": "m0", "p": float(d["p_primary"]), "c_Pa": d["C_primary"].tolist()},
             "M1": {"name": "m1", "p": float(d["p_secondary"]), "c_Pa": d["C_secondary"].tolist()}}
    incl = [{"x": float(t[0]), "y": float(t[1]), "z": float(t[2]), "r": float(t[3])} for t in d["inclusions"]]
    cfg = json.loads(str(d["cfg"]))
    return {"grid": json.loads(str(d["grid_cfg"])), "inclusions": incl,
            "material": {"primary": "M0", "secondary": "M1", "properties": props},
            "simulation": {"courant": float(d["courant"]), "steps": int(d["steps"]),
                           "solver": "hip", "cfg": dict(cfg, write_mode="off")}}


@pytest.mark.parametrize("name", ["default_json_100", "nonuniform_json_60", "fine_json_20"])
def test_load_settings_reproduces_reference_inputs(tmp_path, name):
    d = golden(name)
    path = tmp_path / "s.json"
    path.write_text(json.dumps(_settings_from_fixture(d)))
    common.findSolvers()
    cfg, g, m = common.loadSettings(str(path))
    for a in "xyz":
        assert np.array_equal(getattr(g, a), d[a]), a
    assert np.array_equal(m.ids, d["ids"])
    assert m.dt == d["dt"][()]
    assert common.solver.name == "hip"
    assert common.solver.cfg["wave"] == json.loads(str(d["cfg"]))["wave"]


@pytest.mark.parametrize("name", ["default_json_100", "nonuniform_json_60"])
def test_save_settings_round_trips(tmp_path, name):
    # settings written back (common.py:158-181) load to the same mesh,
    # material map, time step and solver cfg
    d = golden(name)
    first = tmp_path / "a.json"
    first.write_text(json.dumps(_settings_from_fixture(d)))
    common.findSolvers()
    common.loadSettings(str(first))
    common.solver.cfg["precision"] = "fp32c"  # a cfg change the file must carry
    second = tmp_path / "b.json"
    common.saveSettings(str(second))
    saved = json.loads(second.read_text())
    assert saved["simulation"]["solver"] == "hip"
    assert saved["simulation"]["cfg"]["precision"] == "fp32c"
    assert saved["material"]["primary"] == "M0" and saved["material"]["secondary"] == "M1"
    _, g, m = common.loadSettings(str(second))
    for a in "xyz":
        assert np.array_equal(getattr(g, a), d[a]), a
    assert np.array_equal(m.ids, d["ids"])
    assert m.dt == d["dt"][()]
    assert common.solver.cfg["precision"] == "fp32c"


def test_test_defaults_match_reference():
    d = golden("testdefaults_10")
    s = solver_hip.Solver()
    s.cfg["write_mode"] = "off"
    base_solver.BaseSolver.init(s, base_solver.TestDefaults.g, base_solver.TestDefaults.m, 10)
    for a in "xyz":
        assert np.array_equal(getattr(s.g, a), d[a])
    assert s.m.dt == d["dt"][()]
    assert np.array_equal(s.m.ids, d["ids"])
    assert np.array_equal(s.m.C, rm.dense_material(d["ids"], [d["C_primary"], d["C_secondary"]],
                                                   [d["p_primary"], d["p_secondary"]])[0])


@pytest.mark.parametrize("kind,args", [("sin", {"f": 100}), ("ricker", {"f": 100}),
                                       ("ricker", {"f": 250.0, "source_delay": 1e-3}),
                                       ("gauss", {"f": 2e6})])
def test_source_table_matches_oracle(kind, args):
    s = solver_hip.Solver()
    s.m = material.Material()
    s.m.dt = np.float64(2.1135684794138367e-05) if kind != "gauss" else np.float64(4.0477e-9)
    s.cfg["wave"], s.cfg["wave_args"] = kind, args
    a = dict(args)
    f = a.pop("f")
    ref = np.array([rm.source(kind, tt, s.m.dt, f, **a) for tt in range(300)])
    assert np.array_equal(s.source_table(300), ref)


def test_gauss_source_closed_form():
    s = solver_hip.Solver()
    s.m = material.Material()
    s.m.dt = np.float64(1e-8)
    s.cfg["wave"], s.cfg["wave_args"] = "gauss", {"f": 2e6}
    w = s.source_table(800)
    t = np.arange(800) * 1e-8
    sig = 1 / 2e6
    exact = np.sin(2 * np.pi * 2e6 * (t - 4 * sig)) * np.exp(-0.5 * ((t - 4 * sig) / sig) ** 2)
    assert np.allclose(w, exact, rtol=0, atol=1e-12)
    assert abs(w[0]) < 1e-3 and np.argmax(np.abs(w)) in range(150, 250)


@pytest.mark.parametrize("name", CASES)
def test_mur_coefficients_match_oracle(name):
    d = golden(name)
    pb = oracle_problem(d)
    mc = pb.mur_c
    ref = np.array(list(mc["xmax"]) + list(mc["y0"]) + list(mc["ymax"]) + list(mc["zmax"]))
    fd = [np.diff(d[a]) for a in "xyz"]
    sd = [np.mean([f[1:], f[:-1]], axis=0) for f in fd]
    got = problem.mur_coefficients(rows_of(d)[int(d["ids"][0, 0, 0])], d["dt"][()], fd, sd)
    assert np.array_equal(got, ref)


def test_sponge_tables_match_oracle_profile():
    tabs = problem.sponge_tables((30, 20, 10), {"width": 6, "strength": 0.2})
    for t, n, (lo, hi) in zip(tabs, (30, 20, 10), ((False, True), (True, True), (False, True))):
        assert np.array_equal(t, rm.sponge_profile(n, 6, 0.2, lo, hi))


def test_probe_grid_forms():
    assert problem.probe_grid(None, (65, 65, 33)) is None
    assert problem.probe_grid([0, 4, 16, 0, 4, 16], (65, 65, 33)) == (0, 4, 16, 0, 4, 16)
    assert problem.probe_grid({"stride": 4}, (257, 257, 129)) == (0, 4, 65, 0, 4, 65)
    p = problem.probe_grid({"count": 64}, (257, 257, 129))
    assert p[2] == 64 and p[5] == 64 and p[1] * 63 <= 256


def test_find_solvers_registers_hip_plugin():
    d = common.findSolvers()
    assert "hip" in d
    assert isinstance(d["hip"], solver_hip.Solver)


def test_worker_signals_without_qt():
    s = WorkerSignals()
    seen = []
    s.progress.connect(seen.append)
    s.status.connect(seen.append)
    s.progress.emit(5)
    s.status.emit("x")
    assert seen == [5, "x"]


def test_base_solver_has_no_host_stencil():
    b = base_solver.BaseSolver(base_solver.logger)
    for hook in ("update_T", "apply_T_tfbc", "update_u", "apply_u_tfbc", "apply_u_abc", "time_step"):
        with pytest.raises(NotImplementedError):
            getattr(b, hook)()


def test_lazy_reference_arrays_have_reference_shapes():
    g = grid.Grid()
    g.init(5, 6, 7)
    assert g.ux.shape == (5, 7, 8) and g.uy.shape == (6, 6, 8) and g.uz.shape == (6, 7, 7)
    assert g.ux_old.shape == g.ux.shape and g.T4.shape == (6, 6, 7) and g.T6.shape == (5, 6, 8)


def test_void_crystal_gets_pairing_courant():
    from phonomena_amd.simulation import lattice
    g, m = lattice.crystal((24, 24, 8), 8, 3, seed=1, inclusion=lattice.VOID)
    assert m.c_max < 0.02
    g2, m2 = lattice.crystal((24, 24, 8), 8, 3, seed=1, inclusion=lattice.TUNGSTEN)
    assert m2.c_max == 0.5
    assert 0.1 < lattice.fill_fraction(m2) < 0.6


def test_uniform_spacing_extension():
    g = grid.Grid()
    g.setSpacing(1e-4)
    g.init(64, 64, 32)
    assert g.x.size == 65 and g.z.size == 33
    assert np.allclose(np.diff(g.x), 1e-4, rtol=1e-12, atol=0)
    g.addInclusion(32e-4, 32e-4, 5e-4)
    m = material.Material()
    m.init(g, {"A": material.isotropic("a", 2700, 5.1e10, 2.6e10),
               "B": material.isotropic("b", 19300, 2.0e11, 1.6e11)})
    m.setSecondary("B")
    m.update()
    # centre node inside, corner outside; cylinder runs the full depth
    assert m.ids[32, 32].all() and not m.ids[0, 0].any()
    vp = np.sqrt((5.1e10 + 2 * 2.6e10) / 2700)
    vpb = np.sqrt((2.0e11 + 2 * 1.6e11) / 19300)
    assert m.dt == pytest.approx(0.5 * 1e-4 / max(vp, vpb), rel=1e-12)


class _FakeContext:
    """Stand-in for fdtd.Context: counts steps, no GPU."""

    def __init__(self, spec, precision="fp64", kernel="auto"):
        self.spec = spec
        self.n = 0
        self.closed = False

    @property
    def steps_done(self):
        return self.n

    def step(self, n):
        self.n += n

    def sync(self):
        pass

    def get(self, which):
        nx, ny, nz = self.spec.nodes
        shp = {"ux": (nx - 1, ny, nz), "uy": (nx, ny - 1, nz), "uz": (nx, ny, nz - 1)}
        return np.full(shp[which.replace("_old", "")], float(self.n))

    def probes(self):
        return np.zeros((0, 0, 0))

    def close(self):
        self.closed = True


def test_solver_run_chunks_progress_and_results(monkeypatch):
    monkeypatch.setattr(fdtd, "Context", _FakeContext)
    s = solver_hip.Solver()
    s.cfg["write_mode"] = "off"
    s.cfg["chunk"] = 3
    s.init(base_solver.TestDefaults.g, base_solver.TestDefaults.m, 10)
    sig = WorkerSignals()
    prog = []
    sig.progress.connect(prog.append)
    s.run(signals=sig)
    assert s.ctx.n == 10
    assert prog[0] == 0 and prog[-1] == 100 and prog == sorted(prog)
    assert np.all(s.g.ux == 10.0) and np.all(s.g.uz_old == 10.0)


def test_solver_cancel_stops_between_chunks(monkeypatch):
    monkeypatch.setattr(fdtd, "Context", _FakeContext)
    s = solver_hip.Solver()
    s.cfg["write_mode"] = "off"
    s.cfg["chunk"] = 5
    s.init(base_solver.TestDefaults.g, base_solver.TestDefaults.m, 1000)
    sig = WorkerSignals()

    def cancel_at(p):
        if p >= 10:
            s.cancel()
    sig.progress.connect(cancel_at)
    s.run(signals=sig)
    assert 0 < s.ctx.n < 1000
    assert not s.running.is_set()


def test_write_mode_without_h5py_raises_clearly(monkeypatch):
    monkeypatch.setattr(fdtd, "Context", _FakeContext)
    s = solver_hip.Solver()
    s.cfg["write_mode"] = "thread"
    try:
        import h5py  # noqa: F401
        pytest.skip("h5py present")
    except ImportError:
        pass
    with pytest.raises(RuntimeError, match="h5py"):
        s.init(base_solver.TestDefaults.g, base_solver.TestDefaults.m, 10)


def test_cancel_from_another_thread(monkeypatch):
    monkeypatch.setattr(fdtd, "Context", _FakeContext)
    s = solver_hip.Solver()
    s.cfg["write_mode"] = "off"
    s.cfg["chunk"] = 1
    s.init(base_solver.TestDefaults.g, base_solver.TestDefaults.m, 10 ** 6)
    started = threading.Event()
    orig = _FakeContext.step

    def slow_step(self, n):
        started.set()
        orig(self, n)
    monkeypatch.setattr(_FakeContext, "step", slow_step)
    t = threading.Thread(target=s.run)
    t.start()
    started.wait(5)
    s.cancel()
    t.join(10)
    assert not t.is_alive()
    assert s.ctx.n < 10 ** 6


class _DescLib:
    """Stand-in for the library that records the descriptor pf_create gets."""

    def __init__(self):
        self.desc = None

    def pf_create(self, dptr, hptr):
        d = dptr._obj
        self.desc = (d.kernel, d.tile_j, d.tile_z, d.tile_x)
        hptr._obj.value = 1
        return 0

    def pf_destroy(self, h):
        return 0


@pytest.mark.parametrize("precision,tiles", [("fp32", (16, 64)), ("fp32c", (16, 64)), ("fp64", (8, 32))])
def test_context_tune_flag_sets_the_descriptor_tiles(monkeypatch, precision, tiles):
    # tune=False asks for the fixed shape (no measurement at creation); an
    # explicit shape is passed as given; the default leaves it to pf_create
    from phonomena_amd.fdtd import MeshSpec
    lib = _DescLib()
    monkeypatch.setattr(fdtd, "load", lambda path=None: lib)
    spec = MeshSpec((6, 6, 6), [np.ones(5)] * 3, [np.ones(4)] * 3, np.array([[1.0] * 13]),
                    np.zeros((6, 6, 6), np.uint8), 1.0, 1.0, np.zeros(12), np.zeros(4))
    fdtd.Context(spec, precision=precision).close()
    assert lib.desc == (fdtd.KERNEL["auto"], 0, 0, 0)
    fdtd.Context(spec, precision=precision, tune=False).close()
    assert lib.desc == (fdtd.KERNEL["auto"],) + tiles + (0,)
    fdtd.Context(spec, precision=precision, tile_j=32, tile_z=32, tune=False).close()
    assert lib.desc == (fdtd.KERNEL["auto"], 32, 32, 0)
    # the strip and naive sweeps keep their own defaults (no tiled shape
    # forced on them, which the strip sweep would reject)
    for k in ("strip", "naive"):
        fdtd.Context(spec, precision=precision, kernel=k, tune=False).close()
        assert lib.desc == (fdtd.KERNEL[k], 0, 0, 0)


# ------------------------------------------- the reference's own objects --

def _spec_fields(spec):
    return (spec.nodes, spec.materials, spec.ids, spec.dt, spec.dt2, spec.mur, spec.wave)


def _same_spec(a, b):
    for u, v in zip(_spec_fields(a), _spec_fields(b)):
        assert np.array_equal(np.asarray(u), np.asarray(v))


def test_reference_style_material_drives_init_and_run(monkeypatch):
    # a Material with only the reference's attributes (dense C/P, no rows()
    # or ids), as common.loadSettings hands it to a plugin
    from helpers import ReferenceStyleMaterial
    monkeypatch.setattr(fdtd, "Context", _FakeContext)
    ref_m = ReferenceStyleMaterial()
    ref_m.init(base_solver.TestDefaults.g, base_solver.TestDefaults.properties)
    assert not hasattr(ref_m, "rows") and not hasattr(ref_m, "ids")
    s = solver_hip.Solver()
    s.cfg["write_mode"] = "off"
    s.init(base_solver.TestDefaults.g, ref_m, 10)
    s.run()
    assert s.ctx.n == 10
    own = solver_hip.Solver()
    own.cfg["write_mode"] = "off"
    own.init(base_solver.TestDefaults.g, base_solver.TestDefaults.m, 10)
    _same_spec(s.ctx.spec, own.ctx.spec)


@pytest.mark.parametrize("name", ["default_json_100", "nonuniform_json_60", "nonuniform_deep_40"])
def test_reference_style_material_gives_fixture_tables(name, monkeypatch):
    # two materials with inclusions (partial depth in nonuniform_deep_40):
    # the rows are primary then secondary, the map is the fixture's
    from helpers import reference_style_material
    monkeypatch.setattr(fdtd, "Context", _FakeContext)
    d = golden(name)
    settings = _settings_from_fixture(d)
    g = grid.Grid()
    g.init(**{k: settings["grid"][k] for k in ("size_x", "size_y", "size_z")})
    for k in ("min_d", "max_dx", "max_dy", "max_dz", "slope"):
        setattr(g, k, settings["grid"][k])
    for t in settings["inclusions"]:
        g.addInclusion(x=t["x"], y=t["y"], z=t["z"], r=t["r"])
    g.buildMesh()
    g.update()
    m = reference_style_material(g, d)
    m.c_max = float(d["courant"])
    m.update()
    rows, ids = problem.material_table(m, (g.x.size, g.y.size, g.z.size))
    assert np.array_equal(rows, rows_of(d))
    assert np.array_equal(ids, d["ids"])
    assert m.dt == d["dt"][()]
    s = solver_hip.Solver()
    s.cfg.update(write_mode="off", wave=json.loads(str(d["cfg"]))["wave"])
    s.init(g, m, 5)
    assert np.array_equal(s.ctx.spec.materials, rows_of(d)) and np.array_equal(s.ctx.spec.ids, d["ids"])


def test_material_table_takes_any_per_node_assignment():
    # more than two materials, assigned per node: rows are the distinct
    # per-node rows, and rows[ids] rebuilds every node's entries bit for bit
    from helpers import ReferenceStyleMaterial
    rng = np.random.default_rng(3)
    m = ReferenceStyleMaterial()
    nodes = (6, 5, 7)
    base = [rng.uniform(1, 2, (6, 6)) * 1e10 for _ in range(4)]
    rho = [2700.0, 19300.0, 1.2, 5307.0]
    pick = rng.integers(0, 4, nodes)
    pick[0, 0, 0] = 3  # the secondary first in memory: still row 1
    pick[1, 1, 1] = 0
    m.C = np.stack(base)[pick]
    m.P = np.array(rho)[pick]
    m.primary = {"c": base[0], "p": rho[0]}
    m.secondary = {"c": base[3], "p": rho[3]}
    rows, ids = problem.material_table(m, nodes)
    assert rows.shape == (4, 13) and ids.dtype == np.uint8
    assert np.array_equal(rows[ids], problem.dense_rows(m.C, m.P))
    entries = lambda c: [c[a, b] for a, b in problem._ROW_ENTRIES]
    assert np.array_equal(rows[0][:12], entries(base[0])) and rows[0][12] == rho[0]  # primary
    assert np.array_equal(rows[1][:12], entries(base[3])) and rows[1][12] == rho[3]  # secondary
    assert ids[0, 0, 0] == 1 and ids[1, 1, 1] == 0


def test_material_table_rejects_mismatched_arrays():
    from helpers import ReferenceStyleMaterial
    m = ReferenceStyleMaterial()
    m.C = np.zeros((4, 4, 4, 6, 6))
    m.P = np.ones((4, 4, 4))
    with pytest.raises(ValueError, match="nodes"):
        problem.material_table(m, (5, 4, 4))


def test_power_of_two_reciprocal_products_equal_divisions():
    # The FP64 sweep on power-of-two spacings takes x / h as x * (1/h): for
    # h = 2^k both are the correctly rounded value of the same real number,
    # including subnormal, overflowing and signed-zero results; for other
    # spacings the two differ on some operands (a few special divisors, such
    # as 0.3, happen to agree on every operand tried: only powers of two are
    # used).
    rng = np.random.default_rng(11)
    x = np.concatenate([rng.standard_normal(20000) * 10.0 ** rng.integers(-300, 300, 20000),
                        rng.standard_normal(2000) * 1e-310,          # subnormal
                        np.array([0.0, -0.0, 5e-324, -5e-324, 1.7e308, -1.7e308, np.inf, -np.inf])])
    with np.errstate(over="ignore", under="ignore"):
        for k in range(-40, 41):
            h = np.float64(2.0) ** k
            r = np.float64(1.0) / h
            assert (x * r).view(np.int64).tolist() == (x / h).view(np.int64).tolist(), k
        # the reference's physical spacing and another non-power of two
        for h in (np.float64(1e-4), np.float64(0.7)):
            assert not np.array_equal((x * (1.0 / h)).view(np.int64), (x / h).view(np.int64)), h
