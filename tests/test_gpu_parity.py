"""GPU parity: libphonofdtd against the reference's golden trajectories and
the CPU oracle, through the C ABI (ctypes).  FP64 mode must reproduce the
reference arithmetic exactly; FP32 modes are held to stated tolerances."""
import numpy as np
import pytest

from oracle import brute
from oracle import reference_math as rm
from helpers import (CASES, golden, oracle_problem, rel_l2, rows_of, spec_from, spec_from_golden,
                     uniform_crystal)
from phonomena_amd import fdtd, problem
from phonomena_amd.simulation import base_solver
from phonomena_amd.simulation.solvers import solver_hip

pytestmark = pytest.mark.gpu
FIELDS = ["ux", "uy", "uz", "ux_old", "uy_old", "uz_old"]


def _run(spec, precision="fp64", kernel="auto", steps=None):
    ctx = fdtd.Context(spec, precision=precision, kernel=kernel)
    ctx.step(spec.wave.size if steps is None else steps)
    ctx.sync()
    return ctx


@pytest.mark.parametrize("name", CASES)
def test_fp64_bitwise_vs_reference_trajectory(name):
    d = golden(name)
    ctx = _run(spec_from_golden(d))
    for f in FIELDS:
        got = ctx.get(f)
        assert np.array_equal(got, d[f]), (f, rel_l2(got, d[f]))
    assert np.array_equal(ctx.probes(), d["uz_surface"])
    ctx.close()


def test_plugin_self_test_matches_reference():
    d = golden("testdefaults_10")
    s = solver_hip.Solver()
    s.cfg["write_mode"] = "off"
    s.test()
    for f in ("ux", "uy", "uz", "ux_old", "uy_old", "uz_old"):
        assert np.array_equal(getattr(s.g, f), d[f]), f
    s.close()


def _oracle_brute(x, y, z, ids, rows, dt, wave, steps, probe=None, every=0, sponge=None):
    fd = [np.diff(a) for a in (x, y, z)]
    sd = [np.mean([f[1:], f[:-1]], axis=0) for f in fd]
    mur = problem.mur_coefficients(rows[int(ids[0, 0, 0])], np.float64(dt), fd, sd)
    return brute.run(x, y, z, ids, rows, dt, mur, wave, steps, probe=probe, every=every, sponge=sponge)


def test_fp64_homogeneous_64x64x32_1000_steps():
    """Plain reference case: 64x64x32 cells, homogeneous, Courant 0.5, sine
    on the x = 0 wall at 1 MHz for dx = 1e-4 m (100 cycles per unit time in
    the reference's unit-cell coordinates), 1000 steps, surface u_z every 10
    steps; FP64 must match the oracle to rel-L2 <= 1e-10."""
    x, y, z, ids, rows, dt, _ = uniform_crystal((64, 64, 32), 1.0, steps=1)
    wave = rm.source_table("sin", 1000, dt, {"f": 1e6 * 1e-4})
    probe = (0, 1, 65, 0, 1, 65)
    st, frames = _oracle_brute(x, y, z, ids, rows, dt, wave, 1000, probe=probe, every=10)
    ctx = _run(spec_from(x, y, z, ids, rows, dt, wave, probe=probe, every=10))
    for f in ("ux", "uy", "uz"):
        assert rel_l2(ctx.get(f), st[f]) <= 1e-10, f
    assert rel_l2(ctx.probes(), frames) <= 1e-10
    ctx.close()


def test_fp64_crystal_two_materials():
    x, y, z, ids, rows, dt, wave = uniform_crystal(
        (40, 36, 20), 1.0, incl=(19300.0, 2.0e11, 1.6e11, 12, 4), steps=300, seed=2)
    assert ids.any()
    st, _ = _oracle_brute(x, y, z, ids, rows, dt, wave, 300)
    ctx = _run(spec_from(x, y, z, ids, rows, dt, wave))
    for f in ("ux", "uy", "uz"):
        assert np.array_equal(ctx.get(f), st[f]), f
    ctx.close()


@pytest.mark.parametrize("spacing", [(0.5, 2.0, 0.25), (1.0, 1.0, 1.0), (0.3, 0.7, 0.3)])
@pytest.mark.parametrize("sponge", [False, True])
def test_fp64_spacings_bitwise_vs_oracle(spacing, sponge):
    """On power-of-two spacings the FP64 tiled sweep takes each division by
    a spacing as a product with its exact reciprocal; on any other spacing it
    divides.  Both must give the oracle's bits (which divides), on every
    tile shape and with or without the absorbing layer."""
    cells = (36, 30, 20)
    _, _, _, ids, rows, dt, wave = uniform_crystal(
        cells, min(spacing), incl=(19300.0, 2.0e11, 1.6e11, 12, 4), steps=160, seed=5)
    x, y, z = (np.arange(n + 1) * h for n, h in zip(cells, spacing))
    prof = problem.sponge_tables((x.size, y.size, z.size), {"width": 4, "strength": 0.2}) if sponge else None
    st, _ = _oracle_brute(x, y, z, ids, rows, dt, wave, 160, sponge=prof)
    for tj, kz in ((8, 32), (16, 64), (32, 32)):
        ctx = fdtd.Context(spec_from(x, y, z, ids, rows, dt, wave, sponge=prof), precision="fp64",
                           kernel="tiled", tile_j=tj, tile_z=kz)
        ctx.step(160)
        for f in ("ux", "uy", "uz"):
            assert np.array_equal(ctx.get(f), st[f]), (tj, kz, f, rel_l2(ctx.get(f), st[f]))
        ctx.close()


@pytest.mark.parametrize("precision,tol", [("fp32", 1e-4), ("fp32c", 1e-5)])
def test_fp32_modes_vs_fp64_oracle(precision, tol):
    x, y, z, ids, rows, dt, wave = uniform_crystal(
        (48, 40, 24), 1.0, incl=(19300.0, 2.0e11, 1.6e11, 16, 5), steps=600, wave_kind="sin", seed=3)
    st, _ = _oracle_brute(x, y, z, ids, rows, dt, wave, 600)
    ctx = _run(spec_from(x, y, z, ids, rows, dt, wave), precision=precision)
    errs = {f: rel_l2(ctx.get(f), st[f]) for f in ("ux", "uy", "uz")}
    print(precision, errs)
    for f, e in errs.items():
        assert e <= tol, (f, e)
    ctx.close()


def test_sponge_fp64_bitwise_vs_oracle():
    d = golden("homog16_200")
    pb = oracle_problem(d)
    prof = problem.sponge_tables(pb.shape, {"width": 4, "strength": 0.2})
    pbs = oracle_problem(d, sponge=prof)
    s, _ = rm.run(pbs, 200)
    ctx = _run(spec_from_golden(d, sponge=prof))
    for c in "xyz":
        assert np.array_equal(ctx.get("u" + c), getattr(s, "u" + c)), c
    ctx.close()


def test_energy_audit_matches_oracle():
    d = golden("default_json_100")
    pb = oracle_problem(d)
    s, _ = rm.run(pb, int(d["steps"]))
    e_ref = rm.energy(s, pb)
    for precision in ("fp64", "fp32", "fp32c"):
        ctx = _run(spec_from_golden(d), precision=precision)
        e = ctx.energy()
        tol = 1e-12 if precision == "fp64" else 1e-3
        assert abs(e - e_ref) <= tol * abs(e_ref), (precision, e, e_ref)
        ctx.close()


def test_set_field_initial_condition_then_step():
    # arbitrary initial fields, zero source: GPU and oracle evolve them alike.
    # The x = 0 wall of uy/uz is never updated by the reference; it holds 0
    # in every run that starts from rest, so it is 0 here too.
    d = golden("homog16_200")
    pb = oracle_problem(d, wave=np.zeros(30))
    rng = np.random.default_rng(7)
    s = rm.State(*pb.shape)
    for c in "xyz":
        a = rng.standard_normal(getattr(s, "u" + c).shape)
        if c != "x":
            a[0] = 0.0
        getattr(s, "u" + c)[...] = a
        getattr(s, "u" + c + "_new")[...] = a
        getattr(s, "u" + c + "_old")[...] = a * 0.5
    spec = spec_from_golden(d, wave=np.zeros(30), probe=None)
    ctx = fdtd.Context(spec, precision="fp64")
    for c in "xyz":
        ctx.set("u" + c, getattr(s, "u" + c))
        ctx.set("u" + c + "_old", getattr(s, "u" + c + "_old"))
        assert np.array_equal(ctx.get("u" + c), getattr(s, "u" + c))
    for tt in range(30):
        rm.step(s, pb, tt)
    ctx.step(30)
    ctx.sync()
    for c in "xyz":
        assert np.array_equal(ctx.get("u" + c), getattr(s, "u" + c)), c
    ctx.close()


def test_probe_ring_keeps_latest_frames():
    d = golden("homog16_200")
    spec = spec_from_golden(d)
    spec.probe_capacity = 5
    ctx = _run(spec)
    fr = ctx.probes()
    assert fr.shape[0] == 5
    assert np.array_equal(fr, d["uz_surface"][-5:])
    ctx.close()


def test_step_beyond_source_table_is_an_error():
    d = golden("testdefaults_10")
    ctx = fdtd.Context(spec_from_golden(d))
    ctx.step(10)
    with pytest.raises(fdtd.FdtdError, match="PF_ERR_RANGE"):
        ctx.step(1)
    ctx.close()


def test_device_info_names_the_gpu():
    info = fdtd.device_info()
    assert info["cu"] > 0
    print(info)
