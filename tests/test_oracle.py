"""The oracle, pinned: golden trajectories of the reference, two independent
restatements agreeing bit for bit, and closed forms / invariants."""
import json

import numpy as np
import pytest

from oracle import brute
from oracle import reference_math as rm
from helpers import CASES, golden, oracle_problem, record_every, rows_of

FIELDS = ["ux", "uy", "uz", "ux_old", "uy_old", "uz_old"]


@pytest.mark.parametrize("name", CASES)
def test_numpy_oracle_matches_reference_trajectory(name):
    d = golden(name)
    pb = oracle_problem(d)
    x, y = d["x"], d["y"]
    s, frames = rm.run(pb, int(d["steps"]), probe=(0, 1, x.size, 0, 1, y.size),
                       record_every=record_every(name))
    for f in FIELDS:
        assert np.array_equal(getattr(s, f), d[f]), f
    assert np.array_equal(frames, d["uz_surface"])
    if "full_steps" in d.files:
        # intermediate frames too (re-run to the recorded steps)
        for n, t in enumerate(d["full_steps"]):
            s2, _ = rm.run(pb, int(t) + 1)
            for c in "xyz":
                assert np.array_equal(getattr(s2, "u" + c), d["full_u" + c][n])


@pytest.mark.parametrize("omp", [False, True])
@pytest.mark.parametrize("name", CASES)
def test_scalar_oracle_matches_reference_trajectory(name, omp):
    # the serial and the OpenMP build of the C restatement, bit for bit
    d = golden(name)
    pb = oracle_problem(d)
    mc = pb.mur_c
    mur12 = np.array(list(mc["xmax"]) + list(mc["y0"]) + list(mc["ymax"]) + list(mc["zmax"]))
    x, y = d["x"], d["y"]
    st, frames = brute.run(x, y, d["z"], d["ids"], rows_of(d), pb.dt, mur12, pb.wave, int(d["steps"]),
                           probe=(0, 1, x.size, 0, 1, y.size), every=record_every(name), omp=omp)
    for f in FIELDS:
        key = f.replace("_old", "_old")
        assert np.array_equal(st[key], d[f]), f
    assert np.array_equal(frames, d["uz_surface"])


@pytest.mark.parametrize("name", CASES + ["nonuniform_deep_40"])
def test_time_step_size_equals_reference_dt(name):
    # material.py:80-93 over both materials, spacings in SI (SI = 1), the
    # Courant number of the run (0.5 by default, material.py:30)
    d = golden(name)
    c_max = float(d["courant"]) if "courant" in d.files else 0.5
    fd = [np.diff(d[a]) for a in "xyz"]
    mats = [(d["C_primary"], d["p_primary"]), (d["C_secondary"], d["p_secondary"])]
    assert rm.time_step_size(mats, fd, c_max) == d["dt"][()]
    # the spacing that bounds it is the smallest one over all axes
    if name in ("nonuniform_json_60", "nonuniform_deep_40"):
        assert rm.time_step_size(mats, [np.array([np.max(f)]) for f in fd], c_max) != d["dt"][()]


def _start_state(d):
    """Oracle state at the fixture's seeded start (u_new = u)."""
    s = rm.State(d["x"].size, d["y"].size, d["z"].size)
    for c in "xyz":
        getattr(s, "u" + c)[...] = d["start_u" + c]
        getattr(s, "u" + c + "_new")[...] = d["start_u" + c]
        getattr(s, "u" + c + "_old")[...] = d["start_u" + c + "_old"]
    return s


def test_oracles_match_nonuniform_deep_fixture():
    # graded x/y spacing, dz = 10/12, >= 8 nodes per axis, partial-depth
    # inclusions, seeded start: both restatements reproduce the reference
    d = golden("nonuniform_deep_40")
    pb = oracle_problem(d)
    steps = int(d["steps"])
    x, y = d["x"], d["y"]
    s, frames = rm.run(pb, steps, probe=(0, 1, x.size, 0, 1, y.size), record_every=10, state=_start_state(d))
    for f in FIELDS:
        assert np.array_equal(getattr(s, f), d[f]), f
    assert np.array_equal(frames, d["uz_surface"])
    mc = pb.mur_c
    mur12 = np.array(list(mc["xmax"]) + list(mc["y0"]) + list(mc["ymax"]) + list(mc["zmax"]))
    st0 = _start_state(d)
    state = {"u" + c + suf: np.ascontiguousarray(getattr(st0, "u" + c + suf))
             for c in "xyz" for suf in ("", "_new", "_old")}
    st, fr = brute.run(x, y, d["z"], d["ids"], rows_of(d), pb.dt, mur12, pb.wave, steps, state=state,
                       probe=(0, 1, x.size, 0, 1, y.size), every=10)
    for f in FIELDS:
        assert np.array_equal(st[f], d[f]), f
    assert np.array_equal(fr, d["uz_surface"])


def _small_problem(steps=40, sponge=None, scale=1.0, wave_kind="ricker"):
    d = golden("default_json_100")
    cfg = json.loads(str(d["cfg"]))
    dt = d["dt"][()]
    w = rm.source_table(wave_kind, steps, dt, {"f": 5000.0}) * scale
    return d, oracle_problem(d, wave=w, sponge=sponge)


def test_zero_source_keeps_fields_zero():
    d, pb = _small_problem(scale=0.0)
    s, _ = rm.run(pb, 40)
    for c in "xyz":
        assert not np.any(getattr(s, "u" + c))


def test_power_of_two_scaling_is_exact():
    # every operation is linear in u; scaling by 2 commutes exactly with IEEE
    # add/sub/mul/div, so doubling the source doubles every field bit for bit
    _, pb1 = _small_problem(scale=1.0)
    _, pb2 = _small_problem(scale=2.0)
    s1, _ = rm.run(pb1, 40)
    s2, _ = rm.run(pb2, 40)
    for c in "xyz":
        assert np.array_equal(2 * getattr(s1, "u" + c), getattr(s2, "u" + c))


def test_linearity_superposition():
    _, pa = _small_problem(wave_kind="ricker")
    _, pb = _small_problem(wave_kind="sin")
    d, pab = _small_problem()
    pab.wave = pa.wave + pb.wave
    sa, _ = rm.run(pa, 40)
    sb, _ = rm.run(pb, 40)
    sab, _ = rm.run(pab, 40)
    for c in "xyz":
        ref = getattr(sab, "u" + c)
        assert np.allclose(getattr(sa, "u" + c) + getattr(sb, "u" + c), ref,
                           rtol=0, atol=1e-12 * np.abs(ref).max())


def test_y_mirror_symmetry_homogeneous():
    # uniform source line and a y-symmetric homogeneous box: ux, uz even in y,
    # uy odd (at its staggered positions)
    d = golden("homog16_200")
    pb = oracle_problem(d)
    s, _ = rm.run(pb, 120)
    for c, sign, shift in (("x", 1, 0), ("z", 1, 0), ("y", -1, 0)):
        u = getattr(s, "u" + c)
        m = sign * u[:, ::-1, :]
        scale = np.abs(u).max()
        assert np.allclose(u, m, rtol=0, atol=1e-12 * scale), c


def _plane_wave_problem(n=(48, 40, 40), polar="x"):
    """Homogeneous uniform grid, zero source, exact discrete plane wave along x."""
    h = 1e-3
    nx, ny, nz = n
    x, y, z = (np.arange(k) * h for k in n)
    rho, lam, mu = 2700.0, 5.1e10, 2.6e10
    c = np.zeros((6, 6))
    c[:3, :3] = lam
    for a in range(3):
        c[a, a] = lam + 2 * mu
        c[3 + a, 3 + a] = mu
    ids = np.zeros(n, dtype=np.uint8)
    C, P = rm.dense_material(ids, [c], [rho])
    v = np.sqrt((lam + 2 * mu) / rho) if polar == "x" else np.sqrt(mu / rho)
    dt = 0.4 * h / np.sqrt((lam + 2 * mu) / rho)
    steps = 12
    pb = rm.Problem(x, y, z, C, P, dt, np.zeros(steps))
    k = 2 * np.pi / (12 * h)
    nu = v * dt / h
    omega = 2 / dt * np.arcsin(nu * np.sin(k * h / 2))  # discrete dispersion
    return pb, k, omega, h, dt, steps


@pytest.mark.parametrize("polar", ["x", "z"])
def test_plane_wave_follows_discrete_dispersion(polar):
    """P wave (ux) or S wave (uz) along x obeys sin(w dt/2) = (v dt/h) sin(k h/2)
    in the interior, away from every boundary, until boundary effects arrive."""
    pb, k, omega, h, dt, steps = _plane_wave_problem(polar=polar)
    s = rm.State(*pb.shape)
    nx = pb.shape[0]
    if polar == "x":
        xs = (np.arange(nx - 1) + 0.5) * h  # ux lives at i + 1/2
        s.ux[...] = np.cos(k * xs)[:, None, None]
        s.ux_old[...] = np.cos(k * xs + omega * dt)[:, None, None]
        s.ux_new[...] = s.ux
    else:
        xs = np.arange(nx) * h  # uz lives at integer x
        s.uz[...] = np.cos(k * xs)[:, None, None]
        s.uz_old[...] = np.cos(k * xs + omega * dt)[:, None, None]
        s.uz_new[...] = s.uz
    for tt in range(steps):
        rm.step(s, pb, tt)
    exact = np.cos(k * xs - omega * steps * dt)
    c = pb.shape[1] // 2, pb.shape[2] // 2
    sl = slice(steps + 2, nx - steps - 3)
    got = (s.ux if polar == "x" else s.uz)[sl, c[0], c[1]]
    assert np.max(np.abs(got - exact[sl])) < 1e-11
    # and a wrong dispersion (continuum omega) is clearly rejected
    v = (np.sqrt(pb.C[0, 0, 0, 0, 0] / pb.P[0, 0, 0]) if polar == "x"
         else np.sqrt(pb.C[0, 0, 0, 5, 5] / pb.P[0, 0, 0]))
    wrong = np.cos(k * xs - v * k * steps * dt)
    assert np.max(np.abs(got - wrong[sl])) > 1e-6


def test_sponge_reduces_to_reference_when_zero():
    # s = 0 everywhere: the damped form ((2u-(1-0)uo)+c*d)/(1+0) is the
    # reference expression bit for bit
    d = golden("default_json_100")
    pb0 = oracle_problem(d)
    zeros = [np.zeros(pb0.shape[0]), np.zeros(pb0.shape[1]), np.zeros(pb0.shape[2])]
    pbs = oracle_problem(d, sponge=zeros)
    s0, _ = rm.run(pb0, 60)
    s1, _ = rm.run(pbs, 60)
    for c in "xyz":
        assert np.array_equal(getattr(s0, "u" + c), getattr(s1, "u" + c))


def test_sponge_numpy_and_scalar_agree():
    d = golden("homog16_200")
    pb = oracle_problem(d)
    prof = [rm.sponge_profile(n, 4, 0.2, lo, hi) for n, (lo, hi) in
            zip(pb.shape, ((False, True), (True, True), (False, True)))]
    pbs = oracle_problem(d, sponge=prof)
    s, _ = rm.run(pbs, 80)
    mc = pbs.mur_c
    mur12 = np.array(list(mc["xmax"]) + list(mc["y0"]) + list(mc["ymax"]) + list(mc["zmax"]))
    st, _ = brute.run(d["x"], d["y"], d["z"], d["ids"], rows_of(d), pbs.dt, mur12, pbs.wave, 80,
                      sponge=prof)
    for c in "xyz":
        assert np.array_equal(getattr(s, "u" + c), st["u" + c])


def test_energy_conserved_in_closed_box_without_source():
    # interior-only initial bump, no source, before any flux reaches a
    # boundary: the mixed-level leapfrog energy is constant to rounding
    pb, k, omega, h, dt, steps = _plane_wave_problem(n=(24, 24, 24))
    pb.wave = np.zeros(8)
    s = rm.State(*pb.shape)
    X, Y, Z = np.meshgrid(*(np.arange(n) for n in pb.shape), indexing="ij")
    bump = np.exp(-((X - 12) ** 2 + (Y - 12) ** 2 + (Z - 12) ** 2) / 6.0)
    s.ux[...] = bump[:-1]
    s.ux_old[...] = bump[:-1]
    s.ux_new[...] = bump[:-1]
    es = []
    for tt in range(5):
        rm.step(s, pb, tt)
        es.append(rm.energy(s, pb))
    es = np.array(es)
    assert es.min() > 0
    assert (es.max() - es.min()) / es.mean() < 1e-8


def _courant_one_pulse(steps, mur_x=None, n=(50, 100, 100)):
    """A P pulse, uniform in y and z, moving +x at Courant 1 (dt = h / v_p)
    in a homogeneous box; ux on the box's centre line after every step (C
    oracle).  mur_x replaces the x = max face's ux coefficient."""
    x, y, z = (np.arange(k, dtype=np.float64) for k in n)
    rho, lam, mu = 2700.0, 5.1e10, 2.6e10
    c = np.zeros((6, 6))
    c[:3, :3] = lam
    for a in range(3):
        c[a, a] = lam + 2 * mu
        c[3 + a, 3 + a] = mu
    table = brute.material_table([c], [rho])
    ids = np.zeros(n, dtype=np.uint8)
    dt = 1.0 / np.sqrt((lam + 2 * mu) / rho)
    pb = rm.Problem(x, y, z, *rm.dense_material(ids, [c], [rho]), dt, np.zeros(steps))
    mc = pb.mur_c
    mur12 = np.array(list(mc["xmax"]) + list(mc["y0"]) + list(mc["ymax"]) + list(mc["zmax"]))
    if mur_x is not None:
        mur12[0] = mur_x
    xs = np.arange(n[0] - 1) + 0.5
    f = lambda X: np.exp(-((X - 15) / 2.0) ** 2)
    shp = {"x": (n[0] - 1, n[1], n[2]), "y": (n[0], n[1] - 1, n[2]), "z": (n[0], n[1], n[2] - 1)}
    st = {"u" + q + suf: np.zeros(shp[q]) for q in "xyz" for suf in ("", "_new", "_old")}
    st["ux"][...] = f(xs)[:, None, None]
    st["ux_old"][...] = f(xs + 1)[:, None, None]  # one step earlier: moving +x
    st["ux_new"][...] = st["ux"]
    line = []
    for tt in range(steps):
        st, _ = brute.run(x, y, z, ids, table, dt, mur12, pb.wave, 1, state=st, t0=tt, omp=True)
        line.append(st["ux"][:, n[1] // 2, n[2] // 2].copy())
    return np.array(line), f, xs, (dt * np.sqrt((lam + 2 * mu) / rho), mc, np.sqrt(mu / rho) * dt)


def test_mur_absorbs_a_normal_plane_wave_exactly_at_courant_one():
    # At Courant 1 the 1-D leapfrog shifts a wave one cell per step exactly,
    # and the first-order Mur coefficient (v dt - h) / (v dt + h) is 0, so
    # the face passes the wave out with no reflection (base_solver.py:
    # 525-544).  On the box's centre line the 3-D scheme is that 1-D scheme
    # for as many steps as the line is cells away from the y and z faces and
    # the surface (domain of dependence).
    steps = 44
    line, f, xs, (nu, mc, vt_dt) = _courant_one_pulse(steps)
    assert nu == pytest.approx(1.0, abs=1e-15) and abs(mc["xmax"][0]) < 1e-15
    for s in (10, 20, 30):
        assert np.abs(line[s - 1] - f(xs - s)).max() < 1e-12, s  # exact shift
    assert np.abs(line[-1]).max() < 1e-10  # the pulse has left: nothing came back
    # planted mistake: the shear speed in the x face's ux coefficient
    wrong = (vt_dt - 1.0) / (vt_dt + 1.0)
    bad, _, _, _ = _courant_one_pulse(steps, mur_x=wrong)
    assert np.abs(bad[-1]).max() > 0.1
