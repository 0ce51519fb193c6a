"""The absorbing layer (an extension with no reference counterpart: the
reference absorbs with Mur faces only, base_solver.py:519-554), pinned by
numbers and closed forms rather than by restating its formula:

* the damping profile s(d) = strength * ((width - d) / width)^2 at chosen
  distances, as literal numbers, on the faces it belongs to (x = max, both
  y faces, z = max; never the x = 0 source wall or the z = 0 free surface);
* one damped step from a spatially uniform state, where every stress
  difference vanishes and the update reduces to
  u+ = (2 u - (1 - s) u-) / (1 + s), against hand-derived values;
* the decay rate of a wave in a uniformly damped medium against the
  closed form exp(-s x / (v dt)) (damping sigma = 2 s / dt, amplitude
  decay sigma / (2 v) per unit length for sigma << omega).

Each check is also run on a planted mistake (a linear profile, damping on
the wrong faces, a wrong damping factor), stated here on the CPU, to show
that it would fail.  The GPU's FP64 sweep equals the oracle bit for bit with
the layer (test_gpu_parity.test_sponge_fp64_bitwise_vs_oracle), so these
pins hold for the device as well.
"""
import numpy as np
import pytest

from oracle import brute
from oracle import reference_math as rm
from phonomena_amd import problem

# (n, width, strength): a 40-node axis, 16-cell layer, strength 0.15 (the
# benchmark's layer)
N, W, S = 40, 16, 0.15
# distance from the damped face -> s(d), written out
PROFILE_NUMBERS = {0: 0.15, 4: 0.084375, 8: 0.0375, 12: 0.009375, 15: 0.15 / 256, 16: 0.0, 20: 0.0}


def _check_profile(fn):
    """fn(n, width, strength, lo, hi) -> per-node damping, as sponge_axis."""
    hi_only = fn(N, W, S, False, True)
    lo_only = fn(N, W, S, True, False)
    both = fn(N, W, S, True, True)
    for d, v in PROFILE_NUMBERS.items():
        assert hi_only[N - 1 - d] == pytest.approx(v, rel=1e-15, abs=0), ("hi", d)
        assert lo_only[d] == pytest.approx(v, rel=1e-15, abs=0), ("lo", d)
    assert not hi_only[:N - W].any() and not lo_only[W:].any()
    assert np.array_equal(both, lo_only + hi_only)


def _check_faces(tables_fn):
    """tables_fn(nodes, cfg) -> [sx, sy, sz], as problem.sponge_tables."""
    sx, sy, sz = tables_fn((30, 24, 20), {"width": 4, "strength": 0.25})
    # x: the far face only (x = 0 carries the source line)
    assert sx[0] == 0.0 and sx[1] == 0.0 and sx[29] == 0.25 and sx[28] == 0.140625
    # y: both faces
    assert sy[0] == 0.25 and sy[23] == 0.25 and sy[2] == 0.0625 and sy[12] == 0.0
    # z: the bottom only (z = 0 is the free surface)
    assert sz[0] == 0.0 and sz[19] == 0.25 and sz[16] == 0.015625


def test_profile_values_are_the_stated_numbers():
    _check_profile(problem.sponge_axis)
    _check_profile(rm.sponge_profile)


def test_profile_faces():
    _check_faces(problem.sponge_tables)


def _linear_profile(n, width, strength, lo, hi):
    s = np.zeros(n)
    for i in range(n):
        if lo and i < width:
            s[i] += strength * (width - i) / width
        if hi and n - 1 - i < width:
            s[i] += strength * (width - (n - 1 - i)) / width
    return s


def _wrong_face_tables(nodes, cfg):
    # plants the x layer on the source wall instead of the far face
    w, st = cfg["width"], cfg["strength"]
    return [problem.sponge_axis(nodes[0], w, st, True, False), problem.sponge_axis(nodes[1], w, st, True, True),
            problem.sponge_axis(nodes[2], w, st, False, True)]


def test_planted_linear_profile_fails():
    with pytest.raises(AssertionError):
        _check_profile(_linear_profile)


def test_planted_wrong_face_fails():
    with pytest.raises(AssertionError):
        _check_faces(_wrong_face_tables)


# ---------------------------------------------------- uniform state ----

def _homogeneous(n):
    x, y, z = (np.arange(k, dtype=np.float64) for k in n)
    lam, mu, rho = 5.1e10, 2.6e10, 2700.0
    c = np.zeros((6, 6))
    c[:3, :3] = lam
    for a in range(3):
        c[a, a] = lam + 2 * mu
        c[3 + a, 3 + a] = mu
    ids = np.zeros(n, dtype=np.uint8)
    return x, y, z, ids, c, rho


# ux cell -> hand-derived value after one step from u = 1, u_old = 1/2
# (width 4, strength 1/4 on a 16^3-node box: s = 1/4 (3/4)^2 = 9/64 one
# node in from a damped face, 1/16 two nodes in, 1/64 three nodes in; a
# cell's s is the sum over its three axes):
#   s = 0      -> (2 - 1/2) / 1                 = 1.5
#   s = 1/64   -> (2 - (63/64)/2) / (65/64)     = 1.4846153846153847
#   s = 1/16   -> (2 - (15/16)/2) / (17/16)     = 1.4411764705882353
#   s = 9/64   -> (2 - (55/64)/2) / (73/64)     = 1.3767123287671232
#   s = 11/32  -> (2 - (21/32)/2) / (43/32)     = 1.244186046511628
UNIFORM_CASES = {
    (5, 8, 8): 1.5,                    # interior, no layer
    (1, 8, 8): 1.5,                    # next to the source wall: not damped
    (5, 8, 1): 1.5,                    # under the free surface: not damped
    (12, 8, 8): 1.4846153846153847,    # x, three nodes from x = max
    (13, 8, 8): 1.4411764705882353,    # x, two nodes from x = max
    (5, 14, 8): 1.3767123287671232,    # y, one node from y = max
    (5, 1, 8): 1.3767123287671232,     # y, one node from y = 0
    (5, 8, 14): 1.3767123287671232,    # z, one node from z = max
    (13, 14, 14): 1.244186046511628,   # corner: 1/16 + 9/64 + 9/64
}


def _one_uniform_step(tables):
    n = (16, 16, 16)
    x, y, z, ids, c, rho = _homogeneous(n)
    C, P = rm.dense_material(ids, [c], [rho])
    dt = 0.5 / np.sqrt(c[0, 0] / rho)
    pb = rm.Problem(x, y, z, C, P, dt, np.zeros(1), sponge=tables)
    s = rm.State(*n)
    for q in "xyz":
        u = np.ones(getattr(s, "u" + q).shape)
        if q != "x":
            u[0] = 0.0  # the x = 0 wall of uy/uz stays 0 in the reference
        getattr(s, "u" + q)[...] = u
        getattr(s, "u" + q + "_new")[...] = u
        getattr(s, "u" + q + "_old")[...] = 0.5 * u
    rm.step(s, pb, 0)
    return s


def _check_uniform(tables):
    s = _one_uniform_step(tables)
    for cell, v in UNIFORM_CASES.items():
        assert s.ux[cell] == v, (cell, s.ux[cell], v)


def test_damped_step_on_uniform_state_gives_hand_derived_values():
    _check_uniform(problem.sponge_tables((16, 16, 16), {"width": 4, "strength": 0.25}))


def test_planted_mistakes_fail_the_uniform_step():
    with pytest.raises(AssertionError):
        _check_uniform(_wrong_face_tables((16, 16, 16), {"width": 4, "strength": 0.25}))
    lin = [_linear_profile(16, 4, 0.25, False, True), _linear_profile(16, 4, 0.25, True, True),
           _linear_profile(16, 4, 0.25, False, True)]
    with pytest.raises(AssertionError):
        _check_uniform(lin)


# --------------------------------------------------- damped medium ----

def _beam_peaks(s_uniform, probes, nx=220, ny=40, nz=40, x0=120, steps=300):
    """Peak |ux| on the axis of a P-wave beam travelling +x (Gaussian
    envelope of 12 cells times a 16-cell carrier, 10-cell Gaussian cross
    section) at the probe planes, in a homogeneous box whose every node
    carries the damping term s_uniform (C oracle)."""
    n = (nx, ny, nz)
    x, y, z, ids, c, rho = _homogeneous(n)
    table = brute.material_table([c], [rho])
    vp = np.sqrt(c[0, 0] / rho)
    dt = 0.5 / vp
    fd = [np.diff(a) for a in (x, y, z)]
    sd = [np.mean([f[1:], f[:-1]], axis=0) for f in fd]
    mur = problem.mur_coefficients(table[0], dt, fd, sd)
    sponge = [np.full(nx, s_uniform), np.zeros(ny), np.zeros(nz)]
    shp = {"x": (nx - 1, ny, nz), "y": (nx, ny - 1, nz), "z": (nx, ny, nz - 1)}
    st = {"u" + q + suf: np.zeros(shp[q]) for q in "xyz" for suf in ("", "_new", "_old")}
    xs = np.arange(nx - 1) + 0.5
    hy = np.exp(-((np.arange(ny) - (ny - 1) / 2) / 10.0) ** 2)
    hz = np.exp(-((np.arange(nz) - (nz - 1) / 2) / 10.0) ** 2)
    g = lambda X: np.exp(-((X - x0) / 12.0) ** 2) * np.cos(2 * np.pi * (X - x0) / 16.0)
    cross = hy[None, :, None] * hz[None, None, :]
    st["ux"][...] = g(xs)[:, None, None] * cross
    st["ux_old"][...] = g(xs + vp * dt)[:, None, None] * cross  # one step earlier: moving +x
    st["ux_new"][...] = st["ux"]
    peaks = np.zeros(len(probes))
    wave = np.zeros(steps)
    for t0 in range(0, steps, 2):
        st, _ = brute.run(x, y, z, ids, table, dt, mur, wave, 2, state=st, t0=t0, sponge=sponge)
        peaks = np.maximum(peaks, np.abs(st["ux"][probes, ny // 2, nz // 2]))
    return peaks, vp * dt


def _decay_per_10_cells(s_uniform, scale=1.0):
    probes = [129, 139, 149, 159, 169, 179]
    ref, vdt = _beam_peaks(0.0, probes)
    got, _ = _beam_peaks(s_uniform * scale, probes)
    r = got / ref
    return (r[-1] / r[0]) ** (1.0 / 5), np.exp(-10 * s_uniform / vdt)


def test_uniform_damping_follows_the_closed_form_decay():
    # sigma = 2 s / dt; for sigma << omega a plane wave decays as
    # exp(-sigma x / (2 v)) = exp(-s x / (v dt)).  Carrier: omega dt =
    # 2 pi v dt / 16 = 0.196, sigma dt = 0.01, so sigma / omega = 0.05.
    measured, closed = _decay_per_10_cells(0.005)
    assert measured == pytest.approx(closed, rel=1e-2), (measured, closed)
    # a damping term twice too strong (e.g. sigma dt instead of sigma dt / 2)
    # is far outside that tolerance
    wrong, _ = _decay_per_10_cells(0.005, scale=2.0)
    assert abs(wrong - closed) > 3e-2 * closed


def test_sponge_absorbs_more_energy_than_mur_alone():
    # total energy left in a box after a source pulse, with and without the
    # layer: the layer removes most of it
    from helpers import golden, oracle_problem
    d = golden("homog16_200")
    pb = oracle_problem(d)
    steps = 200
    wave = np.r_[pb.wave[:40], np.zeros(steps - 40)]
    prof = [rm.sponge_profile(n, 5, 0.15, lo, hi) for n, (lo, hi) in
            zip(pb.shape, ((False, True), (True, True), (False, True)))]
    e_plain = rm.energy(rm.run(oracle_problem(d, wave=wave), steps)[0], pb)
    pbs = oracle_problem(d, wave=wave, sponge=prof)
    e_sponge = rm.energy(rm.run(pbs, steps)[0], pbs)
    assert e_sponge < 0.5 * e_plain
