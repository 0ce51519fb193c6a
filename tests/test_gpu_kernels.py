"""The LDS-tiled sweep against the naive sweep: same formulas on the same
operands, so identical bits in every precision, on uniform and non-uniform
meshes, with and without the sponge, for every tile shape."""
import numpy as np
import pytest

from helpers import CASES, golden, spec_from, spec_from_golden, uniform_crystal
from phonomena_amd import fdtd, problem

pytestmark = pytest.mark.gpu
FIELDS = ["ux", "uy", "uz", "ux_old", "uy_old", "uz_old"]


def _fields(spec, precision, kernel, tile_j=0, tile_x=0, steps=None, tile_z=0):
    ctx = fdtd.Context(spec, precision=precision, kernel=kernel, tile_j=tile_j, tile_x=tile_x, tile_z=tile_z)
    ctx.step(spec.wave.size if steps is None else steps)
    ctx.sync()
    out = {f: ctx.get(f) for f in FIELDS}
    out["probes"] = ctx.probes()
    ctx.close()
    return out


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp32c"])
def test_tiled_equals_naive_on_golden_meshes(name, precision):
    spec = spec_from_golden(golden(name))
    a = _fields(spec, precision, "naive")
    b = _fields(spec, precision, "tiled")
    for f in a:
        assert np.array_equal(a[f], b[f]), f


@pytest.mark.parametrize("name", ["default_json_100", "nonuniform_json_60"])
@pytest.mark.parametrize("rows", [4, 8])
def test_strip_equals_naive(name, rows):
    spec = spec_from_golden(golden(name))
    a = _fields(spec, "fp32", "naive")
    b = _fields(spec, "fp32", "strip", tile_j=rows)
    for f in a:
        assert np.array_equal(a[f], b[f]), f


def test_fp64_tiled_bitwise_vs_reference():
    for name in CASES:
        d = golden(name)
        out = _fields(spec_from_golden(d), "fp64", "tiled")
        for f in FIELDS:
            assert np.array_equal(out[f], d[f]), (name, f)


@pytest.mark.parametrize("tile_j,tile_x,tile_z", [(4, 0, 64), (4, 2, 64), (8, 1, 64), (8, 7, 64), (16, 0, 64),
                                                 (16, 33, 64), (16, 70, 64), (8, 3, 32), (16, 0, 32), (32, 0, 32),
                                                 (32, 9, 32)])
@pytest.mark.parametrize("precision", ["fp32", "fp32c"])
def test_tiled_shapes_on_crystal(tile_j, tile_x, tile_z, precision):
    # z extent not a multiple of the 62 output lanes, y not of the row count
    x, y, z, ids, rows, dt, wave = uniform_crystal(
        (70, 45, 131), 1.0, incl=(19300.0, 2.0e11, 1.6e11, 14, 5), steps=120, seed=5)
    spec = spec_from(x, y, z, ids, rows, dt, wave, probe=(0, 3, 24, 1, 2, 23), every=7)
    a = _fields(spec, precision, "naive")
    b = _fields(spec, precision, "tiled", tile_j=tile_j, tile_x=tile_x, tile_z=tile_z)
    for f in a:
        assert np.array_equal(a[f], b[f]), f


@pytest.mark.parametrize("tile_j,tile_x,tile_z", [(8, 0, 64), (16, 3, 32), (8, 5, 32)])
def test_tiled_fp64_shapes_on_crystal(tile_j, tile_x, tile_z):
    x, y, z, ids, rows, dt, wave = uniform_crystal(
        (70, 45, 131), 1.0, incl=(19300.0, 2.0e11, 1.6e11, 14, 5), steps=60, seed=5)
    spec = spec_from(x, y, z, ids, rows, dt, wave)
    a = _fields(spec, "fp64", "naive")
    b = _fields(spec, "fp64", "tiled", tile_j=tile_j, tile_x=tile_x, tile_z=tile_z)
    for f in a:
        assert np.array_equal(a[f], b[f]), f


@pytest.mark.parametrize("tile_j,tile_z", [(32, 32), (16, 64)])
def test_tiled_interior_tiles_on_larger_mesh(tile_j, tile_z):
    # big enough that most blocks take the boundary-free (interior) path
    x, y, z, ids, rows, dt, wave = uniform_crystal(
        (120, 100, 140), 1.0, incl=(19300.0, 2.0e11, 1.6e11, 20, 7), steps=40, seed=11)
    spec = spec_from(x, y, z, ids, rows, dt, wave)
    a = _fields(spec, "fp32", "naive")
    b = _fields(spec, "fp32", "tiled", tile_j=tile_j, tile_z=tile_z)
    for f in a:
        assert np.array_equal(a[f], b[f]), f


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp32c"])
def test_tiled_sponge_equals_naive(precision):
    x, y, z, ids, rows, dt, wave = uniform_crystal(
        (40, 40, 40), 1.0, incl=(1.2, 1.4e5, 0.0, 10, 3), steps=150, seed=9)
    sponge = problem.sponge_tables((41, 41, 41), {"width": 8, "strength": 0.15})
    spec = spec_from(x, y, z, ids, rows, dt, wave, sponge=sponge)
    a = _fields(spec, precision, "naive")
    b = _fields(spec, precision, "tiled")
    for f in a:
        assert np.array_equal(a[f], b[f]), f


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp32c"])
def test_measured_shape_starts_clear_and_matches_a_fixed_shape(precision):
    """A context whose tile shape is left open times the built shapes at
    creation on its own fields: it must start at step 0 with every level
    and the probe ring clear, and step to the same bits as a context with a
    fixed shape."""
    x, y, z, ids, rows, dt, wave = uniform_crystal((64, 48, 40), 1.0, steps=30, seed=5)
    spec = spec_from(x, y, z, ids, rows, dt, wave, probe=(0, 4, 8, 0, 4, 8), every=5)
    ctx = fdtd.Context(spec, precision=precision)
    try:
        kernel, tj, tz = ctx.kernel_info()
        assert kernel == "tiled"
        built = {(8, 32), (16, 32), (8, 64)} if precision == "fp64" else {(16, 64), (8, 64), (32, 32), (16, 32),
                                                                           (8, 32)}
        assert (tj, tz) in built
        assert ctx.steps_done == 0
        for f in FIELDS:
            assert not np.any(ctx.get(f)), f
        assert ctx.probes().shape[0] == 0
        ctx.step(30)
        ctx.sync()
        got = {f: ctx.get(f) for f in FIELDS}
        got["probes"] = ctx.probes()
    finally:
        ctx.close()
    tj0, tz0 = fdtd.STATIC_TILES[precision]
    want = _fields(spec, precision, "tiled", tile_j=tj0, tile_z=tz0, steps=30)
    for f in want:
        assert np.array_equal(got[f], want[f]), f


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp32c"])
def test_unmeasured_context_takes_the_fixed_shape(precision):
    """tune=False (contexts that only step inside a batch) skips the shape
    measurement and runs the fixed default shape."""
    x, y, z, ids, rows, dt, wave = uniform_crystal((40, 32, 24), 1.0, steps=4, seed=1)
    spec = spec_from(x, y, z, ids, rows, dt, wave)
    ctx = fdtd.Context(spec, precision=precision, tune=False)
    try:
        assert ctx.kernel_info() == ("tiled",) + fdtd.STATIC_TILES[precision]
    finally:
        ctx.close()


@pytest.mark.parametrize("precision,kernel,tiles", [("fp64", "tiled", (8, 32)), ("fp32", "tiled", (8, 64)),
                                                     ("fp32c", "tiled", (32, 32)), ("fp32", "tiled2", (32, 32)),
                                                     ("fp32c", "tiled2", (16, 64))])
def test_tiled_and_pair_match_naive_with_layer_on_larger_mesh(precision, kernel, tiles):
    # a mesh large enough for boundary-free tiles both inside and across the
    # absorbing layer, from a random start, so every tile kind computes
    # non-trivial values: the tiled sweeps give the naive sweep's bits
    x, y, z, ids, rows, dt, wave = uniform_crystal(
        (150, 120, 150), 1.0, incl=(19300.0, 2.0e11, 1.6e11, 18, 6), steps=30, seed=8, wave_kind="gauss",
        f=1.0 / (12 * 0.5 / 6176.0))
    nodes = (x.size, y.size, z.size)
    sponge = problem.sponge_tables(nodes, {"width": 10, "strength": 0.15})
    rng = np.random.default_rng(3)
    spec = spec_from(x, y, z, ids, rows, dt, wave, probe=(0, 4, 36, 0, 4, 30), every=5, sponge=sponge)
    start = {c: rng.standard_normal(s) for c, s in (("ux", (nodes[0] - 1,) + nodes[1:]),
                                                   ("uy", (nodes[0], nodes[1] - 1, nodes[2])),
                                                   ("uz", nodes[:2] + (nodes[2] - 1,)))}
    for c in ("uy", "uz"):
        start[c][0] = 0.0  # the x = 0 wall stays 0 in the reference

    def run(kern, tj=0, tz=0):
        ctx = fdtd.Context(spec, precision=precision, kernel=kern, tile_j=tj, tile_z=tz)
        for c in ("ux", "uy", "uz"):
            ctx.set(c, start[c])
            ctx.set(c + "_old", 0.5 * start[c])
        ctx.step(30)
        ctx.sync()
        out = {f: ctx.get(f) for f in FIELDS}
        out["probes"] = ctx.probes()
        ctx.close()
        return out

    ref = run("naive")
    got = run(kernel, *tiles)
    for f in ref:
        assert np.array_equal(ref[f], got[f]), (precision, kernel, f)
