"""The two-step tiled sweep (levels n+1 and n+2 per pass) against the
one-step tiled sweep: the same per-value formulas on the same operands, so
identical bits in FP32 and FP32C for every field, both time levels and the
probe frames, over even and odd step counts, both tile shapes, several x
chunkings, with and without the sponge."""
import numpy as np
import pytest

from helpers import CASES, FIELDS, golden, random_start, run_from, spec_from, spec_from_golden, uniform_crystal
from phonomena_amd import fdtd, problem, sweep

pytestmark = pytest.mark.gpu


def _run(spec, precision, kernel, chunks, **tiles):
    ctx = fdtd.Context(spec, precision=precision, kernel=kernel, **tiles)
    for n in chunks:
        ctx.step(n)
    ctx.sync()
    out = {f: ctx.get(f) for f in FIELDS}
    out["probes"] = ctx.probes()
    ctx.close()
    return out


def _same(a, b):
    for f in a:
        assert np.array_equal(a[f], b[f]), f


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("precision", ["fp32", "fp32c"])
def test_pair_equals_tiled_on_golden_meshes(name, precision):
    spec = spec_from_golden(golden(name))
    if min(spec.nodes) < 8:
        pytest.skip("the two-step sweep needs >= 8 nodes per axis")
    steps = spec.wave.size
    a = _run(spec, precision, "tiled", [steps])
    b = _run(spec, precision, "tiled2", [steps])
    _same(a, b)


@pytest.mark.parametrize("tile_j,tile_z,tile_x", [(16, 64, 0), (32, 32, 0), (16, 64, 5), (32, 32, 9),
                                                 (32, 32, 32), (16, 64, 250)])
@pytest.mark.parametrize("precision", ["fp32", "fp32c"])
def test_pair_shapes_on_crystal(tile_j, tile_z, tile_x, precision):
    x, y, z, ids, rows, dt, wave = uniform_crystal(
        (70, 45, 131), 1.0, incl=(19300.0, 2.0e11, 1.6e11, 14, 5), steps=121, seed=5)
    spec = spec_from(x, y, z, ids, rows, dt, wave, probe=(0, 3, 24, 1, 2, 23), every=7)
    a = _run(spec, precision, "tiled", [121])
    # an odd count: 60 pairs and one single step
    b = _run(spec, precision, "tiled2", [121], tile_j=tile_j, tile_z=tile_z, tile_x=tile_x)
    _same(a, b)


@pytest.mark.parametrize("precision", ["fp32", "fp32c"])
def test_pair_single_steps_between_pairs(precision):
    # single steps shift the pair phase and the level rotation
    x, y, z, ids, rows, dt, wave = uniform_crystal(
        (40, 33, 50), 1.0, incl=(19300.0, 2.0e11, 1.6e11, 10, 3), steps=60, seed=2)
    spec = spec_from(x, y, z, ids, rows, dt, wave, probe=(1, 2, 15, 0, 3, 11), every=3)
    a = _run(spec, precision, "tiled", [60])
    b = _run(spec, precision, "tiled2", [1, 6, 3, 1, 1, 17, 2, 29])
    _same(a, b)


# Meshes whose last tile along y or z holds a few cells (65 nodes: 32-wide
# tiles own 0..27, 28..55 and the last one 56..64) or most of its lanes, in
# both tile shapes; from a random start, so every edge cell carries values
# from the first step: the one-step sweep's bits.
@pytest.mark.parametrize("cells", [(40, 64, 64), (40, 45, 64), (40, 64, 45), (40, 45, 45), (40, 64, 70),
                                   (40, 70, 64)])
@pytest.mark.parametrize("tile_j,tile_z", [(32, 32), (16, 64)])
@pytest.mark.parametrize("precision", ["fp32", "fp32c"])
def test_pair_edge_tiles_from_random_start(cells, tile_j, tile_z, precision):
    x, y, z, ids, rows, dt, wave = uniform_crystal(
        cells, 1.0, incl=(19300.0, 2.0e11, 1.6e11, 14, 5), steps=12, seed=4)
    nodes = (x.size, y.size, z.size)
    sponge = problem.sponge_tables(nodes, {"width": 6, "strength": 0.15})
    spec = spec_from(x, y, z, ids, rows, dt, wave, probe=(0, 3, 13, 1, 4, 11), every=3, sponge=sponge)
    start = random_start(nodes, 11)
    a = run_from(spec, start, precision, "tiled", 11)
    b = run_from(spec, start, precision, "tiled2", 11, tile_j=tile_j, tile_z=tile_z)
    _same(a, b)


@pytest.mark.parametrize("kernel", ["tiled", "tiled2"])
def test_profile_steps_like_step(kernel):
    # pf_profile issues pf_step's launches with timing events between some
    # of them: the same fields and probe frames.  Its first call creates
    # the context's events and later calls reuse them; 80 steps are 40
    # pairs, bracketed every fifth to stay within the eight brackets.
    x, y, z, ids, rows, dt, wave = uniform_crystal(
        (40, 33, 50), 1.0, incl=(19300.0, 2.0e11, 1.6e11, 10, 3), steps=157, seed=2)
    spec = spec_from(x, y, z, ids, rows, dt, wave, probe=(1, 2, 15, 0, 3, 11), every=3)
    a = _run(spec, "fp32", kernel, [157])
    ctx = fdtd.Context(spec, precision="fp32", kernel=kernel)
    for n in (5, 1, 80, 71):
        sweep_ms, face_ms = ctx.profile(n)
        assert sweep_ms > 0 and sweep_ms + face_ms > 0
    ctx.sync()
    b = {f: ctx.get(f) for f in FIELDS}
    b["probes"] = ctx.probes()
    ctx.close()
    _same(a, b)


@pytest.mark.parametrize("tile_j,tile_z", [(32, 32), (16, 64)])
@pytest.mark.parametrize("precision", ["fp32", "fp32c"])
def test_pair_interior_tiles_on_larger_mesh(tile_j, tile_z, precision):
    # big enough that most blocks take the boundary-free (interior) path
    x, y, z, ids, rows, dt, wave = uniform_crystal(
        (120, 100, 140), 1.0, incl=(19300.0, 2.0e11, 1.6e11, 20, 7), steps=40, seed=11)
    spec = spec_from(x, y, z, ids, rows, dt, wave, probe=(2, 5, 20, 3, 4, 20), every=4)
    a = _run(spec, precision, "tiled", [40])
    b = _run(spec, precision, "tiled2", [40], tile_j=tile_j, tile_z=tile_z)
    _same(a, b)


@pytest.mark.parametrize("precision", ["fp32", "fp32c"])
def test_pair_sponge_equals_tiled(precision):
    x, y, z, ids, rows, dt, wave = uniform_crystal(
        (40, 40, 40), 1.0, incl=(1.2, 1.4e5, 0.0, 10, 3), steps=151, seed=9)
    sponge = problem.sponge_tables((41, 41, 41), {"width": 8, "strength": 0.15})
    spec = spec_from(x, y, z, ids, rows, dt, wave, sponge=sponge)
    a = _run(spec, precision, "tiled", [151])
    b = _run(spec, precision, "tiled2", [151])
    _same(a, b)


@pytest.mark.parametrize("cells", [(7, 7, 7), (7, 9, 8), (12, 7, 10)])
@pytest.mark.parametrize("precision", ["fp32", "fp32c"])
def test_pair_smallest_meshes(cells, precision):
    # 8 nodes on an axis: the band's planes nx-3, nx-2 and rows 1, ny-3,
    # ny-2 leave the pass two rows and three planes of its own
    x, y, z, ids, rows, dt, wave = uniform_crystal(cells, 1.0, steps=24, seed=4)
    spec = spec_from(x, y, z, ids, rows, dt, wave, probe=(0, 1, cells[0] + 1, 0, 1, cells[1] + 1), every=2)
    a = _run(spec, precision, "tiled", [24])
    for tj, tz in ((16, 64), (32, 32)):
        _same(a, _run(spec, precision, "tiled2", [5, 19], tile_j=tj, tile_z=tz))


def test_pair_rejects_fp64_and_other_tiles():
    x, y, z, ids, rows, dt, wave = uniform_crystal(
        (20, 20, 20), 1.0, incl=(19300.0, 2.0e11, 1.6e11, 6, 2), steps=4, seed=1)
    spec = spec_from(x, y, z, ids, rows, dt, wave)
    with pytest.raises(fdtd.FdtdError, match="FP32"):
        fdtd.Context(spec, precision="fp64", kernel="tiled2")
    with pytest.raises(fdtd.FdtdError, match="16 x 64 or 32 x 32"):
        fdtd.Context(spec, precision="fp32", kernel="tiled2", tile_j=8, tile_z=64)


@pytest.mark.parametrize("cells,precision,want", [
    ((256, 256, 256), "fp32", "tiled2"),   # 257^3 nodes >= 2^24
    ((256, 256, 252), "fp32", "tiled"),    # 257 x 257 x 253 nodes < 2^24
    ((256, 256, 256), "fp32c", "tiled"),
    ((256, 256, 256), "fp64", "tiled"),
])
def test_auto_resolves_by_size_and_precision(cells, precision, want):
    """'auto' runs the two-step sweep for FP32 meshes of at least 2^24 nodes
    and the one-step sweep otherwise, as pf_kernel_info reports."""
    spec, _, _ = sweep.crystal_spec(cells, 24, 8, 3, 4, device_ids=True)
    ctx = fdtd.Context(spec, precision=precision)
    try:
        kernel, tj, tz = ctx.kernel_info()
        assert kernel == want
        if kernel == "tiled2":
            shapes = {(16, 64), (32, 32)}
        elif precision == "fp64":  # one-step shapes, measured at creation
            shapes = {(8, 32), (16, 32), (8, 64)}
        else:
            shapes = {(16, 64), (8, 64), (32, 32), (16, 32), (8, 32)}
        assert (tj, tz) in shapes
        ctx.step(3)  # one pair and one single step
        assert ctx.steps_done == 3
    finally:
        ctx.close()
