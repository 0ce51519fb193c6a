"""Host side of the device rasterizer: the inclusion list it is given
(problem.raster_targets: centres and radii as the reference's float32
target record holds them, and the z nodes each inclusion covers) and the
per-node rule it applies (problem.raster_id), against the reference mirror's
Grid.inclusionIndices + Material.setConstants (grid.py:158-175,
material.py:55-63) and the reference's own masks (mesh_cases.npz)."""
import os

import numpy as np
import pytest

from helpers import GOLDEN
from phonomena_amd import problem
from phonomena_amd.simulation import grid, material


def raster_map(x, y, nz, targets):
    """The device's rule, vectorised on the host: node (i, j, k) is 1 when
    some target has sqrt((x-tx)^2 + (y-ty)^2) < r and k < K."""
    kmax = np.zeros((x.size, y.size), dtype=np.int64)
    for tx, ty, r, kk in targets:
        X = x[:, None] - tx
        Y = y[None, :] - ty
        inside = np.sqrt(np.add(np.square(X), np.square(Y))) < r
        kmax = np.where(inside, np.maximum(kmax, int(kk)), kmax)
    return (np.arange(nz)[None, None, :] < kmax[:, :, None]).astype(np.uint8)


def _crystal_grid(sx, sy, sz, incs, dz=None):
    """Unit spacing, or (dz given) the reference mesher with max_dz = dz:
    graded x/y nodes around the inclusions and size_z / dz z cells."""
    g = grid.Grid()
    if dz is None:
        g.setSpacing(1.0)
    g.init(sx, sy, sz)
    for inc in incs:
        g.addInclusion(*inc)
    if dz is not None:
        g.max_dx, g.max_dy, g.max_dz, g.min_d, g.slope = 1.0, 1.0, dz, 0.5, 1.0
        g.buildMesh()
        g.update()
    return g


def _material_map(g):
    m = material.Material()
    m.init(g, {"a": material.isotropic("a", 2700.0, 5e10, 2.6e10),
               "b": material.isotropic("b", 19300.0, 2e11, 1.6e11)})
    m.setPrimary("a")
    m.setSecondary("b")
    m.update()
    return m.ids


CASES = [
    # full depth, partial depths in rising and falling order, off-grid centres
    (20, 18, 6, [(5.5, 5.25, 2.0), (12.3, 9.7, 3.1)], None),
    (24, 20, 8, [(6, 6, 3.0, 5.0), (14, 10, 4.0, 2.0), (7, 14, 2.5, 7.0)], None),
    (24, 20, 8, [(6, 6, 3.0, 2.0), (14, 10, 4.0, 6.0), (7, 14, 2.5, 8.0)], None),
    # mesher with more z nodes than size_z: from the second inclusion on
    # the depth is compared with the rebound index array, not coordinates
    (16, 16, 4, [(5, 5, 2.0, 4.0), (10, 10, 3.0, 3.5), (11, 4, 2.0, 4.0)], 0.5),
    (16, 16, 4, [(5, 5, 2.0, 3.0), (10, 10, 3.0, 4.0), (11, 4, 2.0, 1.5)], 0.25),
    # exact ties: integer centres and radii on a unit grid (3-4-5 triangles)
    (30, 30, 4, [(10, 10, 5.0), (20, 17, 5.0), (15, 24, 3.0)], None),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_z_counts_follow_inclusion_indices(case):
    sx, sy, sz, incs, dz = CASES[case]
    g = _crystal_grid(sx, sy, sz, incs, dz)
    _, _, tg = problem.raster_targets(g)
    for (yx, z), (_, _, _, k) in zip(g.inclusionIndices(), tg):
        assert np.array_equal(z, np.arange(int(k)))


@pytest.mark.parametrize("case", range(len(CASES)))
def test_rule_matches_material_map(case):
    sx, sy, sz, incs, dz = CASES[case]
    g = _crystal_grid(sx, sy, sz, incs, dz)
    x, y, tg = problem.raster_targets(g)
    ids = _material_map(g)
    assert np.array_equal(raster_map(x, y, g.z.size, tg), ids)
    # the scalar rule the host uses for the Mur row, on a sample of nodes
    rng = np.random.default_rng(case)
    for _ in range(200):
        i, j, k = (int(rng.integers(0, n)) for n in ids.shape)
        assert problem.raster_id(x[i], y[j], k, tg) == ids[i, j, k]


def test_rule_matches_reference_mesh_cases():
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
        x, y, tg = problem.raster_targets(g)
        assert np.array_equal(raster_map(x, y, g.z.size, tg), d["c%d_mask" % n]), n


def test_a_planted_mistake_fails():
    # the rule with <= instead of < differs on the exact-tie case
    sx, sy, sz, incs, dz = CASES[4]
    g = _crystal_grid(sx, sy, sz, incs, dz)
    x, y, tg = problem.raster_targets(g)
    kmax = np.zeros((x.size, y.size), dtype=np.int64)
    for tx, ty, r, kk in tg:
        inside = np.sqrt(np.add(np.square(x[:, None] - tx), np.square(y[None, :] - ty))) <= r
        kmax = np.where(inside, np.maximum(kmax, int(kk)), kmax)
    wrong = (np.arange(g.z.size)[None, None, :] < kmax[:, :, None]).astype(np.uint8)
    assert not np.array_equal(wrong, _material_map(g))


def test_descriptor_without_ids_is_checked_on_the_host():
    # pf_create with ids NULL validates the inclusion list (and never reads
    # the absent map) before any device call, so this runs without a GPU
    from phonomena_amd import fdtd
    g = _crystal_grid(*CASES[0])
    m = material.Material()
    m.init(g, {"a": material.isotropic("a", 2700.0, 5e10, 2.6e10),
               "b": material.isotropic("b", 19300.0, 2e11, 1.6e11)})
    m.setPrimary("a")
    m.setSecondary("b")
    m.update(rasterize=False)
    spec = problem.mesh_spec(g, m, np.zeros(2), {}, device_ids=True)
    x, y, tg = spec.raster
    bad = tg.copy()
    bad[1, 3] = 2.5
    spec.raster = (x, y, bad)
    with pytest.raises(fdtd.FdtdError, match="z nodes"):
        fdtd.Context(spec, precision="fp32")
