"""A mesh whose time level is larger than 4 GiB (FP64, 602 x 600 x 500
nodes, 4.3 GB per level): AUTO takes the naive sweep with 64-bit addressing
there.  A displacement pulse placed near the far corner (the highest
addresses) evolves exactly as the same pulse on a small mesh whose far faces
sit at the same distance: after 6 steps the stencil (one cell per step) has
not reached the faces the two meshes do not share, so the region around the
pulse is bit-identical between the large run and the small tiled run."""
import numpy as np
import pytest

from helpers import spec_from, uniform_crystal
from phonomena_amd import fdtd

pytestmark = pytest.mark.gpu
STEPS = 6
SMALL = (40, 36, 30)       # nodes
BIG = (602, 600, 500)


def _pulse(shape, centre):
    """exp(-r^2/8) within 6 nodes of centre, 0 beyond (compact support, so
    it can be written into a window of a large array)."""
    u = np.zeros(shape)
    lo = [max(0, c - 7) for c in centre]
    hi = [min(s, c + 8) for s, c in zip(shape, centre)]
    idx = np.indices([h - l for l, h in zip(lo, hi)], dtype=np.float64)
    r2 = sum((idx[a] + lo[a] - centre[a]) ** 2 for a in range(3))
    u[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = np.where(r2 <= 36.0, np.exp(-r2 / 8.0), 0.0)
    return u


def _run(nodes, kernel):
    cells = tuple(n - 1 for n in nodes)
    x, y, z, ids, rows, dt, wave = uniform_crystal(cells, 1.0, steps=STEPS, wave_kind="sin", seed=0)
    wave = np.zeros_like(wave)  # the pulse is the only excitation
    spec = spec_from(x, y, z, ids, rows, dt, wave)
    ctx = fdtd.Context(spec, precision="fp64", kernel=kernel)
    out = {}
    for c, shape in ctx._shapes.items():
        # pulse 10 nodes inside the far corner, in each component's array
        u = _pulse(shape, tuple(s - 10 for s in shape))
        ctx.set(c, u)
        ctx.set(c + "_old", u)
    del u
    ctx.step(STEPS)
    ctx.sync()
    for c, shape in ctx._shapes.items():
        a = ctx.get(c)
        # the last (small extent - 8) nodes along each axis: 8 or more nodes
        # from the small mesh's low faces, which the large mesh does not have
        sl = tuple(slice(s - (m - 8), s) for s, m in zip(shape, SMALL))
        out[c] = np.array(a[sl])
        del a
    ctx.close()
    return out


def test_level_above_4gib_runs_naive_and_matches_small_mesh():
    big_level = BIG[0] * BIG[1] * BIG[2] * 3 * 8
    assert big_level > 4 * 1024 ** 3
    small = _run(SMALL, "tiled")
    big = _run(BIG, "auto")
    with pytest.raises(fdtd.FdtdError, match="4 GiB"):
        cells = tuple(n - 1 for n in BIG)
        x, y, z, ids, rows, dt, wave = uniform_crystal(cells, 1.0, steps=2, wave_kind="sin", seed=0)
        fdtd.Context(spec_from(x, y, z, ids, rows, dt, wave), precision="fp64", kernel="tiled")
    for c in small:
        assert np.abs(small[c]).max() > 0.1, c
        assert np.array_equal(small[c], big[c]), c
