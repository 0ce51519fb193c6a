"""pf_get_field_noalloc through the C-ABI: the same values as pf_get_field,
bit for bit, on two calls in a row (the second on a smaller mesh, so the
shared static staging array is reused with less data), in every precision,
for every field selector; fields beyond its limit are refused."""
import numpy as np
import pytest

from helpers import spec_from, uniform_crystal
from phonomena_amd import fdtd

pytestmark = pytest.mark.gpu
FIELDS = ["ux", "uy", "uz", "ux_old", "uy_old", "uz_old"]


def _context(cells, precision, steps=30, seed=1):
    x, y, z, ids, rows, dt, wave = uniform_crystal(cells, 1.0, incl=(19300.0, 2.0e11, 1.6e11, 8, 3),
                                                   steps=steps, seed=seed)
    ctx = fdtd.Context(spec_from(x, y, z, ids, rows, dt, wave), precision=precision)
    ctx.step(steps)
    ctx.sync()
    return ctx


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp32c"])
def test_noalloc_equals_get_field_two_calls(precision):
    big = _context((60, 50, 40), precision, seed=1)
    small = _context((20, 17, 12), precision, steps=25, seed=2)
    for ctx in (big, small):  # the second call stages fewer elements
        nonzero = False
        for f in FIELDS:
            a = ctx.get(f)
            b = ctx.get_noalloc(f)
            assert np.array_equal(a, b), (precision, f)
            nonzero = nonzero or bool(np.any(a != 0))
        assert nonzero
    # and back on the larger mesh after the smaller one
    for f in FIELDS:
        assert np.array_equal(big.get(f), big.get_noalloc(f)), (precision, f)
    big.close()
    small.close()


def test_noalloc_refuses_fields_beyond_its_array():
    nx, ny, nz = 170, 160, 160  # 4.35 M nodes > 4 M
    x, y, z = (np.arange(n, dtype=np.float64) for n in (nx, ny, nz))
    ids = np.zeros((nx, ny, nz), dtype=np.uint8)
    rows = np.array([[1.03e11, 5.1e10, 5.1e10, 5.1e10, 1.03e11, 5.1e10, 5.1e10, 5.1e10, 1.03e11,
                      2.6e10, 2.6e10, 2.6e10, 2700.0]])
    ctx = fdtd.Context(spec_from(x, y, z, ids, rows, 1e-5, np.zeros(2)), precision="fp32", tune=False)
    assert (nx - 1) * ny * nz > fdtd.NOALLOC_FIELD_MAX
    with pytest.raises(fdtd.FdtdError, match="at most"):
        ctx.get_noalloc("ux")
    ctx.close()
