"""The opt-in fixed free surface (pf_desc.surface = 1, ``free_surface:
"fixed"``; no counterpart in the reference) on the GPU, kept apart from the
reference-parity tests:

* FP64 at a physical spacing (h = 1e-4 m) equals the oracle's fixed scheme
  bit for bit, for the tiled, naive and strip sweeps; the two-step pass
  equals the one-step sweep bit for bit in FP32 and FP32C;
* the benchmark's plain CPU case at its physical spacing (64x64x32 cells,
  dx = 1e-4 m, homogeneous, Courant 0.5, 1 MHz sin source on the x = 0
  wall, 8-cell absorbing layer, 1000 steps, surface u_z every 10 steps):
  FP64 with the fixed surface equals the reference's scheme run at unit
  spacing (where the two coincide) to rounding, for every field and probe
  frame, and FP32 / FP32C stay within rel-L2 1e-4 / 1e-5 of it.
"""
import numpy as np
import pytest

from helpers import rel_l2, spec_from, uniform_crystal
from oracle import brute
from oracle import reference_math as rm
from phonomena_amd import fdtd, problem
from test_surface import box_inputs

pytestmark = pytest.mark.gpu
FIELDS = ["ux", "uy", "uz", "ux_old", "uy_old", "uz_old"]


def _box_spec(h, steps):
    x, y, z, ids, cs, rhos, dt, wave = box_inputs(h, steps)
    spec = spec_from(x, y, z, ids, brute.material_table(cs, rhos), dt, wave)
    spec.surface = 1
    return spec


@pytest.mark.parametrize("kernel,h", [("tiled", 1e-4), ("naive", 1e-4), ("strip", 1e-4), ("tiled", 2.0 ** -13)])
def test_fp64_fixed_surface_equals_oracle_at_physical_spacing(kernel, h):
    # h = 2^-13 m (1.22e-4): the tiled sweep's reciprocal-product flavour
    steps = 60
    x, y, z, ids, cs, rhos, dt, wave = box_inputs(h, steps)
    C, P = rm.dense_material(ids, cs, rhos)
    ref, _ = rm.run(rm.Problem(x, y, z, C, P, dt, wave, surface_fixed=True), steps)
    ctx = fdtd.Context(_box_spec(h, steps), precision="fp64", kernel=kernel,
                       tile_j=8 if kernel == "strip" else 0)
    ctx.step(steps)
    for f in FIELDS:
        assert np.array_equal(ctx.get(f), getattr(ref, f)), (kernel, f)
    ctx.close()


@pytest.mark.parametrize("precision", ["fp32", "fp32c"])
def test_pair_equals_tiled_with_fixed_surface(precision):
    steps = 41
    out = []
    for kernel in ("tiled", "tiled2"):
        ctx = fdtd.Context(_box_spec(1e-4, steps), precision=precision, kernel=kernel)
        ctx.step(steps)
        out.append({f: ctx.get(f) for f in FIELDS})
        ctx.close()
    for f in FIELDS:
        assert np.array_equal(out[0][f], out[1][f]), f


def _c1(d, surface, precision, wave=None, steps=1000):
    """The plain CPU case at spacing d (the source frequency scaled so that
    f * dt is the same at every d)."""
    x, y, z, ids, rows, dt, w = uniform_crystal((64, 64, 32), d, steps=steps, wave_kind="sin",
                                                f=1e6 * 1e-4 / d)
    nodes = (x.size, y.size, z.size)
    spec = spec_from(x, y, z, ids, rows, dt, w if wave is None else wave, probe=(0, 2, 32, 0, 2, 32), every=10,
                     sponge=problem.sponge_tables(nodes, {"width": 8}))
    spec.surface = surface
    ctx = fdtd.Context(spec, precision=precision)
    ctx.step(steps)
    out = {f: ctx.get(f) for f in FIELDS}
    out["probes"] = ctx.probes()
    ctx.close()
    return out, spec.wave


def test_c1_case_at_physical_spacing():
    phys, wave = _c1(1e-4, 1, "fp64")
    unit, _ = _c1(1.0, 0, "fp64", wave=wave)
    assert float(np.abs(phys["uz"]).max()) > 0
    for f in phys:
        r = rel_l2(phys[f], unit[f])
        print("fp64 fixed at 1e-4 m vs reference at unit spacing", f, r)
        assert r < 1e-9, (f, r)
    for precision, tol in (("fp32", 1e-4), ("fp32c", 1e-5)):
        got, _ = _c1(1e-4, 1, precision, wave=wave)
        for f in phys:
            r = rel_l2(got[f], phys[f])
            print(precision, "vs fp64, fixed surface at 1e-4 m", f, r)
            assert r < tol, (precision, f, r)


def test_batch_members_keep_their_surface():
    # a fixed-surface member at a physical spacing beside a reference-surface
    # member at unit spacing: each equals its alone run bit for bit
    steps = 32
    specs = [_box_spec(1e-4, steps), _box_spec(1.0, steps)]
    specs[1].surface = 0
    for precision in ("fp64", "fp32c"):
        ctxs = [fdtd.Context(s, precision=precision, tune=False) for s in specs]
        b = fdtd.Batch(precision)
        for c in ctxs:
            b.add(c)
        b.step(steps)
        b.sync()
        batched = [{f: c.get(f) for f in FIELDS} for c in ctxs]
        b.close()
        for c in ctxs:
            c.close()
        for s, got in zip(specs, batched):
            ctx = fdtd.Context(s, precision=precision)
            ctx.step(steps)
            for f in FIELDS:
                assert np.array_equal(ctx.get(f), got[f]), (precision, s.surface, f)
            ctx.close()
