"""Batched stepping through the C-ABI: every member's fields and probes are
bit-identical to stepping it alone, whatever the tile shape, the graph
replay, the member mix, the step budgets, or members joining and leaving."""
import numpy as np
import pytest

from helpers import spec_from, uniform_crystal
from phonomena_amd import fdtd, problem

pytestmark = pytest.mark.gpu
FIELDS = ["ux", "uy", "uz", "ux_old", "uy_old", "uz_old"]

# (cells, inclusion, steps, seed): uneven sizes, both inclusion kinds
MIX = [((40, 36, 20), (19300.0, 2.0e11, 1.6e11, 10, 3), 90, 1),
       ((64, 50, 33), (1.2, 1.4e5, 0.0, 12, 4), 70, 2),
       ((23, 70, 45), (19300.0, 2.0e11, 1.6e11, 8, 2), 110, 3),
       ((90, 31, 64), (1.2, 1.4e5, 0.0, 16, 5), 60, 4),
       ((47, 47, 24), (19300.0, 2.0e11, 1.6e11, 9, 3), 100, 5)]


def _spec(cells, incl, steps, seed, sponge=None):
    x, y, z, ids, rows, dt, wave = uniform_crystal(cells, 1.0, incl=incl, steps=steps, seed=seed,
                                                   wave_kind="gauss")
    nodes = (x.size, y.size, z.size)
    sp = problem.sponge_tables(nodes, sponge) if sponge else None
    return spec_from(x, y, z, ids, rows, dt, wave, probe=(0, 3, nodes[0] // 3, 1, 2, (nodes[1] - 2) // 2),
                     every=5, sponge=sp)


def _state(ctx):
    out = {f: ctx.get(f) for f in FIELDS}
    out["probes"] = ctx.probes()
    out["steps"] = ctx.steps_done
    return out


def _alone(spec, precision, n):
    ctx = fdtd.Context(spec, precision=precision)
    ctx.step(min(n, spec.wave.size))
    ctx.sync()
    s = _state(ctx)
    ctx.close()
    return s


def _same(a, b, tag):
    for k in a:
        if k == "steps":
            assert a[k] == b[k], (tag, k)
        else:
            assert np.array_equal(a[k], b[k]), (tag, k)


@pytest.mark.parametrize("precision", ["fp32", "fp32c", "fp64"])
def test_batch_equals_alone(precision):
    specs = [_spec(*m) for m in MIX]
    n = 120  # beyond some members' budgets: they stop, the rest go on
    ref = [_alone(s, precision, n) for s in specs]
    b = fdtd.Batch(precision)
    ctxs = [fdtd.Context(s, precision=precision) for s in specs]
    for c in ctxs:
        b.add(c)
    b.step(37)      # graph chunks + loose steps
    b.step(n - 37)
    b.sync()
    for i, c in enumerate(ctxs):
        _same(ref[i], _state(c), (precision, i))
    b.close()
    for c in ctxs:
        c.close()


@pytest.mark.parametrize("shape", [(16, 64, 8), (8, 64, 32), (32, 32, 16), (16, 32, 64), (8, 32, 5)])
def test_batch_shapes_and_no_graph(shape):
    specs = [_spec(*m) for m in MIX[:3]]
    ref = [_alone(s, "fp32", 60) for s in specs]
    for graph in (True, False):
        b = fdtd.Batch("fp32")
        ctxs = [fdtd.Context(s, precision="fp32") for s in specs]
        for c in ctxs:
            b.add(c)
        b.set_tiles(*shape, graph=graph)
        b.step(60)
        b.sync()
        for i, c in enumerate(ctxs):
            _same(ref[i], _state(c), (shape, graph, i))
        b.close()
        for c in ctxs:
            c.close()


def test_batch_autotune_keeps_results():
    specs = [_spec(*m) for m in MIX]
    ref = [_alone(s, "fp32", 60) for s in specs]
    b = fdtd.Batch("fp32")
    ctxs = [fdtd.Context(s, precision="fp32") for s in specs]
    for c in ctxs:
        b.add(c)
    best = b.autotune(trial_steps=1)   # 20 shapes x 2 real steps each
    assert best["tile_z"] in (32, 64) and best["ms_per_step"] > 0
    b.step(60 - 40)
    b.sync()
    for i, c in enumerate(ctxs):
        _same(ref[i], _state(c), ("tuned", i))
    b.close()
    for c in ctxs:
        c.close()


def test_members_join_and_leave():
    specs = [_spec(*m) for m in MIX[:4]]
    ref = [_alone(s, "fp32", 50) for s in specs]
    b = fdtd.Batch("fp32")
    c = [fdtd.Context(s, precision="fp32") for s in specs]
    b.add(c[0]); b.add(c[1])
    b.step(20)
    b.add(c[2])                  # joins late: its own step count starts at 0
    b.step(20)
    b.remove(c[0])               # leaves after 40 steps, finishes alone
    c[0].step(10)
    b.add(c[3])
    b.step(10)
    b.remove(c[1])
    b.step(20)                   # c[2] reaches 50, c[3] 30
    b.sync()
    b.remove(c[3])
    c[3].step(20)
    c[0].sync(); c[3].sync()
    for i in range(4):
        _same(ref[i], _state(c[i]), ("join/leave", i))
    b.close()
    for x in c:
        x.close()


def test_batch_with_sponge_and_rules():
    specs = [_spec(*m, sponge={"width": 6}) for m in MIX[:3]]
    ref = [_alone(s, "fp32", 40) for s in specs]
    b = fdtd.Batch("fp32")
    ctxs = [fdtd.Context(s, precision="fp32") for s in specs]
    for x in ctxs:
        b.add(x)
    plain = fdtd.Context(_spec(*MIX[3]), precision="fp32")
    with pytest.raises(fdtd.FdtdError):
        b.add(plain)                       # no absorbing layer: rejected
    other = fdtd.Context(_spec(*MIX[4], sponge={"width": 6}), precision="fp64")
    with pytest.raises(fdtd.FdtdError):
        b.add(other)                       # other precision: rejected
    b.step(40)
    with pytest.raises(fdtd.FdtdError):
        ctxs[0].step(1)                    # stepped by its batch
    with pytest.raises(fdtd.FdtdError):
        ctxs[0].set("ux", np.zeros((specs[0].nodes[0] - 1,) + specs[0].nodes[1:]))
    b.sync()
    for i, x in enumerate(ctxs):
        _same(ref[i], _state(x), ("sponge", i))
    b.close()
    for x in ctxs + [plain, other]:
        x.close()


@pytest.mark.parametrize("precision", ["fp32", "fp32c"])
def test_two_step_context_through_a_batch(precision):
    # a two-step context (four level buffers) steps in pairs, joins a batch
    # (which drives its current and previous level), leaves it and steps in
    # pairs again: the same bits as stepping alone with the one-step sweep
    spec = _spec(*MIX[0])
    ref = _alone(spec, precision, 90)
    ctx = fdtd.Context(spec, precision=precision, kernel="tiled2")
    ctx.step(7)  # three pairs and one single step
    b = fdtd.Batch(precision)
    b.add(ctx)
    b.step(21)
    b.sync()
    b.remove(ctx)
    b.close()
    ctx.step(62)
    ctx.sync()
    out = _state(ctx)
    ctx.close()
    for f in ref:
        assert np.array_equal(out[f], ref[f]), f
