"""Per-step recording, host side (no GPU): the AsyncRecorder's slot cycle
(snapshot -> wait -> write -> release, in step order, with back-pressure)
against a stand-in context, and the .npy sink's schema round trip."""
import json

import numpy as np
import pytest

from phonomena_amd.simulation import recording


class FakeCtx:
    """Snapshot slots with the library's state rules; frame = step count."""

    def __init__(self, nodes):
        self.nodes = nodes
        self.steps = 0
        self.shapes = recording.frame_shapes(nodes)
        self.state = None
        self.taken = {}
        self.calls = []

    def snapshot_open(self, slots):
        assert self.state is None
        self.state = [0] * slots

    def step(self):
        self.steps += 1

    def snapshot(self, slot):
        assert self.state[slot] == 0, "slot in use"
        self.state[slot] = 1
        self.taken[slot] = self.steps
        self.calls.append(("snap", slot, self.steps))

    def snapshot_wait(self, slot):
        assert self.state[slot] in (1, 2)
        self.state[slot] = 2
        v = float(self.taken[slot])
        return {k: np.full(s, v) for k, s in self.shapes.items()}, self.taken[slot]

    def snapshot_release(self, slot):
        assert self.state[slot] == 2
        self.state[slot] = 0


@pytest.mark.parametrize("slots", [1, 2, 4])
def test_frames_land_in_step_order(tmp_path, slots):
    nodes, steps = (5, 4, 3), 23
    ctx = FakeCtx(nodes)
    sink = recording.NpySink(str(tmp_path / "r"), nodes, steps, {"dt": 0.5, "steps": steps})
    rec = recording.AsyncRecorder(ctx, sink, slots=slots)
    for _ in range(steps):
        ctx.step()
        rec.record()
    rec.close()
    data, attrs = recording.load_npy(str(tmp_path / "r"))
    for name, shape in recording.frame_shapes(nodes).items():
        assert data[name].shape == shape + (steps,)
        for tt in range(steps):
            assert np.all(data[name][..., tt] == tt + 1), (name, tt)
    assert attrs["dt"] == 0.5
    assert all(s == 0 for s in ctx.state)


def test_writer_errors_surface(tmp_path):
    class Bad:
        def put(self, frame, tt):
            raise OSError("disk full")

        def close(self):
            pass

    ctx = FakeCtx((4, 4, 4))
    rec = recording.AsyncRecorder(ctx, Bad(), slots=2)
    with pytest.raises(OSError, match="disk full"):
        for _ in range(10):
            ctx.step()
            rec.record()
        rec.close()


def test_schema_attrs_match_writer_keys():
    class G:
        pass

    g = G()
    for k in recording.ATTR_KEYS:
        setattr(g, k, np.arange(3.0))

    class M:
        dt = 0.1
        primary = {"name": "a"}
        secondary = {"name": "b"}

    a = recording.schema_attrs(g, M(), 7, {"wave": "sin"})
    assert set(a) == set(recording.ATTR_KEYS) | {"steps", "dt", "prim_material", "sec_material", "solver_cfg"}
    assert json.loads(a["solver_cfg"]) == {"wave": "sin"}
