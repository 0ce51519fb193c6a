"""Per-step recording from device snapshots: frames equal the fields read
synchronously after every step, and the solver's 'npy' write mode
reproduces the reference's recorded frames (FP64, bit for bit)."""
import json

import numpy as np
import pytest

from helpers import golden, spec_from, uniform_crystal
from test_gpu_raster import _fixture_settings
from phonomena_amd import common, fdtd
from phonomena_amd.simulation import recording

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,slots", [("fp32c", 1), ("fp32", 3), ("fp64", 4)])
def test_frames_equal_synchronous_reads(tmp_path, precision, slots):
    x, y, z, ids, rows, dt, wave = uniform_crystal(
        (30, 22, 18), 1.0, incl=(19300.0, 2.0e11, 1.6e11, 8, 2), steps=40, seed=4)
    spec = spec_from(x, y, z, ids, rows, dt, wave)
    steps = 40
    ref = fdtd.Context(spec, precision=precision)
    expect = []
    for _ in range(steps):
        ref.step(1)
        expect.append({c: ref.get(c) for c in ("ux", "uy", "uz")})
    ref.close()
    ctx = fdtd.Context(spec, precision=precision)
    sink = recording.NpySink(str(tmp_path / "r"), spec.nodes, steps, {"dt": float(dt)})
    rec = recording.AsyncRecorder(ctx, sink, slots=slots)
    for tt in range(steps):
        ctx.step(1)
        rec.record()
    rec.close()
    data, _ = recording.load_npy(str(tmp_path / "r"))
    for tt in range(steps):
        for c in ("ux", "uy", "uz"):
            assert np.array_equal(data[c][..., tt], expect[tt][c]), (tt, c)
    # the context steps on after its snapshots
    ctx.close()


def test_solver_npy_mode_reproduces_reference_frames(tmp_path):
    d = golden("default_json_100")
    path = tmp_path / "s.json"
    settings = _fixture_settings(d)
    settings["simulation"]["cfg"]["write_mode"] = "npy"
    settings["simulation"]["cfg"]["precision"] = "fp64"
    path.write_text(json.dumps(settings))
    solvers = common.findSolvers()
    cfg, g, m = common.loadSettings(str(path))
    s = solvers["hip"]
    s.cfg.update(cfg["simulation"]["cfg"])
    s.file = str(tmp_path / "out")
    s.init(g, m, int(d["steps"]))
    s.run()
    s.close()
    data, attrs = recording.load_npy(s.file + ".d")
    steps = int(d["steps"])
    assert attrs["steps"] == steps and attrs["dt"] == d["dt"][()]
    assert np.array_equal(np.asarray(attrs["x"]), d["x"])
    # the reference's surface frames are its writer frames 9, 19, ..., 99
    surf = np.transpose(data["uz"][:, :, 0, 9::10], (2, 0, 1))
    assert np.array_equal(surf, d["uz_surface"])
    for c in ("ux", "uy", "uz"):
        assert np.array_equal(data[c][..., steps - 1], d[c]), c
    assert data["density"].shape == d["ids"].shape
