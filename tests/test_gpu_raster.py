"""The device rasterizer (material map from the inclusion list, pf_create
with ids NULL) against the reference's masks and material maps, and fields
stepped from a device-rasterized map against the reference trajectories."""
import json
import os

import numpy as np
import pytest

from helpers import GOLDEN, golden
from oracle import reference_math as rm
from test_raster_host import CASES, _crystal_grid, _material_map, raster_map
from phonomena_amd import common, fdtd, problem, sweep
from phonomena_amd.simulation import grid, material

pytestmark = pytest.mark.gpu


def _device_ids(g, m, precision="fp64"):
    wave = np.zeros(2)
    spec = problem.mesh_spec(g, m, wave, {}, device_ids=True)
    assert spec.ids is None
    ctx = fdtd.Context(spec, precision=precision)
    ids = ctx.ids()
    ctx.close()
    return ids


def _materials(g):
    m = material.Material()
    m.init(g, {"a": material.isotropic("a", 2700.0, 5e10, 2.6e10),
               "b": material.isotropic("b", 19300.0, 2e11, 1.6e11)})
    m.setPrimary("a")
    m.setSecondary("b")
    m.update(rasterize=False)
    return m


def test_reference_mesh_cases():
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
        assert np.array_equal(_device_ids(g, _materials(g)), d["c%d_mask" % n]), n


@pytest.mark.parametrize("case", range(len(CASES)))
def test_partial_depths_ties_and_mesher(case):
    g = _crystal_grid(*CASES[case])
    assert np.array_equal(_device_ids(g, _materials(g), "fp32"), _material_map(g))


def _fixture_settings(d):
    props = {"M0": {"name": "m0", "p": float(d["p_primary"]), "c_Pa": d["C_primary"].tolist()},
             "M1": {"name": "m1", "p": float(d["p_secondary"]), "c_Pa": d["C_secondary"].tolist()}}
    incl = [{"x": float(t[0]), "y": float(t[1]), "z": float(t[2]), "r": float(t[3])} for t in d["inclusions"]]
    cfg = json.loads(str(d["cfg"]))
    return {"grid": json.loads(str(d["grid_cfg"])), "inclusions": incl,
            "material": {"primary": "M0", "secondary": "M1", "properties": props},
            "simulation": {"courant": float(d["courant"]), "steps": int(d["steps"]),
                           "solver": "hip", "cfg": dict(cfg, write_mode="off")}}


@pytest.mark.parametrize("name", ["default_json_100", "nonuniform_json_60", "fine_json_20"])
def test_reference_trajectories_from_device_map(tmp_path, name):
    # settings -> grid, material, inclusion list; the map is rasterized on
    # the device and the FP64 run must still equal the reference bit for bit
    d = golden(name)
    path = tmp_path / "s.json"
    path.write_text(json.dumps(_fixture_settings(d)))
    common.findSolvers()
    cfg, g, m = common.loadSettings(str(path))
    scfg = json.loads(str(d["cfg"]))
    wave = rm.source_table(scfg["wave"], int(d["steps"]), d["dt"][()], scfg["wave_args"])
    spec = problem.mesh_spec(g, m, wave, {}, device_ids=True)
    assert spec.dt == d["dt"][()]
    ctx = fdtd.Context(spec, precision="fp64")
    assert np.array_equal(ctx.ids(), d["ids"])
    ctx.step(int(d["steps"]))
    for f in ("ux", "uy", "uz", "ux_old", "uy_old", "uz_old"):
        assert np.array_equal(ctx.get(f), d[f]), f
    ctx.close()


def test_crystal_mix_maps():
    for job in sweep.mix(n=64, seed=64, steps=2)[:12]:
        g, m = sweep.lattice.crystal(job["cells"], job["a"], job["r"], job["seed"],
                                     inclusion=sweep.INCLUSIONS[job["inclusion"]])
        spec = problem.mesh_spec(g, m, np.zeros(2), {}, device_ids=True)
        x, y, tg = spec.raster
        assert np.array_equal(raster_map(x, y, g.z.size, tg), m.ids)
        ctx = fdtd.Context(spec, precision="fp32")
        assert np.array_equal(ctx.ids(), m.ids), job
        ctx.close()


def test_fields_same_from_either_map():
    job = sweep.mix(n=4, seed=3, steps=60)[1]
    a = sweep.job_spec(job, probe={"count": 8}, every=5)
    b = sweep.job_spec(job, probe={"count": 8}, every=5, device_ids=True)
    out = []
    for spec in (a, b):
        ctx = fdtd.Context(spec, precision="fp32c")
        ctx.step(60)
        out.append([ctx.get(f) for f in ("ux", "uy", "uz", "uz_old")] + [ctx.probes()])
        ctx.close()
    for u, v in zip(*out):
        assert np.array_equal(u, v)


def test_rejects_bad_inclusion_lists():
    g = _crystal_grid(*CASES[0])
    m = _materials(g)
    spec = problem.mesh_spec(g, m, np.zeros(2), {}, device_ids=True)
    x, y, tg = spec.raster
    bad = tg.copy()
    bad[0, 3] = g.z.size + 1
    spec.raster = (x, y, bad)
    with pytest.raises(fdtd.FdtdError, match="z nodes"):
        fdtd.Context(spec, precision="fp32")
