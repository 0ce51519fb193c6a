"""Benchmark: FDTD cell-updates/s on one MI355X, with roofline and CPU baseline.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload plain]

Workload "plain" (the headline): the plain FP32 sweep on the 512x512x256-cell
phononic crystal (513x513x257 nodes; square lattice a = 24 cells, r = 8
cells, +-0.5 cell jitter, seed 3; matrix rho 2700, lambda 5.1e10, mu 2.6e10;
tungsten-like inclusions rho 19300, lambda 2.0e11, mu 1.6e11; Courant 0.5;
Gaussian-modulated 2 MHz wall source for dx = 5e-5 m, i.e. 100 cycles per
unit time in the reference's unit-cell coordinates).  A step is one time step
over the whole mesh; cell-updates = nodes x steps (the reference counts
nodes, tests/test_speed.py:52).  Fields start at zero and the source drives
them; inputs are synthetic and resident in HBM before the timed region.

Prints one JSON line.  The roofline entry uses the sweep kernel's mean
duration from HIP events recorded on its own stream during the timed
steps, and the algorithmic bytes of one launch.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic HBM bytes per node per step of the sweep: read u^n (3) and
# u^{n-1} (3), write u^{n+1} (3), read the material index (1 byte)
BYTES_PER_NODE = {"fp32": 9 * 4 + 1, "fp32c": 18 * 4 + 1, "fp64": 9 * 8 + 1}


def crystal_spec(cells, a, r, seed, steps, kind="gauss", f0=2e6, dx_phys=5e-5, probe=None,
                 inclusion=None):
    """Crystal in the reference's unit-cell coordinates; the physical cell
    size dx_phys maps the source frequency to f0 * dx_phys per unit time."""
    from phonomena_amd import problem
    from phonomena_amd.simulation import lattice
    from phonomena_amd.simulation.base_solver import BaseSolver

    g, m = lattice.crystal(cells, a, r, seed, inclusion=inclusion or lattice.TUNGSTEN)
    b = BaseSolver(None)
    b.m = m
    b.cfg["wave"], b.cfg["wave_args"] = kind, {"f": f0 * dx_phys}
    wave = b.source_table(steps)
    spec = problem.mesh_spec(g, m, wave, {"probe": probe, "record_every": 10})
    return spec, g, m


def cpu_baseline(budget_s=20.0):
    """The oracle (NumPy restatement of the reference step, its dense C/P
    layout and expression structure) on a bounded sub-block of the same
    crystal recipe; 1 core (NumPy ufuncs are single-threaded)."""
    from oracle import reference_math as rm
    cells = (128, 128, 64)
    spec, g, m = crystal_spec(cells, 24, 8, 3, 64)
    C, P = m.C, m.P
    pb = rm.Problem(g.x, g.y, g.z, C, P, m.dt, spec.wave)
    s = rm.State(*pb.shape)
    rm.step(s, pb, 0)  # warm
    n = 0
    t0 = time.perf_counter()
    while n < 60 and (time.perf_counter() - t0) < budget_s:
        rm.step(s, pb, 1 + n)
        n += 1
    dt = time.perf_counter() - t0
    nodes = int(np.prod(pb.shape))
    return {"value": nodes * n / dt, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": "oracle (NumPy restatement of base_solver.py's step, dense C[...,6,6] layout) "
                      "on a %dx%dx%d-cell block of the same crystal recipe, %d steps, %.1f s; host has "
                      "%d CPUs (%d usable)" % (cells + (n, dt, os.cpu_count(), len(os.sched_getaffinity(0))))}


def traffic_from_profiles(workload, precision):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as fh:
            t = json.load(fh)
        return t.get("%s/%s" % (workload, precision))
    except (OSError, ValueError):
        return None


def bench_plain(args):
    from phonomena_amd import fdtd
    steps = args.warmup + args.steps
    spec, g, m = crystal_spec((512, 512, 256), 24, 8, 3, steps)
    nodes = spec.cells
    ctx = fdtd.Context(spec, precision=args.precision, kernel=args.kernel, tile_j=args.tile_j,
                       tile_x=args.tile_x, tile_z=args.tile_z)
    ctx.step(args.warmup)
    ctx.sync()
    t0 = time.perf_counter()
    sweep_ms, face_ms = ctx.profile(args.steps)  # issues K steps with HIP events, then syncs
    t1 = time.perf_counter()
    wall = t1 - t0
    value = nodes * args.steps / wall
    bpn = BYTES_PER_NODE[args.precision]
    achieved = bpn * nodes / (sweep_ms * 1e-3) / 1e9
    rec = {
        "metric": "cell-updates/s",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "f32", "fp32c": "f32 (hi+lo compensated)", "fp64": "f64"}[args.precision],
        "data": "synthetic",
        "config": {"workload": "plain %s sweep, 512x512x256-cell phononic crystal (513x513x257 nodes), "
                               "lattice a=24 r=8 cells, seed 3, gauss 2 MHz source" % args.precision,
                   "nodes": nodes, "kernel": args.kernel, "tile_j": args.tile_j, "tile_x": args.tile_x,
                   "tile_z": args.tile_z,
                   "fill_fraction": float(m.ids.mean())},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_from_profiles("plain", args.precision),
                     "kernel": "sweep", "bytes_per_node": bpn, "sweep_ms": sweep_ms, "face_ms": face_ms},
    }
    ctx.close()
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="plain", choices=["plain"])
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp32c", "fp64"])
    ap.add_argument("--kernel", default="auto", choices=["auto", "naive", "tiled", "strip"])
    ap.add_argument("--tile-j", type=int, default=0, help="tiled sweep rows per block (4..32)")
    ap.add_argument("--tile-z", type=int, default=0, help="tiled sweep lanes along z per row (32, 64)")
    ap.add_argument("--tile-x", type=int, default=0, help="tiled sweep x planes per block")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; multi-GPU sharding is a design "
              "note (slab decomposition along x), not implemented", file=sys.stderr)
        sys.exit(2)
    rec = bench_plain(args)
    rec["cpu_baseline"] = None if args.no_cpu else cpu_baseline()
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
