"""Benchmark: FDTD cell-updates/s on one MI355X, with roofline and CPU baseline.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload plain]

Workload "plain" (the headline): the plain FP32 sweep on the 512x512x256-cell
phononic crystal (513x513x257 nodes; square lattice a = 24 cells, r = 8
cells, +-0.5 cell jitter, seed 3; matrix rho 2700, lambda 5.1e10, mu 2.6e10;
tungsten-like inclusions rho 19300, lambda 2.0e11, mu 1.6e11; Courant 0.5;
Gaussian-modulated 2 MHz wall source for dx = 5e-5 m, i.e. 100 cycles per
unit time in the reference's unit-cell coordinates; 16-cell absorbing layer;
surface u_z probes on a 64x64 subgrid every 10 steps).  Workload "c2" is the
same recipe on the 256x256x128-cell crystal (a = 16, r = 5, seed 2).  A step
is one time step over the whole mesh; cell-updates = nodes x steps (the
reference counts nodes, tests/test_speed.py:52).  Fields start at zero and
the source drives them; inputs are synthetic and resident in HBM before the
timed region.

Prints one JSON line.  The roofline entry uses the sweep kernel's mean
duration from HIP events recorded on its own stream during the timed
steps, and the algorithmic bytes of one launch: 48 / 72 / 96 B per
cell-update for FP32 / FP32 compensated / FP64 (u^n and u^{n-1} read,
u^{n+1} written, three per-cell coefficients read; 12 values, plus the three
compensation terms read and written in the compensated mode).  The sweep
itself reads a 1-byte material index instead of three coefficients and
stores compensation terms in 16 bits, so its own minimum is 37 / 55 / 73 B
(``kernel_bytes_per_node``).
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
# algorithmic HBM bytes per cell-update (12 values: u^n, u^{n-1}, u^{n+1},
# 3 coefficients; + 3 compensation terms read and written when compensated)
BYTES_PER_NODE = {"fp32": 12 * 4, "fp32c": 18 * 4, "fp64": 12 * 8}
# what this sweep itself must move: 9 field values (FP32C: 4-byte hi and
# 2-byte lo each) and the 1-byte material index
KERNEL_BYTES_PER_NODE = {"fp32": 9 * 4 + 1, "fp32c": 9 * (4 + 2) + 1, "fp64": 9 * 8 + 1}
# the two-step pass per step: levels n, n-1 read and n+1, n+2 written once
# per two steps (12 field values) and the index once
PAIR_BYTES_PER_NODE = {"fp32": (12 * 4 + 1) / 2, "fp32c": (12 * (4 + 2) + 1) / 2}
SWEEP_KERNEL = {"tiled": "one-step tiled sweep (k_sweep_tiled)", "tiled2": "two-step sweep (k_sweep_fused)",
                "naive": "per-node sweep (k_sweep_naive)", "strip": "strip sweep (k_sweep_strip)"}
WORKLOADS = {
    "plain": dict(cells=(512, 512, 256), a=24, r=8, seed=3),
    "c2": dict(cells=(256, 256, 128), a=16, r=5, seed=2),
}
SPONGE = {"width": 16}
PROBE = {"count": 64}


def crystal_spec(cells, a, r, seed, steps, probe=None, sponge=None):
    """Crystal in the reference's unit-cell coordinates (phonomena_amd.sweep)."""
    from phonomena_amd import sweep
    return sweep.crystal_spec(cells, a, r, seed, steps, probe=probe, sponge=sponge)


def cpu_baseline(budget_s=15.0):
    """The oracle (NumPy restatement of the reference step, its dense C/P
    layout and expression structure) on a bounded sub-block of the same
    crystal recipe; 1 core (NumPy ufuncs are single-threaded)."""
    from oracle import reference_math as rm
    cells = (128, 128, 64)
    spec, g, m = crystal_spec(cells, 24, 8, 3, 160, sponge=SPONGE)
    C, P = m.C, m.P
    pb = rm.Problem(g.x, g.y, g.z, C, P, m.dt, spec.wave, sponge=spec.sponge)
    s = rm.State(*pb.shape)
    rm.step(s, pb, 0)  # warm
    n = 0
    t0 = time.perf_counter()
    while n < 150 and (time.perf_counter() - t0) < budget_s:
        rm.step(s, pb, 1 + n)
        n += 1
    dt = time.perf_counter() - t0
    nodes = int(np.prod(pb.shape))
    return {"value": nodes * n / dt, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": "oracle (NumPy restatement of base_solver.py's step, dense C[...,6,6] layout, "
                      "16-cell absorbing layer) on a %dx%dx%d-cell block of the same crystal recipe, "
                      "%d steps, %.1f s; host has "
                      "%d CPUs (%d usable)" % (cells + (n, dt, os.cpu_count(), len(os.sched_getaffinity(0))))}


def cpu_baseline_omp(budget_s=15.0):
    """The C restatement of the oracle (oracle/brute.c, the reference's step
    as per-cell loops, same bits as the NumPy one) built with OpenMP, on all
    the host cores this process may use: a stronger CPU baseline than the
    reference's own single-threaded NumPy cost model.  Sample: the C2-size
    crystal recipe (256x256x128 cells), as many steps as fit the budget."""
    from oracle import brute
    cells = (256, 256, 128)
    spec, g, m = crystal_spec(cells, 16, 5, 2, 400, sponge=SPONGE)
    threads = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
    state = None
    run = lambda n, st, t0: brute.run(g.x, g.y, g.z, m.ids, spec.materials, spec.dt, spec.mur, spec.wave, n,
                                      state=st, t0=t0, sponge=spec.sponge, omp=True)[0]
    state = run(1, None, 0)  # warm
    n, t = 0, time.perf_counter()
    while n < 399 and time.perf_counter() - t < budget_s:
        state = run(1, state, 1 + n)
        n += 1
    dt = time.perf_counter() - t
    return {"value": spec.cells * n / dt, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": "oracle/brute.c (per-cell C restatement of base_solver.py's step, bit-identical to the "
                      "NumPy oracle) built with OpenMP, %d threads, on the %dx%dx%d-cell crystal recipe "
                      "(a=16 r=5 seed 2, 16-cell absorbing layer), %d steps, %.1f s" % ((threads,) + cells + (n, dt))}


def stream_peaks(nodes):
    """Practical HBM bandwidth on this device for a plain copy and for the
    sweep's mix (6 field arrays read, 3 written), arrays of the mesh's size."""
    from phonomena_amd import fdtd
    n = nodes // 4 * 4
    return {"copy_1r1w": fdtd.stream_probe(1, 1, n, 10)[0], "mix_6r3w": fdtd.stream_probe(6, 3, n, 10)[0]}


def traffic_from_profiles(workload, precision, kernel):
    """HBM bytes per launch of the sweep kernel, from the FETCH_SIZE /
    WRITE_SIZE summaries under profiles/ (None when not measured)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as fh:
            t = json.load(fh)
        return t.get("%s/%s/%s" % (workload, precision, kernel))
    except (OSError, ValueError):
        return None


def bench_plain(args):
    from phonomena_amd import fdtd
    steps = args.warmup + args.steps
    wl = WORKLOADS[args.workload]
    spec, g, m = crystal_spec(wl["cells"], wl["a"], wl["r"], wl["seed"], steps,
                              sponge=None if args.no_sponge else SPONGE, probe=PROBE)
    nodes = spec.cells
    ctx = fdtd.Context(spec, precision=args.precision, kernel=args.kernel, tile_j=args.tile_j,
                       tile_x=args.tile_x, tile_z=args.tile_z)
    kernel, tile_j, tile_z = ctx.kernel_info()
    per_launch = 2 if kernel == "tiled2" else 1  # steps per sweep launch
    # warm-up through the timed path: its first call creates the context's
    # timing events, which the timed call then reuses
    if args.warmup > 0:
        ctx.profile(args.warmup)
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
        "config": {"workload": "plain %s sweep, %dx%dx%d-cell phononic crystal (%dx%dx%d nodes), "
                               "lattice a=%d r=%d cells, seed %d, gauss 2 MHz source, %s, 64x64 surface "
                               "probes every 10 steps" % ((args.precision,) + wl["cells"] + spec.nodes +
                                                          (wl["a"], wl["r"], wl["seed"],
                                                           "no absorbing layer" if args.no_sponge
                                                           else "16-cell absorbing layer")),
                   "nodes": nodes, "kernel": kernel, "kernel_requested": args.kernel, "tile_j": tile_j,
                   "tile_z": tile_z, "tile_x": args.tile_x,
                   "fill_fraction": float(m.ids.mean()),
                   # every spacing a power of two: the FP64 sweep multiplies by
                   # exact reciprocals instead of dividing (same bits)
                   "spacings_power_of_two": bool(all(np.all(np.frexp(a)[0] == 0.5) for a in spec.fd + spec.sd))},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     # the same algorithmic bytes over the whole timed step
                     # (faces, probes and launch gaps included)
                     "frac_whole_step": value * bpn / (HBM_PEAK_GBS * 1e9),
                     "traffic": traffic_from_profiles(args.workload, args.precision, kernel),
                     "kernel": SWEEP_KERNEL[kernel], "steps_per_launch": per_launch,
                     "launch_ms": sweep_ms * per_launch, "bytes_per_node": bpn,
                     "kernel_bytes_per_node": (PAIR_BYTES_PER_NODE if kernel == "tiled2"
                                               else KERNEL_BYTES_PER_NODE)[args.precision],
                     "sweep_ms": sweep_ms, "face_ms": face_ms},
    }
    ctx.close()
    return rec


def bench_other_precisions(args, precisions=("fp32c", "fp64")):
    """The plain workload in the other arithmetic modes (AUTO's sweep for
    each): FP32 compensated, and FP64, the reference's own arithmetic
    (base_solver.py:26), bit-identical to it."""
    out = {}
    for prec in precisions:
        a = argparse.Namespace(**dict(vars(args), precision=prec, kernel="auto", tile_j=0, tile_z=0, tile_x=0))
        r = bench_plain(a)
        out[prec] = {"value": r["value"], "unit": r["unit"], "ms_per_step": r["ms_per_step"], "dtype": r["dtype"],
                     "kernel": r["config"]["kernel"], "tile_j": r["config"]["tile_j"],
                     "tile_z": r["config"]["tile_z"],
                     "spacings_power_of_two": r["config"]["spacings_power_of_two"],
                     "roofline": {k: r["roofline"][k] for k in ("achieved", "peak", "unit", "frac", "frac_whole_step",
                                                                 "bytes_per_node", "kernel_bytes_per_node",
                                                                 "sweep_ms", "face_ms", "traffic")}}
    return out


def bench_setup(workload="plain", precision="fp32"):
    """Set-up of the plain workload's mesh: the material map built on the
    host (the reference's Grid.inclusionIndices + Material.setConstants
    loop, restated in phonomena_amd.simulation) and uploaded, against the
    map rasterized on the device from the inclusion list.  Each time covers
    lattice + material set-up and context creation (allocation, tables,
    map); the two maps are compared byte for byte."""
    from phonomena_amd import fdtd, problem
    from phonomena_amd.simulation import lattice
    wl = WORKLOADS[workload]
    out = {"workload": "%dx%dx%d-cell crystal, lattice a=%d r=%d, seed %d" % (wl["cells"] + (wl["a"], wl["r"], wl["seed"]))}
    maps = []
    for mode in ("host", "device"):
        t0 = time.perf_counter()
        g, m = lattice.crystal(wl["cells"], wl["a"], wl["r"], wl["seed"], host_ids=(mode == "host"))
        t1 = time.perf_counter()
        spec = problem.mesh_spec(g, m, np.zeros(2), {}, device_ids=(mode == "device"))
        ctx = fdtd.Context(spec, precision=precision, tune=False)  # map set-up only, no shape timing
        ctx.sync()
        t2 = time.perf_counter()
        maps.append(ctx.ids())
        ctx.close()
        out[mode] = {"lattice_material_s": t1 - t0, "context_s": t2 - t1, "total_s": t2 - t0}
    out["inclusions"] = int(len(g.targets))
    out["maps_equal"] = bool(np.array_equal(maps[0], maps[1]))
    out["speedup"] = out["host"]["total_s"] / out["device"]["total_s"]
    return out


def bench_batched(steps, precisions=("fp32c", "fp32"), n=64, seed=64, trial_steps=2):
    """The 64-crystal mix: the plain single-instance loop (one context at a
    time, pf_step) against one batch of all of them (autotuned, graph-
    replayed), same steps and precision, and the accuracy of the batched
    fields against an FP64 run of the same meshes (the GPU FP64 mode is
    bit-identical to the reference's arithmetic).  Set-up (meshing, upload)
    is not timed.  The first precision is the one reported as ``value``."""
    from phonomena_amd import fdtd, sweep
    jobs = sweep.mix(n, seed, steps)
    specs = [sweep.job_spec(j, probe={"count": 32}, every=10) for j in jobs]
    nodes = sum(sp.cells for sp in specs)
    # FP64 fields of the same meshes, as one FP64 batch
    c64 = [fdtd.Context(sp, precision="fp64", tune=False) for sp in specs]
    b64 = fdtd.Batch("fp64")
    for c in c64:
        b64.add(c)
    b64.step(steps)
    b64.sync()
    ref = [{f: c.get(f) for f in ("ux", "uy", "uz")} for c in c64]
    b64.close()
    for c in c64:
        c.close()
    out = {}
    for prec in precisions:
        # the plain loop as the library runs it (tile shapes measured at
        # creation, not timed, like the upload) and with the fixed shape
        # the library used before it measured shapes
        t_plain, t_fixed = 0.0, 0.0
        for sp in specs:
            for tune in (True, False):
                c = fdtd.Context(sp, precision=prec, tune=tune)
                t0 = time.perf_counter()
                c.step(steps)
                c.sync()
                if tune:
                    t_plain += time.perf_counter() - t0
                else:
                    t_fixed += time.perf_counter() - t0
                c.close()
        ctxs = [fdtd.Context(sp, precision=prec, tune=False) for sp in specs]
        b = fdtd.Batch(prec)
        for c in ctxs:
            b.add(c)
        t0 = time.perf_counter()
        tuned = b.autotune(trial_steps) if trial_steps > 0 else None
        b.step(steps - ctxs[0].steps_done)
        b.sync()
        t_batch = time.perf_counter() - t0
        worst = max(sweep.rel_l2(c.get(f), r[f]) for c, r in zip(ctxs, ref) for f in ("ux", "uy", "uz"))
        b.close()
        for c in ctxs:
            c.close()
        out[prec] = {"value": nodes * steps / t_batch, "plain_loop": nodes * steps / t_plain,
                     "speedup_vs_plain": t_plain / t_batch, "seconds_batched": t_batch,
                     "seconds_plain": t_plain, "rel_l2_max_vs_fp64": worst, "autotuned": tuned,
                     "plain_loop_fixed_tiles": nodes * steps / t_fixed,
                     "speedup_vs_plain_fixed_tiles": t_fixed / t_batch}
    head = out[precisions[0]]
    return {"workload": "%d-crystal mix, seed %d: 48x48x24..192x192x96 cells, lattice a 8-24 cells, r/a "
                        "0.15-0.45, void-like or tungsten-like, gauss 2 MHz source, 32x32 probes every 10 "
                        "steps, %d steps each" % (n, seed, steps),
            "meshes": n, "nodes_total": nodes, "steps": steps, "precision": precisions[0],
            "value": head["value"], "unit": "cell-updates/s", "plain_loop": head["plain_loop"],
            "speedup_vs_plain": head["speedup_vs_plain"], "rel_l2_max_vs_fp64": head["rel_l2_max_vs_fp64"],
            "speedup_vs_plain_fixed_tiles": head["speedup_vs_plain_fixed_tiles"],
            "by_precision": out}


def bench_precision(steps, workload="plain"):
    """The precision case: the crystal run for `steps` steps in FP32
    compensated (and plain FP32) against a GPU FP64 run of the same mesh and
    source (FP64 here is bit-identical to the reference's arithmetic);
    per-field rel-L2 at the final step and the drift of the FP64 energy
    audit, |E - E_64| / E_64."""
    from phonomena_amd import fdtd, sweep
    wl = WORKLOADS[workload]
    spec, _, _ = crystal_spec(wl["cells"], wl["a"], wl["r"], wl["seed"], steps, sponge=SPONGE, probe=PROBE)
    out = {"workload": "%dx%dx%d-cell crystal (a=%d r=%d seed %d, C2 recipe), %d steps" % (
        wl["cells"] + (wl["a"], wl["r"], wl["seed"], steps)), "steps": steps}
    ref = None
    for prec in ("fp64", "fp32c", "fp32"):
        c = fdtd.Context(spec, precision=prec)
        t0 = time.perf_counter()
        c.step(steps)
        c.sync()
        secs = time.perf_counter() - t0
        e = c.energy()
        f = {q: c.get(q) for q in ("ux", "uy", "uz")}
        c.close()
        if ref is None:
            ref = (f, e)
            out["fp64"] = {"energy": e, "seconds": secs}
            continue
        out[prec] = {"rel_l2": {q: sweep.rel_l2(f[q], ref[0][q]) for q in f},
                     "energy": e, "energy_drift_vs_fp64": abs(e - ref[1]) / abs(ref[1]) if ref[1] else None,
                     "seconds": secs}
        del f
    return out


def bench_physical(steps=1000):
    """The plain CPU case at its physical spacing: 64x64x32 cells at
    dx = 1e-4 m, homogeneous (rho 2700, lambda 5.1e10, mu 2.6e10), Courant
    0.5, 1 MHz sin source on the x = 0 wall, 8-cell absorbing layer, surface
    u_z every 10 steps, with the fixed free surface (the reference's
    precedence diverges at such a spacing).  FP32 and FP32C against FP64 of
    the same scheme, per field and probe frames; and FP64 against the
    reference's scheme run at unit spacing with the same source sequence
    (equal up to rounding: the fixed scheme does not depend on the spacing)."""
    from phonomena_amd import fdtd, sweep
    from phonomena_amd.simulation import lattice
    kw = dict(inclusion=lattice.MATRIX, probe=[0, 2, 32, 0, 2, 32], every=10, sponge={"width": 8},
              kind="sin", f0=1e6, dx_phys=1e-4)
    phys, _, _ = sweep.crystal_spec((64, 64, 32), 16, 5, 0, steps, d=1e-4, free_surface="fixed", **kw)
    unit, _, _ = sweep.crystal_spec((64, 64, 32), 16, 5, 0, steps, d=1.0, **kw)
    unit.wave = phys.wave

    def run(spec, prec):
        c = fdtd.Context(spec, precision=prec)
        c.step(steps)
        out = {q: c.get(q) for q in ("ux", "uy", "uz")}
        out["probes"] = c.probes()
        c.close()
        return out

    ref = run(phys, "fp64")
    rel = lambda got: {q: sweep.rel_l2(got[q], ref[q]) for q in ref}
    return {"workload": "64x64x32 cells at dx = 1e-4 m, homogeneous, Courant 0.5, 1 MHz sin, 8-cell layer, "
                        "%d steps, fixed free surface" % steps,
            "fp64_vs_unit_spacing_reference": rel(run(unit, "fp64")),
            "fp32_vs_fp64": rel(run(phys, "fp32")), "fp32c_vs_fp64": rel(run(phys, "fp32c"))}


def bench_served(n_jobs, capacity_jobs_per_s, precision="fp32c", load=0.7, seed=5, chunk=200, hbm_budget=200e9):
    """Served design-sweep stream: Poisson arrivals at `load` x the batched
    capacity (jobs/s), grids U[32^3 .. 256x256x128] cells, U{500..8000}
    steps, lattice parameters as in the batched mix, an HBM budget for the
    resident jobs (200 GB: never binding for this stream; a small one makes
    arrivals wait for memory), 32x32 probe frames streamed back every 10
    steps."""
    from phonomena_amd import serve
    jobs = serve.stream_jobs(n_jobs, seed)
    rate = load * capacity_jobs_per_s
    arrivals = serve.poisson_arrivals(n_jobs, rate, seed)
    srv = serve.Server(precision=precision, hbm_budget=hbm_budget, chunk=chunk, probe={"count": 32},
                       record_every=10)
    results, stats = srv.run(jobs, arrivals)
    stats.update({"workload": "served stream of %d jobs, seed %d: Poisson arrivals at %.2f jobs/s (%.1f x the "
                              "batched capacity), 32^3..256x256x128-cell crystals, 500-8000 steps, %s, "
                              "%.3g GB HBM budget, 32x32 probes every 10 steps streamed back"
                              % (n_jobs, seed, rate, load, precision, hbm_budget / 1e9),
                  "arrival_rate_jobs_per_s": rate})
    waits = np.array([r.admitted - r.arrival for r in results if r is not None and r.admitted is not None])
    if waits.size:  # arrival to admission (memory and chunk boundaries)
        stats["admission_wait_s"] = {"mean": float(waits.mean()), "p99": float(np.percentile(waits, 99)),
                                     "max": float(waits.max())}
    return stats


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="plain", choices=sorted(WORKLOADS))
    ap.add_argument("--no-sponge", action="store_true", help="Mur faces only, no absorbing layer")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp32c", "fp64"])
    ap.add_argument("--kernel", default="auto", choices=["auto", "naive", "tiled", "strip", "tiled2"])
    ap.add_argument("--tile-j", type=int, default=0, help="tiled sweep rows per block (4..32)")
    ap.add_argument("--tile-z", type=int, default=0, help="tiled sweep lanes along z per row (32, 64)")
    ap.add_argument("--tile-x", type=int, default=0, help="tiled sweep x planes per block")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-stream", action="store_true", help="skip the streaming-bandwidth probe")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched 64-crystal leg")
    ap.add_argument("--batch-steps", type=int, default=2000, help="steps per crystal of the batched leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the served-stream leg")
    ap.add_argument("--precision-steps", type=int, default=10000,
                    help="steps of the FP32C/FP32 vs FP64 precision case (0 = skip)")
    ap.add_argument("--serve-jobs", type=int, default=48, help="jobs in the served stream")
    ap.add_argument("--no-setup", action="store_true", help="skip the mesh set-up (material map) leg")
    ap.add_argument("--no-physical", action="store_true",
                    help="skip the physical-spacing leg (plain CPU case at dx = 1e-4 m, fixed free surface)")
    ap.add_argument("--no-precisions", action="store_true",
                    help="skip the FP32C / FP64 legs of the plain workload")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the plain leg (for profiling the other legs alone; the line then lacks value)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; multi-GPU sharding is a design "
              "note (slab decomposition along x), not implemented", file=sys.stderr)
        sys.exit(2)
    rec = {"metric": "cell-updates/s", "value": None, "config": {"nodes": 0}, "roofline": {}} if args.no_plain \
        else bench_plain(args)
    if not args.no_plain and not args.no_precisions and args.precision == "fp32":
        rec["other_precisions"] = bench_other_precisions(args)
    if not args.no_stream and not args.no_plain:
        rec["roofline"]["stream_peak_gbs"] = stream_peaks(rec["config"]["nodes"])
    if not args.no_batch:
        rec["batched"] = bench_batched(args.batch_steps)
        if not args.no_serve:
            from phonomena_amd import serve, sweep
            # capacity in jobs/s: batched throughput over the mean work of a served job
            jobs = serve.stream_jobs(256, 5)
            mean_work = np.mean([np.prod([c + 1 for c in j["cells"]]) * j["steps"] for j in jobs])
            rec["served"] = bench_served(args.serve_jobs, rec["batched"]["value"] / mean_work)
            # the same stream with a budget of about half its resident peak:
            # admission by memory, arrivals queue for it
            rec["served_budget_bound"] = bench_served(args.serve_jobs, rec["batched"]["value"] / mean_work,
                                                      hbm_budget=0.5e9)
    if args.precision_steps > 0:
        rec["precision_case"] = bench_precision(args.precision_steps)
    if not args.no_setup:
        rec["setup"] = bench_setup(args.workload)
    if not args.no_physical:
        rec["physical_spacing"] = bench_physical()
    rec["cpu_baseline"] = None if args.no_cpu else cpu_baseline()
    if not args.no_cpu:
        rec["cpu_baseline_omp"] = cpu_baseline_omp()
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
