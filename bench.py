#!/usr/bin/env python
"""Throughput of the spherical vertex model on one MI355X.

Workload (default): 256 independent spheres in one batch, N log-uniform in
[200, 20000], p0 uniform in [3.5, 4.0], v0 uniform in [0.01, 0.5],
Dr in {0.1, 1, 10} (tau = 1/Dr), dt = 0.01, mu = 1, K_A = K_P = 1, A0 = 1,
FP64; active Brownian dynamics with T1 topology maintenance on every step;
steps replayed from a captured HIP graph.  Every sphere is first stabilised
as the reference driver does (200 Brownian T=0 steps with (A0, P0) = (1, 1),
untimed), then initialised by 1000 untimed active steps at the target
parameters (the driver's initialisation phase).  A "step" is one timestep of every
sphere; the metric is cell-steps/s = sum_N * steps / wall time, with the
state resident in HBM before the timed region starts.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-steps/s (batched spheres)"   # the FIRE learned-vs-plain part is the "fire" object
PEAK_HBM_GBS = 8000.0
# Algorithmic HBM bytes per cell per step of the vertex pass (k_vertex_active):
# two vertices per cell, each reading its position (24 B) and director (24 B),
# writing its new position (24 B) and rotated director (24 B) and reading its
# vertex-cell and vertex-vertex lists (2 x 12 B): 2 x 120 = 240 B.
VERTEX_BYTES_PER_CELL = 240.0
STEP_BYTES_PER_CELL = 244.0   # whole active step (positions, directors r/w, topology)


def batch_params(n_sys, seed=13377, nmin=200, nmax=20000):
    """Per-sphere (N, p0, v0, Dr), drawn in sphere order from one Philox generator."""
    g = np.random.Generator(np.random.Philox(seed))
    out = []
    for _ in range(n_sys):
        u = g.random(4)
        n = int(round(np.exp(np.log(nmin) + u[0] * (np.log(nmax) - np.log(nmin)))))
        p0 = 3.5 + 0.5 * u[1]
        v0 = 0.01 + 0.49 * u[2]
        dr = [0.1, 1.0, 10.0][min(2, int(u[3] * 3))]
        out.append((n, p0, v0, dr))
    return out


def stabilise(ctx, params, steps, eager=False):
    """The reference driver's stabilisation (sphericalVertex.cpp:100-110):
    Brownian dynamics at T=0 with (A0, P0) = (1, 1), then the target (1, p0).
    From raw random Voronoi starts the dynamics otherwise squeeze small cells
    until they invert (the model has no T2 move)."""
    for i in range(len(params)):
        ctx.set_preferred_parameters(i, 1.0, 1.0)
        ctx.set_brownian(i, dt=0.01, mu=1.0, T=0.0)
    if eager:
        ctx.brownian_steps(steps)
    else:
        ctx.run_graph(0, 10, steps // 10)
    ctx.synchronize()
    for i, (n, p0, v0, dr) in enumerate(params):
        ctx.set_preferred_parameters(i, 1.0, p0)


def health(ctx, params, base, after):
    """Geometry invariants after the run (the angle-sum areas must tile each
    sphere) and the T1 counters of the timed steps: `base` and `after` are
    snapshots of counters and diagnostics taken right before and right after
    them."""
    ctx.compute_geometry()
    dev = 0.0
    for i, (n, *_rest) in enumerate(params):
        dev = max(dev, abs(ctx.area_perimeter(i)[:, 0].sum() / n - 1.0))
    c = after[0] - base[0]
    dg = after[1]
    r = (dg - base[1]).sum(axis=0)
    return {"t1_sequential_reasons": {"flippable_list_overflow": int(r[0]), "group_too_large": int(r[1]),
                                      "group_out_of_room": int(r[2]), "blocked_list_overflow": int(r[3])},
            "t1_candidates": {"flippable": int(r[4]), "blocked": int(r[5])},
            "t1_dry_runs": {"group_members": int(r[6]), "group_attempts": int(r[7]), "group_visits": int(r[8]),
                            "group_longest_walk": int(dg[:, 9].max()), "single_visits": int(r[10]),
                            "single_longest_walk": int(dg[:, 11].max())},"max_rel_area_sum_error": dev, "t1_flips": int(c[:, 0].sum()), "fan_fallbacks": int(c[:, 1].sum()),
            "nan_contributions": int(c[:, 2].sum()), "t1_skipped_no_fourth_cell": int(c[:, 3].sum()),
            "t1_sequential_sphere_steps": int(c[:, 4].sum()), "t1_group_sphere_steps": int(c[:, 5].sum())}


def build_batch(svm, params):
    """Seeded cell points (spatially ordered labels), their Delaunay topology
    built on the GPU, seeded directors.  Returns the context and the seconds
    spent in the GPU triangulation + topology construction."""
    systems = [dict(n_cells=n, a0=1.0, p0=p0, kr=1.0, t1_threshold=0.1, seed=1137 + i)
               for i, (n, p0, v0, dr) in enumerate(params)]
    ctx = svm.Context(systems)
    t_dl = 0.0
    for i, (n, p0, v0, dr) in enumerate(params):
        pts = svm.generate_cell_points(n, 1137 + i)
        pts = pts[svm.spatial_order(pts)]   # spatially coherent labels
        t0 = time.perf_counter()
        ctx.build_from_cell_points(i, pts)
        t_dl += time.perf_counter() - t0
        ctx.set_directors(i, svm.generate_directors(ctx.positions(i), 1137 + i))
        ctx.set_active(i, dt=0.01, mu=1.0, v0=v0, tau=1.0 / dr)
    return ctx, t_dl


def host_cpu():
    """The host's CPU model and the logical CPUs this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"model": model, "logical_cpus_usable": usable, "logical_cpus_total": os.cpu_count()}


def cpu_baseline(ctx, params, budget_s=12.0, threads=None):
    """The CPU oracle (a restatement of the reference's CPU path, which is
    single-threaded) on the host cores: independent spheres of the same batch,
    from the same state, spread over `threads` worker threads (each advances
    its own spheres; ctypes releases the GIL), then the same on one thread."""
    import ctypes as C
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import svmtest
    L = svmtest.oracle_lib()
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
    if threads is None:
        threads = max(1, min(16, os.cpu_count() or 1))

    def make(i):
        n, p0, v0, dr = params[i]
        vn, vcn, ring, cnn = ctx.get_topology(i, 32)
        pos = np.ascontiguousarray(ctx.positions(i))
        h = C.c_void_p(L.or_create(C.c_int(n), C.c_int(2 * n - 4), C.c_double(np.sqrt(n / (4 * np.pi))), dp(pos),
                                   ip(vn), ip(vcn), ip(np.ascontiguousarray(ring)), C.c_int(32), ip(cnn)))
        L.or_set_preferences(h, dp(np.tile(np.array([1.0, p0]), n)))
        L.or_set_directors(h, dp(np.ascontiguousarray(ctx.directors(i))))
        return (h, n, v0, dr, i)

    def run(models, budget):
        steps, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < budget:
            for h, n, v0, dr, i in models:
                L.or_active_steps(h, C.c_int(10), C.c_double(0.01), C.c_double(1.0), C.c_double(v0),
                                  C.c_double(1.0 / dr), C.c_uint64(1137 + i), C.c_uint64(steps))
            steps += 10
        return sum(m[1] for m in models) * steps, time.perf_counter() - t0

    # one group of spheres (>= 20000 cells) per thread, in batch order
    groups, cur, cells, i = [], [], 0, 0
    while len(groups) < threads and i < len(params):
        cur.append(make(i))
        cells += params[i][0]
        i += 1
        if cells >= 20000:
            groups.append(cur)
            cur, cells = [], 0
    if cur:
        groups.append(cur)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(len(groups)) as ex:
        res = list(ex.map(lambda g: run(g, budget_s), groups))
    wall = time.perf_counter() - t0
    multi = sum(r[0] for r in res) / wall
    single_work, single_t = run(groups[0], budget_s / 4)
    for g in groups:
        for h, *_ in g:
            L.or_destroy(h)
    ncells = sum(m[1] for g in groups for m in g)
    return dict(value=multi, unit="cell-steps/s", cores=len(groups), kind="port", host_cpu=host_cpu(),
                single_thread_value=single_work / single_t,
                sample=f"{sum(len(g) for g in groups)} spheres of the batch ({ncells} cells) in {len(groups)} "
                       f"threads, each advancing its own spheres with the oracle's active step for "
                       f"{budget_s:.0f} s; single-thread figure: the first thread's spheres alone")


SCHEDULES = os.path.join(ROOT, "curvedspacevertexmodel_amd", "schedules.txt")


def schedule_table(path):
    """The persisted schedule table's lines (svm_tuning_save format), or None."""
    try:
        return [l.strip() for l in open(path) if l.strip()]
    except OSError:
        return None


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic_batched-active-256.txt")
FP64_FILE = os.path.join(ROOT, "profiles", "pmc_fp64_batched-active-256.txt")
PEAK_FP64_TFLOPS = 78.6


def profile_figure(path, key, kernel):
    """A per-launch figure of `kernel` on this workload from a committed
    rocprofv3 summary line `key kernel value`; None if there is none."""
    try:
        for line in open(path):
            f = line.split()
            if len(f) == 3 and f[0] == key and f[1] == kernel:
                return float(f[2])
    except OSError:
        pass
    return None


def measured_traffic(kernel):
    """HBM bytes per launch (rocprofv3 FETCH_SIZE / WRITE_SIZE, calibrated
    against a plain copy of known size)."""
    return profile_figure(TRAFFIC_FILE, "traffic_bytes_per_launch", kernel)


def fp64_roofline(kernel, launch_ms):
    """The FP64 side of the dominant kernel, which is VALU-bound in the
    parity arithmetic: its FP64 FLOP per launch (rocprofv3
    SQ_INSTS_VALU_{ADD,MUL,FMA}_F64 on this workload, committed summary) over
    the launch time measured here, against the FP64 vector peak."""
    flop = profile_figure(FP64_FILE, "fp64_flop_per_launch", kernel)
    if flop is None or not launch_ms > 0:
        return None
    tf = flop / (launch_ms * 1e-3) / 1e12
    return {"flop_per_launch": flop, "achieved": tf, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "frac": tf / PEAK_FP64_TFLOPS}


def copy_probe(svm, n_bytes=1 << 28):
    """A plain streaming copy of n_bytes (read n_bytes, write n_bytes), three
    launches: the yardstick for the FETCH_SIZE / WRITE_SIZE counters."""
    ms = svm.copy_probe(n_bytes, 3)
    print(f"copy probe: 3 x k_copy_probe of {n_bytes} B, {ms:.4f} ms each, "
          f"{2 * n_bytes / ms / 1e6:.0f} GB/s", flush=True)


def phase_probe(svm):
    """What a dependency between two phases of a step costs: a kernel boundary
    in a HIP graph (the library's form) against a grid barrier inside one
    persistent cooperative launch, 200 phases each, with 2 and 4 resident
    workgroups per compute unit (the vertex pass's and the geometry pass's
    residency) and 1 MiB or 64 MiB written per phase."""
    out = []
    for wg in (512, 1024):
        for mb in (1, 64):
            launch_us, barrier_us, errors = svm.phase_boundary_probe(wg, 200, mb << 20)
            row = {"workgroups": wg, "mib_per_phase": mb, "us_per_kernel_boundary": launch_us,
                   "us_per_grid_barrier": barrier_us, "wrong_values": errors}
            print(f"phase probe: {row}", flush=True)
            out.append(row)
    return out


def single_sphere(svm, n=262144, steps=400):
    """The single-system workload: one sphere of n seeded cells (p0 3.8),
    stabilised, then Brownian dynamics (dt 0.01, mu 1, T 0.001, the reference
    driver's temperature) replayed from HIP graphs with the FP64 parity forces
    and with the FP32 forces.  Also the GPU triangulation + topology build of
    the sphere and one more full retriangulation of its cell points (the
    per-move work of a Voronoi-type model)."""
    ctx = svm.Context([dict(n_cells=n, a0=1.0, p0=3.8, kr=1.0, t1_threshold=0.1, seed=1137)])
    pts = svm.generate_cell_points(n, 1137)
    pts = pts[svm.spatial_order(pts)]
    t0 = time.perf_counter()
    ctx.build_from_cell_points(0, pts)
    t_build = time.perf_counter() - t0
    svm.delaunay_triangles(pts)
    t0 = time.perf_counter()
    svm.delaunay_triangles(pts)
    t_dl = time.perf_counter() - t0
    ctx.set_directors(0, svm.generate_directors(ctx.positions(0), 1137))
    ctx.set_preferred_parameters(0, 1.0, 1.0)
    ctx.set_brownian(0, dt=0.01, mu=1.0, T=0.0)
    ctx.run_graph(0, 10, 10)
    ctx.set_preferred_parameters(0, 1.0, 3.8)
    ctx.set_brownian(0, dt=0.01, mu=1.0, T=0.001)
    rate = {}
    for prec, name in ((0, "fp64"), (1, "fp32")):
        ctx.set_force_precision(prec)
        ctx.run_graph(0, 10, 5)
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.run_graph(0, 10, steps // 10)
        ctx.synchronize()
        rate[name] = n * (steps // 10 * 10) / (time.perf_counter() - t0)
    ctx.set_force_precision(0)
    del ctx
    return {"workload": "single-262144", "cells": n, "updater": "brownian", "dt": 0.01, "T": 0.001, "p0": 3.8,
            "steps": steps // 10 * 10, "cell_steps_per_s": rate, "topology_build_s": t_build,
            "retriangulation_ms": 1e3 * t_dl, "voronoi_model": voronoi_model(svm, pts)}


def voronoi_model(svm, pts, vsteps=20):
    """The cell points as a Voronoi model: every step moves the generators
    (FP64 forces through the circumcentres, Brownian, T 0.001) and
    retriangulates them on the GPU."""
    n = len(pts)
    vctx = svm.Context([dict(n_cells=n, a0=1.0, p0=3.8, kr=1.0, t1_threshold=0.1, seed=1137)])
    vctx.build_voronoi_from_cell_points(0, pts)
    vctx.set_brownian(0, dt=0.01, mu=1.0, T=0.001)
    vctx.voronoi_brownian_steps(2)
    vctx.synchronize()
    t0 = time.perf_counter()
    vctx.voronoi_brownian_steps(vsteps)
    vctx.synchronize()
    t_vor = (time.perf_counter() - t0) / vsteps
    return {"steps": vsteps, "ms_per_step": 1e3 * t_vor, "cell_steps_per_s": n / t_vor,
            "note": "FP64 forces on the generators, Brownian move, GPU retriangulation every step"}


def large_batch(svm, steps=2000, schedules=None, fire_iterations=200):
    """The large batched workload: 32 spheres, N log-uniform in [50,000,
    500,000] (one Philox generator, seed 13377), p0 uniform in [3.7, 3.85],
    FP32 forces on FP64 positions with FP64 predicates.  GPU Delaunay set-up,
    the reference driver's stabilisation (100 Brownian steps at (A0, P0) =
    (1, 1)), the learned warm start, `fire_iterations` FIRE iterations, then
    `steps` active steps (v0 0.05) replayed from HIP graphs over the stored
    or autotuned lanes, timed after 40 warm-up steps.
    FIRE to rms force < 1e-10 is not reachable on these spheres: with the
    reference driver's FIRE parameters it diverges between 500 and 1,000
    iterations in every force precision, and on a 50,000-cell sphere the FP64
    parity run equals the oracle's FIRE bit for bit through the divergence
    (tests/test_gpu_large.py), so the FIRE phase stops where it is still
    descending."""
    from curvedspacevertexmodel_amd import warmstart as ws
    params = ws.batch_params(32, seed=13377, nmin=50000, nmax=500000, p0_range=(3.7, 3.85))
    total = int(sum(n for n, _ in params))
    ctx = svm.Context([dict(n_cells=n, a0=1.0, p0=p0, kr=1.0, t1_threshold=0.1, seed=1137 + i)
                       for i, (n, p0) in enumerate(params)])
    t0 = time.perf_counter()
    for i, (n, _) in enumerate(params):
        pts = svm.generate_cell_points(n, 1137 + i)
        ctx.build_from_cell_points(i, pts[svm.spatial_order(pts)])
        ctx.set_directors(i, svm.generate_directors(ctx.positions(i), 1137 + i))
    t_topo = time.perf_counter() - t0
    print(f"large batch: {len(params)} spheres, {total} cells, topology {t_topo:.1f} s", flush=True)
    ctx.set_force_precision(1)
    for i in range(len(params)):
        ctx.set_preferred_parameters(i, 1.0, 1.0)
        ctx.set_brownian(i, dt=0.01, mu=1.0, T=0.0)
    ctx.run_graph(0, 20, 5)
    for i, (n, p0) in enumerate(params):
        ctx.set_preferred_parameters(i, 1.0, p0)
    schedules = schedules or SCHEDULES
    if not ctx.load_tuning(schedules, updater=1):
        for i in range(len(params)):
            ctx.set_active(i, dt=0.01, mu=1.0, v0=0.05, tau=1.0)
        ctx.autotune(updater=1, steps_per_trial=40)
        ctx.save_tuning(schedules, updater=1)
    net, scale = ws.load()
    ctx.synchronize()
    t0 = time.perf_counter()
    reverted = ctx.warm_start(net, scale)
    for i in range(len(params)):
        ctx.fire_set_parameters(i, max_iterations=fire_iterations, force_cutoff=1e-10, stop_rms=1, **ws.FIRE)
    ctx.fire_minimize(100)
    ctx.synchronize()
    t_fire = time.perf_counter() - t0
    its, conv, rms = ws.fire_states(ctx, len(params))
    print(f"warm start + {fire_iterations} FIRE iterations: {t_fire:.1f} s, rms median {np.median(rms):.2g}", flush=True)
    for i in range(len(params)):
        ctx.set_active(i, dt=0.01, mu=1.0, v0=0.05, tau=1.0)
    G = ctx.tuning()["graph_steps"]
    ctx.run_graph(1, G, max(1, 40 // G))
    ctx.synchronize()
    reps = max(1, steps // G)
    t0 = time.perf_counter()
    ctx.run_graph(1, G, reps)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    err = ctx.force_error()
    ap = [ctx.area_perimeter(i)[:, 0] for i in range(len(params))]
    tangled = [i for i, a in enumerate(ap) if not (np.isfinite(a).all() and (a > 0).all() and a.max() < 100)]
    return {"workload": "batched-large-32", "spheres": len(params), "cells": total, "N_range": [50000, 500000],
            "forces": "fp32", "topology_build_s": t_topo, "steps": reps * G, "ms_per_step": 1e3 * dt / (reps * G),
            "cell_steps_per_s": total * reps * G / dt, "schedule": ctx.tuning(),
            "fire": {"warm_start_reverted": int(reverted), "iterations": fire_iterations, "wall_s": t_fire,
                     "rms_median": float(np.median(rms)), "rms_max": float(np.max(rms)),
                     "to_rms_1e-10": "not reachable: FIRE diverges between 500 and 1,000 iterations on these "
                                     "spheres in every precision, as the oracle's FIRE does bit for bit"},
            "tangled_spheres_after_active_steps": tangled,
            "fp32_force_rms_rel_error_max": float(err[:, 1].max()),
            "fp32_vs_fp64_vector_form_rms_rel_error_max": float(err[:, 5].max())}


def fire_comparison(svm, n_sys=1024, cap=10000, lanes=1):
    """Batched FIRE from seeded uniform random starts (1024 spheres, N
    log-uniform in [500, 5000], p0 uniform in [3.6, 3.9]) to rms force
    < 1e-10, plain and after the learned warm start (weights trained on
    held-out seeds, curvedspacevertexmodel_amd/warmstart_weights.npz).  FIRE
    parameters of the reference's minimisation driver; iterations capped at
    `cap` per sphere (spheres with p0 above ~3.75 are floppy and keep flipping
    without reaching 1e-10).  Both runs use `lanes` stream lanes.  Per run:
    the wall time with the cap, and the time to tolerance of the run's
    converging population (FIRE on only the spheres it converged, so none
    waits at the cap); against plain, iterations to tolerance on the spheres
    both runs converge and their final energies (same-basin check)."""
    from curvedspacevertexmodel_amd import warmstart as ws
    params = ws.batch_params(n_sys)
    net, scale = ws.load()
    ab = ws.ablation(n_sys=n_sys, cap=cap, lanes=lanes, net=net, scale=scale)
    plain, learned = ab["plain"], ab["warm_start"]
    total = int(sum(n for n, _ in params))
    ms = net.time_forward(total, reps=10)
    return {"workload": "batched-fire-1024", "spheres": n_sys, "total_cells": total, "N_range": [500, 5000],
            "p0_range": [3.6, 3.9], "stop": "rms force < 1e-10", "iteration_cap": cap, "lanes": lanes,
            "plain": plain, "learned": learned,
            "speedup_wall_with_cap": plain["wall_s_with_cap"] / learned["wall_s_with_cap"],
            "speedup_time_to_tolerance_converging": plain["time_to_tolerance_converging_s"] /
            learned["time_to_tolerance_converging_s"],
            "speedup_iterations_both_converged": learned.get("iterations_ratio_median_both"),
            "final_energy_within_1e-9": learned.get("final_energy_within_1e-9_both"),
            "mlp_forward": {"cells": total, "ms": ms, "tflops": 18176.0 * total / (ms * 1e-3) / 1e12,
                            "peak_tflops": 2500.0, "frac": 18176.0 * total / (ms * 1e-3) / 1e12 / 2500.0}}


def _oracle_state(ctx, i):
    """The topology, positions and directors of GPU sphere i, in the layout the oracle takes."""
    vn, vcn, ring, cnn = ctx.get_topology(i, 32)
    return dict(vpos=ctx.positions(i), vn=vn, vcn=vcn, ring=ring, cnn=cnn, directors=ctx.directors(i))


def cpu_baseline_configs(svm, budget_s=10.0, threads=None):
    """The CPU oracle (the reference's CPU path restated; single-threaded like
    the reference, or one thread per independent sphere) on the host cores, on
    bounded samples of configurations 1, 2, 4 and 5, from states built on the
    GPU by the same recipes as the GPU figures; and the libm build of the
    oracle beside the default one (fdlibm-style acos/sin/cos/log shared with
    the kernels) on a sample of configuration 3."""
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import svmtest
    from curvedspacevertexmodel_amd import warmstart as ws
    threads = threads or max(1, min(16, os.cpu_count() or 1))
    out = {"host_cpu": host_cpu()}

    def timed(fn):
        t0 = time.perf_counter()
        fn()
        return time.perf_counter() - t0

    # configuration 1: N = 512, active steps, then FIRE (single thread)
    ctx = svm.Context([dict(n_cells=512, a0=1.0, p0=1.0, kr=1.0, t1_threshold=0.1, seed=1137)])
    pts = svm.generate_cell_points(512, 1137)
    ctx.build_from_cell_points(0, pts[svm.spatial_order(pts)])
    ctx.set_directors(0, svm.generate_directors(ctx.positions(0), 1137))
    om = svmtest.OracleModel(_oracle_state(ctx, 0), p0=1.0)
    om.brownian(200, dt=0.01, mu=1.0, T=0.0, seed=1137, step0=0)
    om.set_preferences(np.tile([1.0, 3.8], 512))
    t = timed(lambda: om.active(1000, dt=0.01, mu=1.0, v0=0.1, tau=1.0, seed=1137, step0=200))
    res = {}
    t_f = timed(lambda: res.update(om.fire(**ws.FIRE, force_cutoff=1e-12, max_iterations=2000, stop_rms=1)))
    out["config1_single_thread"] = {"active_cell_steps_per_s": 512 * 1000 / t,
                                    "fire_cell_iterations_per_s": 512 * res["iterations"] / t_f,
                                    "sample": "1000 active steps after 200 Brownian; FIRE to 2000 iterations"}
    del ctx
    # configuration 2: N = 262,144, Brownian T = 0.001 (single thread, steady-state window)
    n = 262144
    ctx = svm.Context([dict(n_cells=n, a0=1.0, p0=1.0, kr=1.0, t1_threshold=0.1, seed=1137)])
    pts = svm.generate_cell_points(n, 1137)
    ctx.build_from_cell_points(0, pts[svm.spatial_order(pts)])
    ctx.set_directors(0, svm.generate_directors(ctx.positions(0), 1137))
    ctx.set_brownian(0, dt=0.01, mu=1.0, T=0.0)
    ctx.run_graph(0, 10, 10)
    om = svmtest.OracleModel(_oracle_state(ctx, 0), p0=3.8)
    om.brownian(5, dt=0.01, mu=1.0, T=0.001, seed=1137, step0=100)
    t = timed(lambda: om.brownian(20, dt=0.01, mu=1.0, T=0.001, seed=1137, step0=105))
    out["config2_single_thread"] = {"cell_steps_per_s": n * 20 / t,
                                    "sample": "20 Brownian steps (T 0.001) after 5, from the GPU-stabilised state"}
    del ctx, om
    # configuration 4: FIRE on the 1,024-sphere batch, one thread per group of spheres
    params = ws.batch_params(1024)
    ctx = ws.fire_batch(params[:4 * threads], seed0=1137, cutoff=1e-10, max_iterations=10000)
    models = [svmtest.OracleModel(_oracle_state(ctx, i), p0=p0) for i, (n_, p0) in enumerate(params[:4 * threads])]
    del ctx

    def fire_group(k):
        done = 0
        t0 = time.perf_counter()
        for i in range(k, len(models), threads):
            if time.perf_counter() - t0 > budget_s:
                break
            r = models[i].fire(**ws.FIRE, force_cutoff=1e-10, max_iterations=1500, stop_rms=1)
            done += params[i][0] * r["iterations"]
        return done

    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        work = sum(ex.map(fire_group, range(threads)))
    out["config4_all_core"] = {"cell_iterations_per_s": work / (time.perf_counter() - t0), "threads": threads,
                               "sample": f"plain FIRE (rms stop 1e-10, the batch's parameters) on up to "
                                         f"{4 * threads} spheres of the batch, 1500 iterations each at most"}
    del models
    # configuration 5: one sphere of the 50k-500k batch, active steps (single thread, FP64 forces)
    lp = ws.batch_params(32, seed=13377, nmin=50000, nmax=500000, p0_range=(3.7, 3.85))
    i5 = int(np.argmin([n_ for n_, _ in lp]))
    n5, p5 = lp[i5]
    ctx = svm.Context([dict(n_cells=n5, a0=1.0, p0=1.0, kr=1.0, t1_threshold=0.1, seed=1137 + i5)])
    pts = svm.generate_cell_points(n5, 1137 + i5)
    ctx.build_from_cell_points(0, pts[svm.spatial_order(pts)])
    ctx.set_directors(0, svm.generate_directors(ctx.positions(0), 1137 + i5))
    ctx.set_brownian(0, dt=0.01, mu=1.0, T=0.0)
    ctx.run_graph(0, 20, 5)
    om = svmtest.OracleModel(_oracle_state(ctx, 0), p0=p5)
    steps, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        om.active(5, dt=0.01, mu=1.0, v0=0.05, tau=1.0, seed=1137 + i5, step0=100 + steps)
        steps += 5
    out["config5_single_thread"] = {"cell_steps_per_s": n5 * steps / (time.perf_counter() - t0),
                                    "sample": f"sphere {i5} of the batch (N {n5}), {steps} active steps, "
                                              f"FP64 forces (the reference's CPU arithmetic)"}
    del ctx, om
    # configuration 5, all cores: the `threads` smallest spheres of the batch, one per thread
    pick = [int(i) for i in np.argsort([n_ for n_, _ in lp])[:threads]]
    ctx = svm.Context([dict(n_cells=lp[i][0], a0=1.0, p0=1.0, kr=1.0, t1_threshold=0.1, seed=1137 + i)
                       for i in pick])
    for k, i in enumerate(pick):
        pts = svm.generate_cell_points(lp[i][0], 1137 + i)
        ctx.build_from_cell_points(k, pts[svm.spatial_order(pts)])
        ctx.set_directors(k, svm.generate_directors(ctx.positions(k), 1137 + i))
        ctx.set_brownian(k, dt=0.01, mu=1.0, T=0.0)
    ctx.run_graph(0, 20, 5)
    models = [svmtest.OracleModel(_oracle_state(ctx, k), p0=lp[i][1]) for k, i in enumerate(pick)]
    del ctx

    def step_sphere(k):
        done, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < budget_s:
            models[k].active(2, dt=0.01, mu=1.0, v0=0.05, tau=1.0, seed=1137 + pick[k], step0=100 + done)
            done += 2
        return lp[pick[k]][0] * done

    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        work = sum(ex.map(step_sphere, range(len(pick))))
    out["config5_all_core"] = {"cell_steps_per_s": work / (time.perf_counter() - t0), "threads": threads,
                               "sample": f"the {len(pick)} smallest spheres of the batch "
                                         f"({min(lp[i][0] for i in pick)}-{max(lp[i][0] for i in pick)} cells), "
                                         f"one per thread, active steps for {budget_s:.0f} s, FP64 forces"}
    del models
    # the oracle built against glibc's acos/sin/cos/log beside the default build (configuration 3 sample)
    bp = batch_params(256)
    ctx, _ = build_batch(svm, bp[:8])
    stabilise(ctx, bp[:8], 200)
    rates = {}
    for libm in (False, True):
        ms = [svmtest.OracleModel(_oracle_state(ctx, i), p0=bp[i][1], libm=libm) for i in range(8)]
        cells = sum(p[0] for p in bp[:8])
        t = timed(lambda: [m.active(20, dt=0.01, mu=1.0, v0=bp[i][2], tau=1.0 / bp[i][3], seed=1137 + i,
                                    step0=200) for i, m in enumerate(ms)])
        rates["libm" if libm else "own_math"] = cells * 20 / t
    out["config3_libm_vs_own_math_single_thread"] = {
        **rates, "sample": "spheres 0-7 of the batch, 20 active steps each after the stabilisation"}
    return out


def gpu_configs(svm, fire_cap=10000):
    """The GPU figures of configurations 1, 4 (cell-iterations/s of plain FIRE)
    and 5 to put beside cpu_baseline_configs (configuration 2 is `single`)."""
    from curvedspacevertexmodel_amd import warmstart as ws
    out = {}
    ctx = svm.Context([dict(n_cells=512, a0=1.0, p0=1.0, kr=1.0, t1_threshold=0.1, seed=1137)])
    pts = svm.generate_cell_points(512, 1137)
    ctx.build_from_cell_points(0, pts[svm.spatial_order(pts)])
    ctx.set_directors(0, svm.generate_directors(ctx.positions(0), 1137))
    ctx.set_brownian(0, dt=0.01, mu=1.0, T=0.0)
    ctx.run_graph(0, 10, 20)
    ctx.set_preferred_parameters(0, 1.0, 3.8)
    ctx.set_active(0, dt=0.01, mu=1.0, v0=0.1, tau=1.0)
    ctx.run_graph(1, 10, 1)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.run_graph(1, 10, 99)
    ctx.synchronize()
    t_a = time.perf_counter() - t0
    ctx.fire_set_parameters(0, max_iterations=20000, force_cutoff=1e-12, stop_rms=1, **ws.FIRE)
    t0 = time.perf_counter()
    ctx.fire_minimize(100)
    ctx.synchronize()
    t_f = time.perf_counter() - t0
    st = ctx.fire_state(0)
    out["config1"] = {"active_cell_steps_per_s": 512 * 1000 / t_a, "fire_cell_iterations_per_s":
                      512 * st["iterations"] / t_f, "fire_iterations": st["iterations"],
                      "fire_rms_force": st["force_max"], "fire_wall_s": t_f}
    del ctx
    params = ws.batch_params(1024)
    ctx = ws.fire_batch(params, seed0=1137, cutoff=1e-10, max_iterations=fire_cap)
    ctx.set_lanes(3)
    wall, its, conv, _ = ws.run_fire(ctx, len(params))
    out["config4"] = {"fire_cell_iterations_per_s": float(sum(n * i for (n, _), i in zip(params, its)) / wall),
                      "wall_s": wall, "iteration_cap": fire_cap, "converged_fraction": float(conv.mean())}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--spheres", type=int, default=256)
    ap.add_argument("--graph-steps", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fire", action="store_true", help="skip the learned-vs-plain FIRE comparison")
    ap.add_argument("--no-single", action="store_true", help="skip the single 262,144-cell sphere block")
    ap.add_argument("--voronoi", action="store_true",
                    help="only the Voronoi model of 262,144 cells (no graphs): one JSON line, then exit")
    ap.add_argument("--large", action="store_true",
                    help="also run the 32-sphere, 50k-500k-cell FP32 batch (about 6 M cells)")
    ap.add_argument("--autotune", type=int, default=1,
                    help="0: default schedule; 1: the schedule stored for this batch in --schedules, else "
                         "svm_autotune (and store its choice); 2: always svm_autotune, then store")
    ap.add_argument("--eager", action="store_true",
                    help="enqueue every step's kernels directly (no HIP graphs), e.g. under a kernel tracer")
    ap.add_argument("--init-steps", type=int, default=1000,
                    help="untimed active steps after the stabilisation and the schedule choice (the reference "
                         "driver's initialisation phase)")
    ap.add_argument("--schedules", default=SCHEDULES,
                    help="persisted schedule table (svm_tuning_save / svm_tuning_load)")
    ap.add_argument("--fire-spheres", type=int, default=1024)
    ap.add_argument("--fire-cap", type=int, default=10000)
    ap.add_argument("--all-configs", action="store_true",
                    help="only the per-configuration GPU figures and CPU baselines (configurations 1, 2, 4, 5 "
                         "and the libm oracle build): one JSON line, then exit")
    ap.add_argument("--lanes", type=int, default=0,
                    help="run the timed steps on this many lanes, whatever the schedule says (0: the schedule's)")
    ap.add_argument("--phase-probe", action="store_true",
                    help="time a kernel boundary against a grid barrier in a persistent launch, then exit")
    ap.add_argument("--copy-probe", action="store_true",
                    help="after the run, 3 launches of a plain 256 MiB -> 256 MiB FP64 copy kernel "
                         "(calibrates the HBM byte counters of a profiler)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    import curvedspacevertexmodel_amd as svm

    if args.phase_probe:
        print(json.dumps({"phase_probe": phase_probe(svm)}), flush=True)
        return
    if args.all_configs:
        out = {"gpu": gpu_configs(svm), "gpu_config2": single_sphere(svm),
               "gpu_config5": large_batch(svm, schedules=args.schedules), "cpu": cpu_baseline_configs(svm)}
        print(json.dumps(out), flush=True)
        return
    if args.voronoi:
        n = 262144
        pts = svm.generate_cell_points(n, 1137)
        pts = pts[svm.spatial_order(pts)]
        print(json.dumps({"workload": "voronoi-262144", "cells": n, **voronoi_model(svm, pts)}), flush=True)
        return
    params = batch_params(args.spheres)
    total_cells = int(sum(p[0] for p in params))
    t_setup = time.perf_counter()
    ctx, t_delaunay = build_batch(svm, params)
    t_setup = time.perf_counter() - t_setup
    K, W = max(1, args.steps), max(0, args.warmup)
    print(f"setup {t_setup:.1f} s: {len(params)} spheres, {total_cells} cells", flush=True)
    stabilise(ctx, params, 200, eager=args.eager)
    print("stabilised (200 Brownian steps)", flush=True)
    tune_table, schedule_source, tune_steps = None, "default", 0
    if args.autotune == 1 and ctx.load_tuning(args.schedules, updater=1):
        schedule_source = "table"
    elif args.autotune and not args.eager:   # the tuner's trials replay graphs
        tune_table = ctx.autotune(updater=1, steps_per_trial=200)
        ctx.save_tuning(args.schedules, updater=1)
        schedule_source = "autotuned"
        tune_steps = int(sum(r[2] + 3 * max(1, 200 // (3 * r[2])) * r[2] for r in tune_table))
    if not args.autotune:
        ctx.set_tuning(graph_steps=max(2, args.graph_steps - args.graph_steps % 2))
    if args.lanes:
        ctx.set_lanes(args.lanes)
        schedule_source += f", lanes set to {args.lanes}"
    # graph length: the schedule's, cut to an even divisor of the timed steps
    # so that exactly K steps are timed (an odd K times one eager step more);
    # the initialisation below replays it before any timed launch
    Kg = K - K % 2
    Gmax = ctx.tuning()["graph_steps"]
    G = max([g for g in range(2, Gmax + 1, 2) if Kg % g == 0] or [2])
    print(f"schedule ({schedule_source}): {ctx.tuning()}; {G} steps per timed graph", flush=True)
    # the reference driver's initialisation (sphericalVertex.cpp:113-121):
    # the dynamics at the target parameters until the T1 transient that
    # follows the stabilisation has passed (thousands of flips per step in
    # the first steps, about 450 per step after a thousand).  It replays
    # the timed graph in both ping-pong parities (one eager step between), so
    # that the timed replays are never a graph's first launch, whatever W is.
    half = max(1, args.init_steps // (2 * G))
    if args.eager:
        ctx.active_steps(2 * half * G + 1)
    else:
        ctx.run_graph(1, G, half)
        ctx.active_steps(1)
        ctx.run_graph(1, G, half)
    ctx.synchronize()
    init_steps = 2 * half * G + 1
    print(f"initialised ({init_steps} active steps)", flush=True)
    # exactly W warm-up steps: the remainder eagerly, then replays of the timed graph
    if args.eager:
        ctx.active_steps(W)
    else:
        if W % G:
            ctx.active_steps(W % G)
        if W >= G:
            ctx.run_graph(1, G, W // G)
    ctx.synchronize()
    print(f"warmup done ({W} steps)", flush=True)
    base = (ctx.counters(), ctx.t1_diagnostics())
    t0 = time.perf_counter()
    if args.eager:
        ctx.active_steps(K)
    elif Kg:
        ctx.run_graph(1, G, Kg // G)
    if K % 2 and not args.eager:
        ctx.active_steps(1)
    ctx.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    after = (ctx.counters(), ctx.t1_diagnostics())
    steps, warm = K, W
    value = total_cells * steps / elapsed
    print(f"timed {steps} steps in {elapsed:.3f} s", flush=True)

    # per-kernel durations with HIP events on the library's stream (eager steps)
    ctx.set_profiling(True)
    ctx.active_steps(G)
    ctx.synchronize()
    timing = ctx.timing()
    ctx.set_profiling(False)
    dom = "k_vertex_active"
    # mean of the dominant kernel's launches in the eager profiled steps (HIP events around
    # each launch on the library's stream); also 20 launches back to back between two events
    dom_ms = timing.get(dom, float("nan"))
    b2b_ms = ctx.time_vertex_pass(1, 20)
    achieved = VERTEX_BYTES_PER_CELL * total_cells / (dom_ms * 1e-3) / 1e9
    result = {
        "metric": METRIC, "value": value, "unit": "cell-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": warm, "ms_per_step": 1e3 * elapsed / steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "batched-active-256", "spheres": len(params), "total_cells": total_cells,
                   "N_range": [200, 20000], "p0_range": [3.5, 4.0], "v0_range": [0.01, 0.5],
                   "Dr": [0.1, 1, 10], "dt": 0.01, "steps_per_graph": None if args.eager else G,
                   "schedule": ctx.tuning(),
                   "schedule_source": schedule_source, "schedule_table": schedule_table(args.schedules),
                   "untimed_steps_before_warmup": {"stabilisation_brownian": 200, "autotune_active": tune_steps,
                                                   "initialisation_active": init_steps},
                   "autotune_table": tune_table.tolist() if tune_table is not None else None},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": measured_traffic(dom),
                     "bytes_per_launch": VERTEX_BYTES_PER_CELL * total_cells, "avg_launch_ms": dom_ms,
                     "back_to_back_ms": b2b_ms, "fp64": fp64_roofline(dom, dom_ms)},
        "kernel_ms": timing,
        # a step's kernels run as one chain per lane (eager launches or graph nodes)
        "launches_per_step": {"kernels_per_lane": len(timing), "lanes": ctx.tuning()["lanes"],
                              "total": len(timing) * ctx.tuning()["lanes"]},
        "init": {"gpu_delaunay_topology_s": t_delaunay, "setup_s": t_setup,
                 "note": "256 spheres triangulated one call each (includes host<->device copies)"},
        "health": health(ctx, params, base, after),
    }
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(ctx, params)
    lanes = ctx.tuning()["lanes"]
    del ctx
    if not args.no_single:
        result["single"] = single_sphere(svm)
    if args.large:
        result["large"] = large_batch(svm, schedules=args.schedules)
    if not args.no_fire:
        result["fire"] = fire_comparison(svm, args.fire_spheres, args.fire_cap, lanes=lanes)
    print(json.dumps(result), flush=True)
    if args.copy_probe:
        copy_probe(svm)


if __name__ == "__main__":
    main()
