"""The larger workloads at their real sizes.

* One sphere of 262,144 cells (the single-system GPU configuration): GPU
  Delaunay topology equal to the CPU construction from Qhull, array for
  array; 10 active steps bit-exact against the CPU oracle; the FP32 force
  variant's error report.
* 32 spheres of 50,000-500,000 cells (5.8 M cells): GPU topology, FP32 forces
  with FP64 geometry and predicates, stabilisation, warm start, FIRE, then
  active steps.  Areas tile each sphere after stabilisation and after the
  warm start (which undoes itself on any sphere where a cell would fold).
  FIRE moves enforce the topology after every step, as the reference's
  moveParticles does (sphericalVertexModel.cpp:133,146), and at these sizes
  and shape indices its inertia folds a few cells now and then in every
  force precision, the FP64 parity path included; most folds unfold again
  within tens of iterations.  A fold that does not (one sphere of 32 in the
  FP32 run, none in the FP64 runs of the same batch) tangles that sphere, so
  the test allows at most two such spheres and reports them (small folds
  still unwinding at the end are reported too).  On every other sphere,
  areas are finite and positive at the end, and the FP32
  forces agree with the FP64 vector form of the same energy to 1e-5 rms
  (1e-7 measured along the whole FIRE trajectory of the sphere that
  tangles, so the tangle is the fold dynamics, not FP32 rounding).
  Topology counts hold on every sphere.
"""
import time

import numpy as np
import pytest

import svmtest
import curvedspacevertexmodel_amd as svm
from test_gpu_parity import assert_topology_equal

pytestmark = pytest.mark.gpu


def test_single_sphere_262144():
    n, seed = 262144, 1137
    pts = svm.generate_cell_points(n, seed)
    ctx = svm.Context([dict(n_cells=n, a0=1.0, p0=3.8, kr=1.0, t1_threshold=0.1, seed=seed)])
    t0 = time.perf_counter()
    ctx.build_from_cell_points(0, pts)
    print(f"GPU Delaunay + topology of {n} cells: {time.perf_counter() - t0:.2f} s", flush=True)
    tris = svmtest.qhull_triangles(pts)
    np.testing.assert_array_equal(ctx.delaunay_triangles(0), tris)
    topo = svmtest.topology_from_triangles(pts, tris)
    topo["directors"] = svmtest.directors(topo["vpos"], seed)
    topo["R"] = np.sqrt(n / (4 * np.pi))
    om = svmtest.OracleModel(topo)
    assert_topology_equal(ctx, 0, om)
    np.testing.assert_array_equal(ctx.positions(0), om.positions())
    ctx.set_directors(0, topo["directors"])
    # FP32 forces against the FP64 parity forces (the reference's expressions),
    # and against the well-conditioned FP64 evaluation of the same model
    rep = ctx.force_error()
    f32 = ctx.forces(0)
    om.compute_forces()
    fv = svmtest.vector_forces(om.positions(), om.cell_positions(), om.area_perimeter(), topo["vn"], topo["vcn"],
                               topo["ring"], topo["cnn"], topo["R"])
    d = np.linalg.norm(f32 - fv, axis=1)
    a = np.linalg.norm(fv, axis=1)
    rms_v = np.sqrt((d ** 2).sum() / (a ** 2).sum())
    print(f"FP32 force error at N={n}: vs FP64 reference expressions rms {rep[0, 1]:.2e} (max {rep[0, 0]:.2e}); "
          f"vs well-conditioned FP64 rms {rms_v:.2e} (max {d.max() / a.max():.2e})", flush=True)
    assert rms_v < 1e-5
    assert rep[0, 1] < 5e-3   # the reference's FP64 expressions lose ~2e-3 rms at this radius
    assert rep[0, 5] < 1e-5   # the library's own FP64 vector form agrees with FP32
    ctx.compute_forces(precision=2)
    g = ctx.forces(0)
    assert np.linalg.norm(g - fv) / np.linalg.norm(fv) < 1e-10
    ctx.set_active(0, dt=0.01, mu=1.0, v0=0.1, tau=1.0)
    ctx.active_steps(10)
    om.active(10, dt=0.01, mu=1.0, v0=0.1, tau=1.0, seed=seed, step0=0)
    np.testing.assert_array_equal(ctx.positions(0), om.positions())
    assert_topology_equal(ctx, 0, om)


def check_areas(ctx, params, phase):
    ctx.compute_geometry()
    bad = []
    for i, (n, _) in enumerate(params):
        a = ctx.area_perimeter(i)[:, 0]
        # the angle-sum areas tile the sphere to rounding, which grows with N
        # (the CPU oracle drifts by 3e-9 relative at N = 384k)
        if not (np.isfinite(a).all() and (a > 0).all() and abs(a.sum() / n - 1.0) < 1e-8):
            bad.append((i, n, float(np.nanmin(a)), float(np.nanmax(a)), float(a.sum() / n - 1.0)))
    print(f"{phase}: spheres with invalid areas {bad}", flush=True)
    return bad


def test_large_batch_fp32_warm_start_fire_active():
    from curvedspacevertexmodel_amd import warmstart as ws
    params = ws.batch_params(32, seed=13377, nmin=50000, nmax=500000, p0_range=(3.7, 3.85))
    total = sum(n for n, _ in params)
    ctx = svm.Context([dict(n_cells=n, a0=1.0, p0=p0, kr=1.0, t1_threshold=0.1, seed=1137 + i)
                       for i, (n, p0) in enumerate(params)])
    t0 = time.perf_counter()
    for i, (n, _) in enumerate(params):
        pts = svm.generate_cell_points(n, 1137 + i)
        ctx.build_from_cell_points(i, pts[svm.spatial_order(pts)])
        ctx.set_directors(i, svm.generate_directors(ctx.positions(i), 1137 + i))
    print(f"{len(params)} spheres, {total} cells: topology in {time.perf_counter() - t0:.1f} s", flush=True)
    err = ctx.force_error()
    print("FP32 vs FP64-reference force error per sphere (rms):", np.array2string(err[:, 1], precision=2), flush=True)
    assert (err[:, 1] < 2e-2).all()   # dominated by the FP64 reference expressions' own error at these radii
    ctx.set_force_precision(1)
    # the reference driver's stabilisation (sphericalVertex.cpp:100-110): raw random
    # starts otherwise squeeze small cells until they invert
    t0 = time.perf_counter()
    for i in range(len(params)):
        ctx.set_preferred_parameters(i, 1.0, 1.0)
        ctx.set_brownian(i, dt=0.01, mu=1.0, T=0.0)
    ctx.run_graph(0, 20, 5)
    for i, (n, p0) in enumerate(params):
        ctx.set_preferred_parameters(i, 1.0, p0)
    ctx.synchronize()
    print(f"100 stabilisation steps: {time.perf_counter() - t0:.1f} s", flush=True)
    assert not check_areas(ctx, params, "after stabilisation")
    net, scale = ws.load()
    t0 = time.perf_counter()
    rev = ctx.warm_start(net, scale)
    print(f"warm start undone on {rev} of {len(params)} spheres (a cell would fold)", flush=True)
    assert not check_areas(ctx, params, "after warm start")
    for i in range(len(params)):
        ctx.fire_set_parameters(i, max_iterations=200, force_cutoff=1e-10, stop_rms=1, **ws.FIRE)
    ctx.fire_minimize(100)
    ctx.synchronize()
    print(f"warm start + 200 FIRE iterations: {time.perf_counter() - t0:.1f} s", flush=True)
    # FIRE's inertia folds a few cells for some tens of iterations at a time on
    # these spheres, in every force precision including the FP64 parity path
    # (the model's own behaviour), so mid-run tiling is reported, not asserted
    check_areas(ctx, params, "after FIRE")
    for i, (n, p0) in enumerate(params):
        ctx.set_active(i, dt=0.01, mu=1.0, v0=0.05, tau=1.0)
    t0 = time.perf_counter()
    ctx.run_graph(1, 20, 100)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    print(f"2000 active steps (FP32 forces): {dt:.2f} s, {total * 2000 / dt:.3g} cell-steps/s", flush=True)
    bad = check_areas(ctx, params, "after active steps")
    # small folds still unwinding leave the tiling off by ~1e-6 with every area
    # in [0.7, 1.3]; a tangle has non-finite, non-positive or huge areas
    tangled = {b[0] for b in bad if not (np.isfinite(b[2]) and np.isfinite(b[3]) and 0 < b[2] and b[3] < 100
                                         and abs(b[4]) < 1e-2)}
    print(f"tangled spheres {sorted(tangled)}", flush=True)
    assert len(tangled) <= 2, bad
    err = ctx.force_error()
    print("FP32 vs FP64 vector-form force error per sphere (rms):", np.array2string(err[:, 5], precision=2),
          flush=True)
    for i, (n, _) in enumerate(params):
        vn, vcn, ring, cnn = ctx.get_topology(i, 32)
        assert cnn.sum() == 3 * (2 * n - 4)        # every vertex in three cells
        assert (cnn >= 3).all()
        if i in tangled:
            continue
        a = ctx.area_perimeter(i)[:, 0]
        assert np.isfinite(a).all() and (a > 0).all()
        assert err[i, 5] < 1e-5, (i, err[i])
    c = ctx.counters()
    print("T1 flips", int(c[:, 0].sum()), "sequential sphere-steps", int(c[:, 4].sum()), flush=True)


def test_t1_commit_overflow_matches_sequential():
    """A 262,144-cell sphere right after stabilisation flips ~280 edges per
    step, so the commit's vertices to rebuild outnumber its LDS list and go
    through the global list; the trajectory equals the sequential pass's bit
    for bit."""
    n, seed = 262144, 1137
    pts = svm.generate_cell_points(n, seed)
    pts = pts[svm.spatial_order(pts)]
    runs = []
    for sequential in (False, True):
        ctx = svm.Context([dict(n_cells=n, a0=1.0, p0=3.8, kr=1.0, t1_threshold=0.1, seed=seed)])
        ctx.build_from_cell_points(0, pts)
        ctx.set_t1_mode(sequential)
        ctx.set_preferred_parameters(0, 1.0, 1.0)
        ctx.set_brownian(0, dt=0.01, mu=1.0, T=0.0)
        ctx.brownian_steps(100)
        ctx.set_preferred_parameters(0, 1.0, 3.8)
        ctx.set_brownian(0, dt=0.01, mu=1.0, T=0.001)
        ctx.brownian_steps(50)              # the heaviest rearrangement, often resolved sequentially
        c0 = ctx.counters()[0].copy()
        t0 = time.perf_counter()
        ctx.brownian_steps(20)
        ctx.synchronize()
        dt = time.perf_counter() - t0
        c = ctx.counters()[0] - c0
        flips, seq = int(c[0]), int(c[4])
        print(f"{'sequential' if sequential else 'dry runs'}: {flips / 20:.0f} flips per step, "
              f"{seq} sequential passes in 20 steps, {dt / 20 * 1e3:.2f} ms per step", flush=True)
        runs.append((ctx.positions(0), ctx.get_topology(0, 32), flips, seq))
    # ~280 flips a step: with no LDS deduplication above 131,072 vertices the
    # commit's list of vertices to rebuild (about 20 ring entries per flip)
    # exceeds its 4096 entries whenever the step is committed from dry runs
    assert runs[0][2] == runs[1][2] > 200 * 20
    assert runs[0][3] < 10 and runs[1][3] == 20
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    for a, b in zip(runs[0][1], runs[1][1]):
        np.testing.assert_array_equal(a, b)


def test_fire_divergence_on_a_large_sphere_is_the_reference():
    """FIRE with the reference driver's parameters does not reach rms 1e-10 on
    the spheres of the large configuration: from the stabilised state of a
    50,000-cell sphere it diverges before 1,000 iterations.  The GPU's FP64
    parity run and the oracle's FIRE (the reference's CPU minimiser) reach the
    same rms force and the same positions, bit for bit, so the divergence is
    the minimiser's on this model, not the library's."""
    from curvedspacevertexmodel_amd import warmstart as ws
    from test_gpu_configs import oracle_from_gpu
    n, p0 = 50000, 3.75
    ctx = svm.Context([dict(n_cells=n, a0=1.0, p0=p0, kr=1.0, t1_threshold=0.1, seed=1137)])
    pts = svm.generate_cell_points(n, 1137)
    ctx.build_from_cell_points(0, pts[svm.spatial_order(pts)])
    ctx.set_preferred_parameters(0, 1.0, 1.0)
    ctx.set_brownian(0, dt=0.01, mu=1.0, T=0.0)
    ctx.brownian_steps(100)
    ctx.set_preferred_parameters(0, 1.0, p0)
    om = oracle_from_gpu(ctx, 0, p0)
    ctx.fire_set_parameters(0, max_iterations=1000, force_cutoff=1e-10, stop_rms=1, **ws.FIRE)
    ctx.fire_minimize(100)
    o = om.fire(**ws.FIRE, max_iterations=1000, force_cutoff=1e-10, stop_rms=1)
    s = ctx.fire_state(0)
    print(f"50,000 cells, 1,000 FIRE iterations: rms {s['force_max']:.4g} (oracle {o['force_max']:.4g})", flush=True)
    assert s["iterations"] == o["iterations"] == 1000
    assert s["force_max"] == o["force_max"] and s["force_max"] > 1e3
    np.testing.assert_array_equal(ctx.positions(0), om.positions())
