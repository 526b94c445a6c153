"""Seeded inputs: the library's generators against the oracle's.

sphericalModel::setParticlePositionsRandomly + setRadius
(src/model/sphericalModel.cpp:11-51) and the director draws of the
vertex-model constructor (src/model/sphericalVertexModel.cpp:58-73), with
Philox draws keyed by the sphere's seed.  The library's generators
(svm_generate_cell_points, svm_generate_directors) feed the benchmark; the
parity tests' inputs come from the oracle's.  They must be the same arrays,
bit for bit, and the benchmark's own pipeline (library points, GPU
triangulation, library directors) must give the oracle's trajectory.
"""
import os

import numpy as np
import pytest

import svmtest

LIB = os.path.join(svmtest.ROOT, "curvedspacevertexmodel_amd", "libsvm.so")


@pytest.mark.parametrize("n,seed", [(4, 1), (200, 1137), (5000, 1137 + 17), (20000, 1137 + 255), (262144, 1137)])
def test_library_generators_equal_the_oracle(n, seed):
    if not os.path.exists(LIB):
        pytest.skip("libsvm.so not built")
    import curvedspacevertexmodel_amd as svm
    L = svmtest.oracle_lib()
    pts = svm.generate_cell_points(n, seed)
    np.testing.assert_array_equal(pts, svmtest.cell_points(n, seed, L))
    # directors at arbitrary points of the sphere (the generator only projects)
    np.testing.assert_array_equal(svm.generate_directors(pts, seed + 5), svmtest.directors(pts, seed + 5, L))


@pytest.mark.gpu
@pytest.mark.parametrize("sphere", [3, 17, 101])
def test_bench_pipeline_inputs_give_the_oracle_trajectory(sphere):
    """One sphere of the batched workload built exactly as the benchmark
    builds it (library points in Z-order labels, GPU Delaunay, library
    directors) runs the oracle's active trajectory bit for bit."""
    import curvedspacevertexmodel_amd as svm
    sys_path_bench = svmtest.ROOT
    import sys
    sys.path.insert(0, sys_path_bench)
    import bench
    n, p0, v0, dr = bench.batch_params(sphere + 1)[sphere]
    seed = 1137 + sphere
    pts = svm.generate_cell_points(n, seed)
    pts = pts[svm.spatial_order(pts)]
    ctx = svm.Context([dict(n_cells=n, a0=1.0, p0=p0, kr=1.0, t1_threshold=0.1, seed=seed)])
    ctx.build_from_cell_points(0, pts)
    d = svm.generate_directors(ctx.positions(0), seed)
    ctx.set_directors(0, d)
    L = svmtest.oracle_lib()
    # the oracle's own inputs: same points, Qhull triangles, oracle topology and directors
    tris = svmtest.qhull_triangles(pts)
    np.testing.assert_array_equal(ctx.delaunay_triangles(0), tris)
    topo = svmtest.topology_from_triangles(pts, tris, L)
    np.testing.assert_array_equal(ctx.positions(0), topo["vpos"])
    topo["directors"] = svmtest.directors(topo["vpos"], seed, L)
    np.testing.assert_array_equal(d, topo["directors"])
    om = svmtest.OracleModel(topo, p0=p0)
    ctx.set_active(0, dt=0.01, mu=1.0, v0=v0, tau=1.0 / dr)
    ctx.active_steps(30)
    om.active(30, dt=0.01, mu=1.0, v0=v0, tau=1.0 / dr, seed=seed, step0=0)
    np.testing.assert_array_equal(ctx.positions(0), om.positions())
    np.testing.assert_array_equal(ctx.directors(0), om.directors())
