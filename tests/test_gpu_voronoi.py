"""Voronoi-model forces on the GPU against the oracle (oracle/voronoi_oracle.py,
pinned by finite differences in test_voronoi.py).

Bar: the topology equals the CPU construction; vertex positions (circumcentres)
and vertex forces bit-exact; generator forces (the chain-rule sums, in another
order) within 1e-12 relative; overdamped dynamics with a retriangulation after
every move lower the energy of every sphere.
"""
import numpy as np
import pytest

import svmtest
import voronoi_oracle as vo
import curvedspacevertexmodel_amd as svm
from curvedspacevertexmodel_amd.voronoi import overdamped_steps

pytestmark = pytest.mark.gpu


def test_voronoi_forces_match_oracle():
    sizes = (300, 1200, 4000)
    pts = [svmtest.cell_points(n, 1190 + i) for i, n in enumerate(sizes)]
    ctx = svm.Context([dict(n_cells=n, a0=1.0, p0=3.8, kr=1.0, t1_threshold=0.1, seed=1190 + i)
                       for i, n in enumerate(sizes)])
    for i, p in enumerate(pts):
        ctx.build_voronoi_from_cell_points(i, p)
    forces = ctx.voronoi_forces()
    for i, p in enumerate(pts):
        R = np.sqrt(p[0, 0] * p[0, 0] + p[0, 1] * p[0, 1] + p[0, 2] * p[0, 2])
        tris = svmtest.qhull_triangles(p)
        np.testing.assert_array_equal(ctx.delaunay_triangles(i), tris)
        topo = svmtest.topology_from_triangles(p, tris)
        topo["vpos"] = vo.circumcentres(p, tris, R)
        np.testing.assert_array_equal(ctx.positions(i), topo["vpos"])
        om = svmtest.OracleModel(topo)
        om.compute_forces()
        np.testing.assert_array_equal(ctx.forces(i), om.forces())
        F = vo.generator_forces(p, tris, om.forces(), R)
        assert np.abs(forces[i] - F).max() <= 1e-12 * np.abs(F).max()


def test_voronoi_relaxation_lowers_the_energy():
    sizes = (500, 2000)
    pts = [svmtest.cell_points(n, 1195 + i) for i, n in enumerate(sizes)]
    ctx = svm.Context([dict(n_cells=n, a0=1.0, p0=3.7, kr=1.0, t1_threshold=0.1, seed=1195 + i)
                       for i, n in enumerate(sizes)])
    _, E = overdamped_steps(ctx, pts, 40, dt=0.02)
    assert (np.diff(E, axis=0) < 0).all(), E[:5]
    assert (E[-1] < 0.5 * E[0]).all()
