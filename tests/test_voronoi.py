"""Voronoi-model forces: the oracle (oracle/voronoi_oracle.py) against finite
differences of the energy of the Voronoi cells, on the CPU."""
import numpy as np

import svmtest
import voronoi_oracle as vo


def voronoi_model(points, p0=3.8):
    """The oracle's model of the Voronoi cells of `points` (Qhull triangles,
    circumcentre vertices); returns (triangles, model with forces computed)."""
    tris = svmtest.qhull_triangles(points)
    topo = svmtest.topology_from_triangles(points, tris)
    topo["vpos"] = vo.circumcentres(points, tris, np.linalg.norm(points[0]))
    om = svmtest.OracleModel(topo, p0=p0)
    om.compute_forces()
    return tris, om


def test_generator_forces_are_minus_the_energy_gradient():
    pts = svmtest.cell_points(300, 1180)
    R = np.linalg.norm(pts[0])
    tris, om = voronoi_model(pts)
    F = vo.generator_forces(pts, tris, om.forces(), R)
    assert np.abs(np.sum(F * pts, 1)).max() < 1e-9 * np.abs(F).max() * R
    rng = np.random.default_rng(5)
    h = 1e-5
    errs = []
    for i in rng.choice(len(pts), 24, replace=False):
        p = pts[i]
        e1 = np.cross(p, rng.normal(size=3))
        e1 /= np.linalg.norm(e1)
        e2 = np.cross(p / R, e1)
        for e in (e1, e2):
            E = []
            for sgn in (1, -1):
                q = pts.copy()
                q[i] = R * (p + sgn * h * e) / np.linalg.norm(p + sgn * h * e)
                t2, m2 = voronoi_model(q)
                assert np.array_equal(t2, tris)
                E.append(m2.energy(0))
            errs.append(abs(F[i] @ e + (E[0] - E[1]) / (2 * h)))
    assert max(errs) < 1e-6 * np.abs(F).max(), max(errs)
