"""Voronoi-model generator forces on the CPU.  Test infrastructure only: the
tests and nothing else import it, as the checker of the GPU path.

The reference's sphericalVoronoi (src/model/sphericalVoronoi.cpp:8-12) keeps
only its topology, the convex hull of the generators.  Here the Voronoi cells
carry the vertex model's energy E = sum Kr (A - A0)^2 + (P - P0)^2, their
vertices being the circumcentres of the Delaunay triangles on the sphere,
v = R n / |n| with n = (b - a) x (c - a) for the counter-clockwise triangle
(a, b, c).  The generator forces follow from the vertex forces F_v = -dE/dv
(the oracle's vertex-model forces, oracle.cpp) by the chain rule:
dn/da = [c - b]x, dn/db = [a - c]x, dn/dc = [b - a]x and
dv/dn = R (I - n^ n^T) / |n|, so with y = (R / |n|) (F_v - n^ (n^ . F_v)) the
generators a, b, c receive (b - c) x y, (c - a) x y and (a - b) x y; the sums
are projected on the tangent plane.

Pinned by central finite differences of the energy along tangent directions
(tests/test_voronoi.py).
"""
import numpy as np


def circumcentres(points, tris, radius):
    """Vertex positions: the circumcentre of each triangle on the sphere,
    (radius / |n|) n with |n| summed in component order."""
    a, b, c = (points[tris[:, k]] for k in range(3))
    n = np.cross(b - a, c - a)
    nrm = np.sqrt(n[:, 0] * n[:, 0] + n[:, 1] * n[:, 1] + n[:, 2] * n[:, 2])
    return (radius / nrm)[:, None] * n


def generator_forces(points, tris, vertex_forces, radius):
    """Forces [N][3] on the generators from the vertex forces [2N-4][3]."""
    a, b, c = (points[tris[:, k]] for k in range(3))
    n = np.cross(b - a, c - a)
    nn = np.linalg.norm(n, axis=1)
    nh = n / nn[:, None]
    s = np.where(np.sum(n * (a + b + c), 1) < 0, -1.0, 1.0)
    y = (s * radius / nn)[:, None] * (vertex_forces - nh * np.sum(nh * vertex_forces, 1)[:, None])
    F = np.zeros_like(points)
    np.add.at(F, tris[:, 0], np.cross(b - c, y))
    np.add.at(F, tris[:, 1], np.cross(c - a, y))
    np.add.at(F, tris[:, 2], np.cross(a - b, y))
    return F - (np.sum(F * points, 1) / np.sum(points * points, 1))[:, None] * points
