"""Voronoi-model dynamics: the generators move, and every move is followed by
the GPU retriangulation (the reference's sphericalVoronoi recomputes its
convex hull after each move, src/model/sphericalModel.cpp:96-121), then the
forces of the Voronoi cells' energy on the generators (svm_voronoi_forces).

    overdamped_steps   x <- move(x, dt mu F(x)) on every sphere of a context
"""
import numpy as np


def sphere_move(points, disp):
    """sphericalDomain::move (sphericalDomain.h:203-209): project the
    displacement on the tangent plane, add, put back on the sphere."""
    R = np.sqrt(np.sum(points[0] * points[0]))
    d = disp - (np.sum(disp * points, 1) / np.sum(points * points, 1))[:, None] * points
    q = points + d
    return R * q / np.linalg.norm(q, axis=1)[:, None]


def overdamped_steps(ctx, points, steps, dt=0.01, mu=1.0):
    """`steps` overdamped steps of every sphere's generators (list of [N][3]
    arrays, one per sphere of ctx); returns the new generators and the
    potential energy per sphere before each step."""
    points = [np.ascontiguousarray(p, np.float64) for p in points]
    energies = []
    for _ in range(steps):
        for s, p in enumerate(points):
            ctx.build_voronoi_from_cell_points(s, p)
        forces = ctx.voronoi_forces()
        energies.append(ctx.energies()[0])
        points = [sphere_move(p, dt * mu * f) for p, f in zip(points, forces)]
    return points, np.array(energies)
