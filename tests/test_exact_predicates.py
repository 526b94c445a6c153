"""The orientation predicates of the GPU Delaunay code (svm_exact.h), built
for the host from the same header, against exact rational arithmetic
(svmtest.orient_exact).  Near-degenerate inputs -- four points on a small
circle of the sphere, whose in-circle determinant is a few ulps after
rounding -- are past the floating-point filter (checked with the filter's own
formula) and exercise the exact expansion arithmetic; exactly degenerate
inputs (cube faces, repeated points) must give 0."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import svmtest

ROOT = svmtest.ROOT


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("no g++")
    out = str(tmp_path_factory.mktemp("exact") / "exact_predicates")
    subprocess.run([gxx, "-std=c++17", "-O2", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Werror",
                    os.path.join(ROOT, "tests", "cpp", "exact_predicates.cpp"), "-o", out],
                   check=True, capture_output=True, text=True)
    return out


def run(exe, rows):
    text = "".join(" ".join(float(x).hex() for x in r) + "\n" for r in rows)
    r = subprocess.run([exe], input=text, capture_output=True, text=True, check=True)
    return [int(x) for x in r.stdout.split()]


def filter_decides(a, b, c, d):
    """The floating-point filter of orient3d_sign (same operations, same bound)."""
    bx, by, bz = b - a
    cx, cy, cz = c - a
    dx, dy, dz = d - a
    m0, m1, m2, m3, m4, m5 = cy * dz, cz * dy, cz * dx, cx * dz, cx * dy, cy * dx
    det = bx * (m0 - m1) + by * (m2 - m3) + bz * (m4 - m5)
    perm = abs(bx) * (abs(m0) + abs(m1)) + abs(by) * (abs(m2) + abs(m3)) + abs(bz) * (abs(m4) + abs(m5))
    return abs(det) > 1e-14 * perm


def small_circle_quads(n, R, seed):
    """n groups of four points on small circles of a sphere of radius R."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        e1 = np.cross(u, rng.normal(size=3))
        e1 /= np.linalg.norm(e1)
        e2 = np.cross(u, e1)
        rho = rng.uniform(0.02, 0.3)
        ang = rng.uniform(0, 2 * np.pi, 4)
        out.append(np.array([R * (np.cos(rho) * u + np.sin(rho) * (np.cos(t) * e1 + np.sin(t) * e2)) for t in ang]))
    return out


def test_near_cocircular_quadruples_take_the_exact_path(exe):
    quads = small_circle_quads(400, 7.3, 11)
    rows = [np.concatenate(q) for q in quads]
    got = run(exe, rows)
    want = [svmtest.orient_exact(*q) for q in quads]
    assert got == want
    undecided = sum(not filter_decides(*q) for q in quads)
    assert undecided >= 300            # the exact arithmetic decided these
    assert len(set(want)) >= 2         # both signs occur


def test_exactly_degenerate_inputs_give_zero(exe):
    c = np.array([[x, y, z] for x in (-1.0, 1.0) for y in (-1.0, 1.0) for z in (-1.0, 1.0)]) * 3.25
    faces = [[0, 1, 3, 2], [4, 5, 7, 6], [0, 1, 5, 4], [2, 3, 7, 6], [0, 2, 6, 4], [1, 3, 7, 5]]
    rows = [np.concatenate(c[f]) for f in faces]
    rows.append(np.concatenate([c[0], c[1], c[2], c[1]]))     # a repeated point
    rows.append(np.concatenate([c[0], -c[0], c[3]]))           # a, -a and c: det[a, -a, c] = 0
    assert run(exe, rows) == [0] * len(rows)


def test_random_and_origin_predicates(exe):
    rng = np.random.default_rng(5)
    P = rng.normal(size=(600, 3))
    P *= 4.0 / np.linalg.norm(P, axis=1)[:, None]
    quads = P.reshape(-1, 4, 3)
    assert run(exe, [q.ravel() for q in quads]) == [svmtest.orient_exact(*q) for q in quads]
    tri = P.reshape(-1, 3, 3)
    assert run(exe, [t.ravel() for t in tri]) == [svmtest.orient_exact(*t) for t in tri]
    # three nearly collinear points on a great circle: det[a, b, c] of a few ulps
    t = rng.uniform(0, 2 * np.pi, (200, 3))
    u, v = np.array([1.0, 2.0, 2.0]) / 3.0, np.array([2.0, -2.0, 1.0]) / 3.0
    rows = [np.concatenate([np.cos(x) * u + np.sin(x) * v for x in ts]) for ts in t]
    assert run(exe, rows) == [svmtest.orient_exact(r[:3], r[3:6], r[6:]) for r in rows]
