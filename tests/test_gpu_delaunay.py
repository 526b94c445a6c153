"""GPU spherical Delaunay triangulation, cell list and topology construction.

Bar: bit-exact.  The triangles equal the committed Qhull fixtures (which the
CPU tests pin by brute force and exact empty-circumcircle checks); the
topology built on the GPU (vertex positions, vertex lists, cell rings) equals
the CPU construction from Qhull triangles array for array; the cell list
equals a brute-force binning.
"""
import glob
import os

import numpy as np
import pytest

import svmtest
import curvedspacevertexmodel_amd as svm
from test_gpu_parity import make_ctx, assert_topology_equal
from test_oracle import load_fixture

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURES = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "delaunay_N*.txt")))


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_triangles_match_fixtures(path):
    pts, tris, _ = load_fixture(path)
    print(f"fixture {os.path.basename(path)}", flush=True)
    np.testing.assert_array_equal(svm.delaunay_triangles(pts), tris)


def test_triangles_of_scaled_and_rotated_points():
    """Points not at unit density (radius 1) and rotated: same combinatorics."""
    pts, tris, _ = load_fixture(os.path.join(ROOT, "tests", "golden", "delaunay_N512_s1137.txt"))
    unit = pts / np.linalg.norm(pts[0])
    np.testing.assert_array_equal(svm.delaunay_triangles(unit), tris)
    th = 0.7
    Q = np.array([[np.cos(th), -np.sin(th), 0.0], [np.sin(th), np.cos(th), 0.0], [0.0, 0.0, 1.0]])
    rot = pts @ Q.T
    np.testing.assert_array_equal(svm.delaunay_triangles(rot), svmtest.qhull_triangles(rot))


def test_large_triangulation_against_qhull():
    n = 100000
    pts = svmtest.cell_points(n, 4242)
    t = svm.delaunay_triangles(pts)
    np.testing.assert_array_equal(t, svmtest.qhull_triangles(pts))


def test_cell_list_matches_brute_force():
    pts = svmtest.cell_points(5000, 77)
    h = 2.5
    G, bin_of, start, order = svm.cell_list(pts, h)
    L = (1.0 + 1e-9) * np.abs(pts).max()
    side = 2 * L / G
    g = np.clip(np.floor((pts + L) / side).astype(np.int64), 0, G - 1)
    ref = (g[:, 0] * G + g[:, 1]) * G + g[:, 2]
    np.testing.assert_array_equal(bin_of, ref)
    counts = np.bincount(ref, minlength=G ** 3)
    np.testing.assert_array_equal(np.diff(start), counts)
    # each bin lists its points in index order (the sort is stable by index)
    np.testing.assert_array_equal(order, np.argsort(ref, kind="stable"))
    # and the construction is deterministic
    G2, b2, s2, o2 = svm.cell_list(pts, h)
    assert G2 == G and np.array_equal(o2, order) and np.array_equal(s2, start)


@pytest.mark.parametrize("n,seed", [(512, 1137), (3000, 1201)])
def test_build_from_cell_points_matches_cpu_construction(n, seed):
    st = svmtest.seeded_state(n, seed)
    om = svmtest.OracleModel(st)
    systems = [dict(n_cells=n, a0=1.0, p0=3.8, kr=1.0, t1_threshold=0.1, seed=seed)]
    ctx = svm.Context(systems)
    ctx.build_from_cell_points(0, st["points"])
    assert_topology_equal(ctx, 0, om)
    np.testing.assert_array_equal(ctx.positions(0), om.positions())
    np.testing.assert_array_equal(ctx.delaunay_triangles(0), st["tris"])
    # the same dynamics as a context given the CPU-built topology
    ref = make_ctx([st], seeds=[seed])
    for c in (ctx, ref):
        c.set_directors(0, st["directors"])
        c.set_active(0, dt=0.01, mu=1.0, v0=0.1, tau=1.0)
        c.active_steps(20)
    np.testing.assert_array_equal(ctx.positions(0), ref.positions(0))


def test_build_batch_of_spheres():
    sizes = [(200, 31), (700, 32), (1500, 33)]
    states = [svmtest.seeded_state(n, s) for n, s in sizes]
    ctx = svm.Context([dict(n_cells=n, a0=1.0, p0=3.8, kr=1.0, t1_threshold=0.1, seed=s) for n, s in sizes])
    for i, st in enumerate(states):
        ctx.build_from_cell_points(i, st["points"])
    for i, st in enumerate(states):
        assert_topology_equal(ctx, i, svmtest.OracleModel(st))


def test_large_empty_circle_needs_a_wider_search():
    """Sphere 100 of the batched-FIRE workload (N = 2513, seed 1237, spatially
    ordered): point 682 has Delaunay neighbours beyond the first search radius
    (a large empty circle), so its first star is unproven and is redone wider."""
    pts = svm.generate_cell_points(2513, 1237)
    pts = pts[svm.spatial_order(pts)]
    np.testing.assert_array_equal(svm.delaunay_triangles(pts), svmtest.qhull_triangles(pts))


def test_many_random_spheres_against_qhull():
    from curvedspacevertexmodel_amd import warmstart as ws
    for i, (n, _) in enumerate(ws.batch_params(96, seed=2024)):
        pts = svm.generate_cell_points(n, 5000 + i)
        if i % 2:
            pts = pts[svm.spatial_order(pts)]
        np.testing.assert_array_equal(svm.delaunay_triangles(pts), svmtest.qhull_triangles(pts), err_msg=f"sphere {i}")


def test_cocircular_points_are_rejected():
    """The eight vertices of a cube: each face's four vertices lie on one empty
    circle, so the Delaunay triangulation is not unique and the library reports
    degenerate input (SVM_ERR_TOPOLOGY) rather than pick a diagonal."""
    s = 2.0 / np.sqrt(3.0)
    cube = np.array([[x, y, z] for x in (-s, s) for y in (-s, s) for z in (-s, s)])
    with pytest.raises(svm.SVMError) as e:
        svm.delaunay_triangles(cube)
    assert e.value.code == 5


@pytest.mark.parametrize("seed,eps", [(1, 1e-13), (2, 3e-15), (3, 3e-15)])
def test_nearly_cocircular_points_are_triangulated_exactly(seed, eps):
    """A latitude-longitude grid (every grid quadrilateral is an isosceles
    trapezoid, so its corners are co-circular) moved off degeneracy by relative
    perturbations of 1e-13 or 3e-15 (tens of ulps): the floating-point filters leave many
    orientation tests to the exact arithmetic, and the angle keys of the wrap
    are nearly tied.  The triangulation must be closed and pass the
    empty-circumcircle test on every edge in exact rational arithmetic."""
    rng = np.random.default_rng(seed)
    lat = np.linspace(-80.0, 80.0, 17) * np.pi / 180.0
    lon = np.arange(36) * 2.0 * np.pi / 36.0
    pts = [[np.cos(a) * np.cos(b), np.cos(a) * np.sin(b), np.sin(a)] for a in lat for b in lon]
    pts = np.array(pts + [[0.0, 0.0, 1.0], [0.0, 0.0, -1.0]]) * 7.0
    pts *= 1.0 + eps * rng.standard_normal(pts.shape)
    tris = svm.delaunay_triangles(pts)
    assert tris.shape == (2 * len(pts) - 4, 3)
    assert svmtest.locally_delaunay_violations(pts, tris, exact=True) == []


def clustered_points(n, crowd, spread, seed=4):
    rng = np.random.default_rng(seed)
    R = np.sqrt(n / (4 * np.pi))
    v = rng.normal(size=(n, 3))
    v[:crowd] = np.array([0.0, 0.0, 1.0]) + spread * v[:crowd]     # a dense cap near the pole
    return R * v / np.linalg.norm(v, axis=1)[:, None]


@pytest.mark.parametrize("n,crowd,spread", [(2000, 1000, 0.15), (4000, 2000, 0.12)])
def test_dense_cap_beside_a_sparse_rest(n, crowd, spread):
    """Half the points in a dense polar cap: stars there hold 850 to 1,500
    candidates within the first search radius, beyond the 512-candidate
    capacity, so their radius is bisected down until the star fits and is
    proven (at most about 300 candidates lie inside each star's proving ball,
    checked on the CPU).  The triangulation equals Qhull's."""
    P = clustered_points(n, crowd, spread)
    np.testing.assert_array_equal(svm.delaunay_triangles(P), svmtest.qhull_triangles(P))


def test_cluster_beyond_the_search_capacity_is_rejected():
    """500 of 700 points in a tiny cap (about 450 times the mean density): the
    smallest ball that proves a star on the cap's rim holds more than 512
    points, so no search radius works; the input is rejected cleanly."""
    with pytest.raises(svm.SVMError) as e:
        svm.delaunay_triangles(clustered_points(700, 500, 0.08))
    assert e.value.code == 5


def test_large_void_needs_a_wide_search():
    """2,000 points with an empty polar cap of half-angle 0.42 rad: stars on its
    rim need search radii up to 3.3 r0 (at most 358 candidates inside the
    proving ball, checked on the CPU), on a sphere with far more than 512
    points, so no whole-sphere search can stand in.  The triangulation equals
    Qhull's."""
    rng = np.random.default_rng(11)
    n = 2000
    R = np.sqrt(n / (4 * np.pi))
    pts = []
    while len(pts) < n:
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        if v[2] < np.cos(0.42):
            pts.append(R * v)
    P = np.array(pts)
    np.testing.assert_array_equal(svm.delaunay_triangles(P), svmtest.qhull_triangles(P))


def test_cocircular_quadruple_on_a_nonempty_circle_is_triangulated():
    """Four exactly co-planar (so co-circular) points (+-s, +-s, h) whose
    circle holds another point, the pole (0, 0, R), inside its cap: the
    triangulation is unique (the four never share a Delaunay facet), so exact
    ties among them in the wrap's comparisons must not reject the input.  The
    triangulation equals Qhull's and passes the exact empty-circumcircle test."""
    rng = np.random.default_rng(21)
    n = 400
    R = np.sqrt(n / (4 * np.pi))
    s, h = 1.5, np.sqrt(R * R - 2 * 1.5 * 1.5)
    v = rng.normal(size=(n - 5, 3))
    v = R * v / np.linalg.norm(v, axis=1)[:, None]
    v = v[v[:, 2] < h - 1.0][: n - 5]        # keep the cap above the square's plane empty but for the pole
    square = np.array([[s, s, h], [-s, s, h], [-s, -s, h], [s, -s, h]])
    P = np.concatenate([square, [[0.0, 0.0, R]], v])
    tris = svm.delaunay_triangles(P)
    assert svmtest.locally_delaunay_violations(P, tris, exact=True) == []
    np.testing.assert_array_equal(tris, svmtest.qhull_triangles(P))
