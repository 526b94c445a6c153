"""Metric neighbour lists on the GPU (svm_neighbor_list) against the oracle's
restatement of neighborList::computeCPU (src/utilities/neighborList.cpp:102-166),
bit for bit: counts, neighbour indices, difference vectors and distances;
at full size against an independent k-d tree radius search."""
import ctypes as C

import numpy as np
import pytest
from scipy.spatial import cKDTree

import svmtest
import curvedspacevertexmodel_amd as svm

pytestmark = pytest.mark.gpu


def assert_lists_equal(got, want):
    counts, idx, vec, dist = got
    ocounts, oidx, ovec, odist = want
    np.testing.assert_array_equal(counts, ocounts)
    assert idx.shape == oidx.shape
    np.testing.assert_array_equal(idx, oidx)
    np.testing.assert_array_equal(vec, ovec)
    np.testing.assert_array_equal(dist, odist)


@pytest.mark.parametrize("n,seed", [(64, 5), (5000, 1137), (20000, 9)])
def test_cell_points_match_the_oracle(n, seed):
    """The list sphericalModel builds: range 1, box half-side R + 2.5."""
    pts = svm.generate_cell_points(n, seed)
    assert_lists_equal(svm.neighbor_list(pts, 1.0), svmtest.oracle_neighbor_list(pts, 1.0))


def test_vertices_and_other_ranges_match_the_oracle():
    st = svmtest.seeded_state(3000, 1200)
    v = st["vpos"]
    for rng in (0.3, 1.0, 2.5):
        assert_lists_equal(svm.neighbor_list(v, rng), svmtest.oracle_neighbor_list(v, rng))


def test_capacity_then_recompute():
    """A width below the largest count returns SVM_ERR_CAPACITY with the
    width needed and the counts; the binding then recomputes, as the
    reference does with a larger Nmax."""
    pts = svm.generate_cell_points(2000, 3)
    n = len(pts)
    counts = np.empty(n, np.int32)
    idx = np.empty((n, 1), np.int32)
    vec = np.empty((n, 1, 3))
    dist = np.empty((n, 1))
    needed = C.c_int32(0)
    rc = svm.lib().svm_neighbor_list(C.c_int32(n), svm._dp(np.ascontiguousarray(pts)), C.c_double(1.0),
                                     C.c_double(float(np.abs(pts).max()) + 2.5), C.c_int32(1), svm._ip(counts),
                                     svm._ip(idx), svm._dp(vec), svm._dp(dist), C.byref(needed))
    want = svmtest.oracle_neighbor_list(pts, 1.0)
    assert rc == 8 and needed.value == want[0].max() > 1
    np.testing.assert_array_equal(counts, want[0])
    assert_lists_equal(svm.neighbor_list(pts, 1.0, max_neighbors=1), want)


def test_full_size_against_a_kdtree():
    """262,144 cell points: every list equals the k-d tree's radius search
    (sets, and distances by the reference's expression)."""
    n = 262144
    pts = svm.generate_cell_points(n, 1137)
    counts, idx, vec, dist = svm.neighbor_list(pts, 1.0)
    tree = cKDTree(pts)
    pairs = tree.query_pairs(1.0 + 1e-9, output_type="ndarray")
    d = pts[pairs[:, 0]] - pts[pairs[:, 1]]
    r = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
    pairs = pairs[r < 1.0]
    assert counts.sum() == 2 * len(pairs)
    # every listed entry is a pair of the tree's, with the reference's d and |d|
    rows = np.repeat(np.arange(n), counts)
    used = np.arange(idx.shape[1])[None, :] < counts[:, None]
    cols = idx[used]
    lo, hi = np.minimum(rows, cols), np.maximum(rows, cols)
    key = lo.astype(np.int64) * n + hi
    want = pairs[:, 0].astype(np.int64) * n + pairs[:, 1]
    np.testing.assert_array_equal(np.unique(key), np.sort(want))
    e = vec[used]
    np.testing.assert_array_equal(e, pts[rows] - pts[cols])
    dd = dist[used]
    np.testing.assert_array_equal(dd, np.sqrt(e[:, 0] * e[:, 0] + e[:, 1] * e[:, 1] + e[:, 2] * e[:, 2]))
