"""svm_delaunay_triangles_noalloc against svm_delaunay_triangles.

The allocation-free variant runs the same kernels in the library's static
per-GPU work array.  Its triangles must equal the allocating entry point's
(and Qhull's) on consecutive calls, the second one smaller than the first, so
a stale larger result left in the shared array cannot pass for the new one.
"""
import numpy as np
import pytest

import svmtest
import curvedspacevertexmodel_amd as svm

pytestmark = pytest.mark.gpu


def test_noalloc_triangles_equal_the_allocating_entry_point():
    big = svmtest.cell_points(50000, 1137)
    small = svmtest.cell_points(3001, 1138)
    for pts in (big, small):
        ref = svm.delaunay_triangles(pts)
        got = svm.delaunay_triangles_noalloc(pts)
        np.testing.assert_array_equal(got, ref)
    np.testing.assert_array_equal(svm.delaunay_triangles_noalloc(small), svmtest.qhull_triangles(small))


def test_noalloc_on_a_caller_stream_and_at_the_size_limit():
    import torch
    s = torch.cuda.Stream()
    pts = svmtest.cell_points(svm.NOALLOC_MAX_POINTS, 1139)
    got = svm.delaunay_triangles_noalloc(pts, stream=s.cuda_stream)
    np.testing.assert_array_equal(got, svm.delaunay_triangles(pts))
    pts2 = svmtest.cell_points(777, 1140)
    np.testing.assert_array_equal(svm.delaunay_triangles_noalloc(pts2, stream=s.cuda_stream),
                                  svm.delaunay_triangles(pts2))


def test_noalloc_refuses_input_beyond_its_limit():
    pts = svmtest.cell_points(svm.NOALLOC_MAX_POINTS + 1, 1141)
    with pytest.raises(svm.SVMError) as e:
        svm.delaunay_triangles_noalloc(pts)
    assert e.value.code == 8   # SVM_ERR_CAPACITY
