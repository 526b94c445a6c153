"""The vertex pass's short root and quotient sequences (FastOps) against the
operators, bit for bit wherever they accept an argument, and forces through
them against the oracle where contributions are NaN (a cell centre on the
pole) and where the geometry is exactly symmetric (the octahedron)."""
import math

import numpy as np
import pytest

import svmtest
import curvedspacevertexmodel_amd as svm
from test_oracle import octahedron_state

pytestmark = pytest.mark.gpu


def test_fast_ops_equal_the_operators():
    n = 1 << 26
    dq, aq, bad_q, dr, ar, bad_r, far_q, far_r = svm.fast_ops_check(n, seed=20261204)
    print(f"quotients: {dq} in precondition, {aq} accepted, {bad_q} differ, {far_q} not nearest; "
          f"roots: {dr} drawn, {ar} accepted, {bad_r} differ, {far_r} not nearest", flush=True)
    assert bad_q == 0 and bad_r == 0
    assert far_q == 0 and far_r == 0
    assert dr == n and 0.4 * n < dq < n
    assert aq > 0.8 * dq and 0.6 * n < ar < 0.9 * n


def _forces(topo):
    n = topo["cnn"].shape[0]
    ctx = svm.Context([dict(n_cells=n, a0=1.0, p0=3.8, kr=1.0, t1_threshold=0.1, seed=1)])
    ctx.set_topology(0, topo["vpos"], topo["vn"], topo["vcn"], topo["ring"], topo["cnn"])
    ctx.compute_forces()
    om = svmtest.OracleModel(topo)
    om.compute_forces()
    return ctx, om


def test_pole_nan_contributions_match_the_oracle():
    R = math.sqrt(6 / (4 * math.pi))
    pts = np.array([[R, 0, 0], [-R, 0, 0], [0, R, 0], [0, -R, 0], [0, 0, R], [0, 0, -R]])
    simplices = [(a, b, c) for a in (0, 1) for b in (2, 3) for c in (4, 5)]
    topo = svmtest.topology_from_triangles(pts, svmtest.canonical_triangles(pts, simplices))
    ctx, om = _forces(topo)
    assert om.L.or_num_nan(om.h) == 8
    np.testing.assert_array_equal(ctx.forces(0), om.forces())
    assert ctx.counters()[0, 2] == 8


def test_octahedron_forces_match_the_oracle():
    topo, R, pts = octahedron_state()
    ctx, om = _forces(topo)
    np.testing.assert_array_equal(ctx.forces(0), om.forces())
    assert ctx.counters()[0, 2] == om.L.or_num_nan(om.h) == 0
