"""Trajectory output and dynamical analysis (the reference's runTrajectory
driver, runTrajectory.cpp:128-169).

CPU: the NetCDF writer against scipy's reader of the classic format; the
log-spaced save schedule against its definition; the oracle's MSD and overlap
(dynamicalFeatures.cpp:17-53) pinned by a rigid rotation, where every cell's
geodesic displacement has a closed form.  GPU (test_gpu_trajectory.py): the
library's kernels against the oracle.
"""
import math

import numpy as np
import pytest

import svmtest
import curvedspacevertexmodel_amd as svm
from curvedspacevertexmodel_amd.trajectory import logSpacedIntegers
from test_oracle import move_np


def test_netcdf_records_read_back(tmp_path):
    from scipy.io import netcdf_file
    path = str(tmp_path / "cells.nc")
    rng = np.random.default_rng(3)
    vecs = rng.normal(size=(5, 12))
    vals = rng.normal(size=5)
    with svm.VectorValueNetCDF(path, 12) as f:
        for v, x in zip(vecs, vals):
            f.writeState(v, x)
        assert f.num_records == 5
    with netcdf_file(path, "r", mmap=False) as nc:
        assert nc.dimensions["record"] is None          # the unlimited dimension
        assert nc.dimensions["unit"] == 1 and nc.dimensions["N"] == 12
        assert nc.variables["value"].dimensions == ("record", "unit")
        assert nc.variables["vector"].dimensions == ("record", "N")
        np.testing.assert_array_equal(nc.variables["vector"][:], vecs)
        np.testing.assert_array_equal(nc.variables["value"][:, 0], vals)
    with pytest.raises(ValueError):
        svm.VectorValueNetCDF(str(tmp_path / "x.nc"), 3).writeState([1.0, 2.0], 0.0)


def test_log_spaced_save_points():
    """Save points are the distinct round(10^(e k)), k >= 1, in order (C rounding)."""
    lsi = logSpacedIntegers(0, 0.05)
    got = []
    for _ in range(60):
        lsi.update()
        got.append(lsi.nextSave)
    want = sorted({int(math.floor(10 ** (0.05 * k) + 0.5)) for k in range(1, 200)})
    assert got == want[:60]
    # starting later: the first save is the first point at or after firstSave
    lsi = logSpacedIntegers(100, 0.05)
    assert lsi.nextSave == 100
    lsi.update()
    assert lsi.nextSave == 112


def rot_z(P, theta):
    c, s = math.cos(theta), math.sin(theta)
    return np.stack([c * P[:, 0] - s * P[:, 1], s * P[:, 0] + c * P[:, 1], P[:, 2]], 1)


def test_oracle_msd_and_overlap_of_a_rigid_rotation():
    """Cells rotated by theta about z move by R acos(cos^2 t + sin^2 t cos theta)
    along the sphere (t: polar angle); the MSD is the mean square of that and
    the overlap the fraction below the cutoff."""
    st = svmtest.seeded_state(600, 1150)
    om = svmtest.OracleModel(st)
    om.compute_forces()
    C = om.cell_positions()
    R = st["R"]
    theta = 0.3
    ref = rot_z(C, -theta)
    cos_t = C[:, 2] / np.linalg.norm(C, axis=1)
    alpha = np.arccos(cos_t ** 2 + (1 - cos_t ** 2) * math.cos(theta))
    dist = R * alpha
    for fraction in (1.0, 0.5):
        n = len(C) if fraction >= 1 else int(math.floor(len(C) * fraction))
        for cutoff in (0.5, 1.5):
            msd, ov = om.msd_overlap(ref, fraction, cutoff, reduce_mode=0)
            np.testing.assert_allclose(msd, np.mean(dist[:n] ** 2), rtol=1e-9)
            assert ov == np.mean(dist[:n] < cutoff)
            msd_t, ov_t = om.msd_overlap(ref, fraction, cutoff, reduce_mode=1)
            np.testing.assert_allclose(msd_t, msd, rtol=1e-13)
            assert ov_t == ov
    # no motion: zero MSD.  A cell that has not moved can still have its
    # cosine with itself rounded above 1, so acos gives NaN: the reference
    # neither adds such a distance to the MSD nor counts it as overlapping
    msd, ov = om.msd_overlap(C, 1.0, 0.5)
    assert msd < 1e-12
    assert 0.9 < ov <= 1.0


# ------------------------------------------------ spherical vectorial Vicsek --
def polar_order(n):
    return np.linalg.norm(n.sum(0)) / len(n)


def test_oracle_vicsek_alignment_noise_and_motion():
    """sphericalVectorialVicsek on the vertex model (oracle): directors stay unit
    tangents; without noise the alignment with the neighbours raises the polar
    order; with overwhelming noise (eta 1e6, rate dt/tau = 1) each new director
    is the random direction alone, isotropic in the tangent plane; with mu = 0
    the first move is the sphere move by dt v0 n."""
    st = svmtest.seeded_state(400, 1170)
    R = st["R"]
    om = svmtest.OracleModel(st)
    P0, N0 = om.positions(), om.directors()
    om.vicsek(1, dt=0.01, mu=0.0, v0=0.5, tau=1.0, eta=0.0)
    # every vertex but the two of each T1 flip (which repositions them) made the plain move
    moved = np.abs(om.positions() - move_np(P0, 0.01 * 0.5 * N0, R)).max(1) > 1e-12
    assert moved.sum() <= 2 * om.n_t1
    om.vicsek(60, dt=0.01, mu=1.0, v0=0.05, tau=0.05, eta=0.0, step0=1)
    P, N = om.positions(), om.directors()
    np.testing.assert_allclose(np.linalg.norm(N, axis=1), 1.0, atol=1e-12)
    assert np.abs(np.sum(N * P, 1)).max() < 1e-9
    # local alignment: neighbouring directors agree far more than at the start
    vn = om.topology(svmtest.RING_CAP)[0]
    def neighbour_agreement(n):
        return np.mean(np.sum(n[:, None, :] * n[vn], 2))
    assert neighbour_agreement(N) > 0.8 and neighbour_agreement(N0) < 0.3
    # noise only: a random tangent direction per vertex
    om2 = svmtest.OracleModel(st)
    om2.vicsek(1, dt=0.01, mu=0.0, v0=0.0, tau=0.01, eta=1e6)
    P2, N2 = om2.positions(), om2.directors()
    e1 = np.cross(P2, [0.0, 0.0, 1.0])
    e1 /= np.linalg.norm(e1, axis=1)[:, None]
    e2 = np.cross(P2 / R, e1)
    ang = np.arctan2(np.sum(N2 * e2, 1), np.sum(N2 * e1, 1))
    from scipy.stats import kstest
    assert kstest((ang + np.pi) / (2 * np.pi), "uniform").pvalue > 1e-3
    assert polar_order(N2) < 0.15
