"""Dynamical analysis on the GPU (dynamicalFeatures, src/analysis/dynamicalFeatures.cpp:17-53)
against the oracle, and the trajectory driver (runTrajectory.cpp:128-169).

Bar: MSD and overlap bit-exact against the oracle with the device's
fixed-tree sums, per sphere of a batch, for the whole sphere and for a
fraction of its cells; the driver's files hold what the kernels measured.
"""
import numpy as np
import pytest

import svmtest
from test_gpu_parity import make_ctx
from curvedspacevertexmodel_amd.trajectory import run_trajectory

pytestmark = pytest.mark.gpu


def test_msd_overlap_match_oracle():
    states = [svmtest.seeded_state(n, 1137 + i) for i, n in enumerate((512, 1500))]
    ctx = make_ctx(states)
    for i in range(len(states)):
        ctx.set_active(i, dt=0.01, mu=1.0, v0=0.1, tau=1.0)
    ctx.compute_geometry()
    ctx.dyn_set_reference()
    ctx.active_steps(100)
    ctx.compute_geometry()
    oms, refs = [], []
    for i, st in enumerate(states):
        om = svmtest.OracleModel(st)
        om.compute_forces()
        refs.append(om.cell_positions())
        om.active(100, dt=0.01, mu=1.0, v0=0.1, tau=1.0, seed=1137 + i, step0=0)
        om.compute_forces()
        oms.append(om)
    for fraction, cutoff in ((1.0, 0.5), (0.5, 0.2)):
        msd, ov = ctx.dyn_measure(fraction, cutoff)
        for i, om in enumerate(oms):
            np.testing.assert_array_equal(ctx.cell_positions(i), om.cell_positions())
            m, o = om.msd_overlap(refs[i], fraction, cutoff, reduce_mode=1)
            assert msd[i] == m and ov[i] == o, (i, fraction, msd[i], m, ov[i], o)
        assert (msd > 0).all() and (ov < 1).all()


def test_trajectory_files_hold_the_measurements(tmp_path):
    from scipy.io import netcdf_file
    st = svmtest.seeded_state(800, 1160)
    ctx = make_ctx([st])
    ctx.set_active(0, dt=0.01, mu=1.0, v0=0.2, tau=1.0)
    saves, msd, ov = run_trajectory(ctx, 1, 120, 0.01, str(tmp_path), tag="t")
    assert list(saves[:5]) == [1, 2, 3, 4, 5] and saves[-1] == 112
    with netcdf_file(str(tmp_path / "t_msd_s0.nc"), "r", mmap=False) as f:
        v = f.variables["vector"][:].copy()
        np.testing.assert_array_equal(v[:, 0], saves * 0.01)
        np.testing.assert_array_equal(v[:, 1], msd[:, 0])
        np.testing.assert_array_equal(f.variables["value"][:, 0], saves)
    with netcdf_file(str(tmp_path / "t_overlap_s0.nc"), "r", mmap=False) as f:
        np.testing.assert_array_equal(f.variables["vector"][:, 1], ov[:, 0])
    with netcdf_file(str(tmp_path / "t_cellPositions_s0.nc"), "r", mmap=False) as f:
        assert f.variables["vector"].shape == (len(saves), 3 * 800)
    assert np.all(np.diff(msd[:, 0]) > 0)          # active cells spread out


def test_vicsek_matches_oracle():
    """sphericalVectorialVicsek on the GPU: positions, directors and topology
    bit-exact against the oracle over 120 steps with T1 flips."""
    from test_gpu_parity import assert_topology_equal
    st = svmtest.seeded_state(512, 1137)
    om = svmtest.OracleModel(st)
    ctx = make_ctx([st])
    ctx.set_vicsek(0, dt=0.01, mu=1.0, v0=0.1, tau=1.0, eta=0.5)
    for chunk in range(3):
        om.vicsek(40, dt=0.01, mu=1.0, v0=0.1, tau=1.0, eta=0.5, seed=1137, step0=40 * chunk)
        ctx.vicsek_steps(40)
        np.testing.assert_array_equal(ctx.positions(0), om.positions())
        np.testing.assert_array_equal(ctx.directors(0), om.directors())
        assert_topology_equal(ctx, 0, om)
    assert om.n_t1 > 0


def test_vicsek_batch_graph_equals_eager():
    states = [svmtest.seeded_state(n, 1137 + i) for i, n in enumerate((300, 900, 2000))]
    runs = []
    for graph in (False, True):
        ctx = make_ctx(states)
        for i in range(len(states)):
            ctx.set_vicsek(i, dt=0.01, mu=1.0, v0=0.05 * (i + 1), tau=0.5 * (i + 1), eta=0.3 * i)
        if graph:
            ctx.run_graph(2, 10, 4)
        else:
            ctx.vicsek_steps(40)
        runs.append([(ctx.positions(i), ctx.directors(i)) for i in range(len(states))])
    for (pa, na), (pb, nb) in zip(*runs):
        np.testing.assert_array_equal(pa, pb)
        np.testing.assert_array_equal(na, nb)
