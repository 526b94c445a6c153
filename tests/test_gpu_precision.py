"""FP32 force variant: error against the FP64 oracle, reported by the library.

The FP32 path evaluates the reference's force expression trees in single
precision from FP64 positions, cell centres and (A - A0, P - P0).  Its error
is dominated by 1/sqrt(1 - cos^2) of short edges, a few 1e-6 relative on
these states; the bars below (rms 1e-4, max 1e-3 relative to the largest
force) leave an order of magnitude of room.
"""
import numpy as np
import pytest

import svmtest
import curvedspacevertexmodel_amd as svm
from test_gpu_parity import make_ctx

pytestmark = pytest.mark.gpu

RMS_TOL = 1e-4
MAX_TOL = 1e-3


@pytest.fixture(scope="module")
def states():
    return [svmtest.seeded_state(n, 1137 + i) for i, n in enumerate((512, 1500, 3000))]


def test_fp32_forces_against_oracle(states):
    ctx = make_ctx(states)
    ctx.compute_forces(precision=1)
    for i, st in enumerate(states):
        om = svmtest.OracleModel(st)
        om.compute_forces()
        f64 = om.forces()
        f32 = ctx.forces(i)
        d = np.linalg.norm(f32 - f64, axis=1)
        a = np.linalg.norm(f64, axis=1)
        rms = np.sqrt((d ** 2).sum() / (a ** 2).sum())
        mx = d.max() / a.max()
        print(f"sphere {i} (N={st['cnn'].shape[0]}): FP32 force error rms {rms:.3e}, max {mx:.3e}")
        assert 0.0 < rms < RMS_TOL
        assert mx < MAX_TOL
        # FP32 forces are not FP64 forces: the variant really runs in single precision
        assert not np.array_equal(f32, f64)


def test_force_error_report_matches_host_computation(states):
    ctx = make_ctx(states)
    rep = ctx.force_error()
    ctx.compute_forces(precision=0)
    f64 = [ctx.forces(i) for i in range(len(states))]
    ctx.compute_forces(precision=1)
    f32 = [ctx.forces(i) for i in range(len(states))]
    for i in range(len(states)):
        d = np.linalg.norm(f32[i] - f64[i], axis=1)
        a = np.linalg.norm(f64[i], axis=1)
        np.testing.assert_allclose(rep[i, 0], d.max() / a.max(), rtol=1e-12)
        np.testing.assert_allclose(rep[i, 1], np.sqrt((d ** 2).sum() / (a ** 2).sum()), rtol=1e-12)
        np.testing.assert_allclose(rep[i, 2], d.max(), rtol=1e-12)
        np.testing.assert_allclose(rep[i, 3], np.sqrt((a ** 2).sum() / len(a)), rtol=1e-12)
        assert rep[i, 1] < RMS_TOL


def test_fp32_dynamics_track_fp64(states):
    """Active steps with FP32 forces stay close to the FP64 trajectory over a
    short horizon (before T1 decisions can diverge)."""
    st = states[0]
    runs = []
    for prec in (0, 1):
        ctx = make_ctx([st])
        ctx.set_force_precision(prec)
        ctx.set_active(0, dt=0.01, mu=1.0, v0=0.1, tau=1.0)
        ctx.active_steps(5)
        runs.append(ctx.positions(0))
    dev = np.abs(runs[1] - runs[0]).max()
    assert 0.0 < dev < 1e-6
    # precision 0 restores the parity path bit for bit
    ctx = make_ctx([st])
    ctx.set_force_precision(1)
    ctx.set_force_precision(0)
    ctx.set_active(0, dt=0.01, mu=1.0, v0=0.1, tau=1.0)
    ctx.active_steps(5)
    np.testing.assert_array_equal(ctx.positions(0), runs[0])


def test_fp64_vector_form_against_numpy(states):
    """precision 2: the vector-form gradients in FP64 on the GPU equal the same
    formulas in numpy (tests/svmtest.vector_forces) to rounding."""
    ctx = make_ctx(states)
    ctx.compute_forces(precision=2)
    rep = ctx.force_error()
    for i, st in enumerate(states):
        om = svmtest.OracleModel(st)
        om.compute_forces()
        fv = svmtest.vector_forces(om.positions(), om.cell_positions(), om.area_perimeter(), st["vn"], st["vcn"],
                                   st["ring"], st["cnn"], st["R"])
        ctx.compute_forces(precision=2)
        g = ctx.forces(i)
        assert np.linalg.norm(g - fv) / np.linalg.norm(fv) < 1e-12
        # FP32 against the FP64 vector form, and the two FP64 forms against each other, on these small spheres
        assert rep[i, 5] < 1e-5
        assert rep[i, 6] < 1e-8
