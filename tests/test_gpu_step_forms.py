"""Launch structures of the step (svm_set_step_form) give identical results.

Flag 1 moves the geometry's total-area check and its triangle-fan fallback
from a launch of their own into the vertex pass; flag 2 closes the T1 step
without the group kernel (blocked candidates checked in the commit kernel,
overlapping dry runs redone by the sequential pass).  Every form must give
the same trajectories, topology and physics counters, bit for bit, and the
oracle's where the oracle is run.
"""
import numpy as np
import pytest

import svmtest
import curvedspacevertexmodel_amd as svm
from test_gpu_parity import _t1_batch_run, folded_positions, make_ctx

pytestmark = pytest.mark.gpu


def batch(params, step_form, lanes):
    ctx = svm.Context([dict(n_cells=n, a0=1.0, p0=p0, kr=1.0, t1_threshold=0.1, seed=3100 + i)
                       for i, (n, p0) in enumerate(params)])
    for i, (n, p0) in enumerate(params):
        pts = svm.generate_cell_points(n, 3100 + i)
        ctx.build_from_cell_points(i, pts[svm.spatial_order(pts)])
        ctx.set_directors(i, svm.generate_directors(ctx.positions(i), 3100 + i))
    ctx.set_step_form(step_form)
    ctx.set_lanes(lanes)
    assert ctx.step_form() == step_form
    return ctx


def test_step_forms_agree_on_a_batch():
    """Brownian steps with noise from raw starts (T1 transient: overlapping
    dry runs and group runs), then active steps eager and from graphs."""
    rng = np.random.default_rng(11)
    params = [(int(n), float(p)) for n, p in zip(np.exp(rng.uniform(np.log(200), np.log(6000), 24)).round(),
                                                  rng.uniform(3.5, 4.0, 24))]
    out = []
    for sform, lanes in ((0, 1), (1, 1), (2, 1), (3, 1), (3, 3), (2, 2)):
        ctx = batch(params, sform, lanes)
        for i in range(len(params)):
            ctx.set_preferred_parameters(i, 1.0, 1.0)
            ctx.set_brownian(i, dt=0.01, mu=1.0, T=0.002)
        ctx.run_graph(0, 10, 6)
        for i, (n, p0) in enumerate(params):
            ctx.set_preferred_parameters(i, 1.0, p0)
            ctx.set_active(i, dt=0.01, mu=1.0, v0=0.3, tau=1.0)
        ctx.active_steps(40)
        ctx.run_graph(1, 10, 4)
        out.append(([ctx.positions(i) for i in range(len(params))],
                    [ctx.directors(i) for i in range(len(params))],
                    [ctx.get_topology(i, svmtest.RING_CAP) for i in range(len(params))], ctx.counters()))
    c0 = out[0][3]
    assert c0[:, 0].sum() > 0 and c0[:, 5].sum() > 0      # flips, and group runs in form 0
    for pos, dirs, topo, cnt in out[1:]:
        for i in range(len(params)):
            np.testing.assert_array_equal(pos[i], out[0][0][i])
            np.testing.assert_array_equal(dirs[i], out[0][1][i])
            for a, b in zip(topo[i], out[0][2][i]):
                np.testing.assert_array_equal(a, b)
        # flips, fan fallbacks, NaN contributions, skipped flips
        np.testing.assert_array_equal(cnt[:, :4], c0[:, :4])
        # a group run of form 0 is a sequential pass without the group kernel
        np.testing.assert_array_equal(cnt[:, 4] + cnt[:, 5], c0[:, 4] + c0[:, 5])


@pytest.fixture(scope="module")
def state512():
    return svmtest.seeded_state(512, 1137)


@pytest.mark.parametrize("step_form", [0, 1, 3])
def test_fan_check_in_the_step_matches_oracle(state512, step_form):
    """One active step from a folded sphere (the middle one of three): the
    area check fails at the step's geometry, every cell of that sphere takes
    its fan area before the forces, exactly as in the oracle's step, whether
    the fallback is its own launch or inside the vertex pass."""
    folded, _ = folded_positions(state512)
    om = svmtest.OracleModel(state512)
    om.set_positions(folded)
    ctx = make_ctx([state512, state512, state512])
    ctx.set_step_form(step_form)
    ctx.set_positions(1, folded)
    for i in range(3):
        ctx.set_active(i, dt=0.01, mu=1.0, v0=0.1, tau=1.0)
    c0 = ctx.counters()[:, 1].copy()
    ctx.active_steps(1)
    om.active(1, dt=0.01, mu=1.0, v0=0.1, tau=1.0, seed=1138, step0=0)
    np.testing.assert_array_equal(ctx.counters()[:, 1] - c0, [0, 1, 0])
    np.testing.assert_array_equal(ctx.positions(1), om.positions())
    np.testing.assert_array_equal(ctx.directors(1), om.directors())
    # the areas the step's forces used: the fan areas of the folded sphere
    geo = svmtest.OracleModel(state512)
    geo.set_positions(folded)
    geo.compute_geometry()
    np.testing.assert_array_equal(ctx.area_perimeter(1), geo.area_perimeter())
    plain = svmtest.OracleModel(state512)
    plain.active(1, dt=0.01, mu=1.0, v0=0.1, tau=1.0, seed=1137, step0=0)
    np.testing.assert_array_equal(ctx.positions(0), plain.positions())


@pytest.fixture(scope="module")
def t1_batch_reference():
    return _t1_batch_run(True)


def test_no_group_kernel_matches_sequential(t1_batch_reference):
    """One dry run per candidate (component distance 0) overlaps often; with
    step form 2 those sphere-steps take the sequential pass instead of group
    runs, with the sequential pass's trajectory and topology."""
    seq = t1_batch_reference
    run = _t1_batch_run(False, 0.0, step_form=2)
    for i in range(len(seq[0])):
        np.testing.assert_array_equal(run[0][i], seq[0][i])
        for a, b in zip(run[1][i], seq[1][i]):
            np.testing.assert_array_equal(a, b)
    c0, c1 = run[2], seq[2]
    np.testing.assert_array_equal(c0[:, 0], c1[:, 0])
    assert c0[:, 0].sum() > 2000
    assert c0[:, 5].sum() == 0                    # no group runs
    assert 0 < c0[:, 4].sum() < c1[:, 4].sum()    # some sequential passes, far fewer than mode 1


def test_fire_with_step_forms_matches_oracle(state512):
    """FIRE's T1 passes close without the group kernel under flag 2."""
    fp = dict(dt=0.01, alpha_start=0.99, dt_max=0.5, dt_inc=1.1, dt_dec=0.95, alpha_dec=0.9, n_min=4,
              force_cutoff=1e-12, alpha_min=0.0, max_iterations=600)
    om = svmtest.OracleModel(state512, p0=3.6)
    ores = om.fire(**fp)
    ctx = make_ctx([state512], p0=3.6)
    ctx.set_step_form(3)
    ctx.fire_set_parameters(0, **fp)
    ctx.fire_minimize(check_every=64)
    st = ctx.fire_state(0)
    assert st["iterations"] == ores["iterations"] and st["force_max"] == ores["force_max"]
    np.testing.assert_array_equal(ctx.positions(0), om.positions())
