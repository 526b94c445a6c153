"""Autotuner: every schedule gives the same trajectory, bit for bit, and the
tuner's steps are ordinary steps."""
import numpy as np
import pytest

import svmtest
from test_gpu_parity import make_ctx

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def states():
    return [svmtest.seeded_state(n, 1400 + i) for i, n in enumerate((300, 900, 2000))]


def run(states, steps, tuning=None):
    ctx = make_ctx(states, p0=3.6)
    for i in range(len(states)):
        ctx.set_active(i, dt=0.01, mu=1.0, v0=0.05, tau=1.0)
    if tuning:
        ctx.set_tuning(**tuning)
    ctx.run_graph(1, 0, steps // ctx.tuning()["graph_steps"])
    return ctx, [ctx.positions(i) for i in range(len(states))]


def test_schedules_give_identical_trajectories(states):
    _, ref = run(states, 40)
    for t in (dict(cluster_blocks=64, graph_steps=4, xcd_mapping=0),
              dict(cluster_blocks=1024, graph_steps=20, xcd_mapping=1),
              dict(cluster_blocks=128, graph_steps=8, xcd_mapping=0)):
        _, got = run(states, 40, t)
        for a, b in zip(got, ref):
            np.testing.assert_array_equal(a, b)


def test_autotune_steps_are_ordinary_steps(states):
    ctx = make_ctx(states, p0=3.6)
    for i in range(len(states)):
        ctx.set_active(i, dt=0.01, mu=1.0, v0=0.05, tau=1.0)
    tab = ctx.autotune(updater=1, steps_per_trial=30)
    # lanes 1, 2, 3 (capped at the 3 spheres) x 2 grids, 3 step forms, 2 geometry forms, 4 vertex forms,
    # 6 more kernel schedules, 3 graph lengths
    assert len(tab) == 24 and (tab[:, 7] > 0).all()
    np.testing.assert_array_equal(tab[:6, 0], [1, 1, 2, 2, 3, 3])
    np.testing.assert_array_equal(tab[:11, 4], 0)
    np.testing.assert_array_equal(tab[:6, 5], 0)
    np.testing.assert_array_equal(tab[:9, 6], 0)
    np.testing.assert_array_equal(tab[6:9, 5], [1, 2, 3])
    np.testing.assert_array_equal(tab[9:11, 6], [1, 2])
    np.testing.assert_array_equal(tab[11:15, 4], [1, 2, 3, 4])
    assert (tab[6:, 0] == tab[np.argmin(tab[:6, 7]), 0]).all()
    t = ctx.tuning()
    best = tab[np.argmin(tab[:, 7])]
    assert (t["lanes"], t["cluster_blocks"], t["graph_steps"], t["xcd_mapping"], t["vertex_form"],
            t["step_form"], t["geometry_form"]) == tuple(int(x) for x in best[:7])
    # the number of steps the tuner took, from the trial table: capture warm-up + 3 samples per candidate
    steps = 0
    for lanes, cb, G, x, form, sform, gform, ms in tab:
        G = int(G)
        steps += G + 3 * max(1, 30 // (3 * G)) * G
    ref = make_ctx(states, p0=3.6)
    for i in range(len(states)):
        ref.set_active(i, dt=0.01, mu=1.0, v0=0.05, tau=1.0)
    ref.active_steps(steps)
    for i in range(len(states)):
        np.testing.assert_array_equal(ctx.positions(i), ref.positions(i))


def test_persisted_schedule_round_trip(states, tmp_path):
    """A schedule saved under a batch's key is found and applied for a batch
    of the same key, not for another key, and graphs captured after it run
    the same trajectory."""
    path = str(tmp_path / "schedules.txt")
    a = make_ctx(states, p0=3.6)
    for i in range(len(states)):
        a.set_active(i, dt=0.01, mu=1.0, v0=0.05, tau=1.0)
    assert not a.load_tuning(path)                 # no file yet
    a.set_lanes(2)
    a.set_tuning(cluster_blocks=512, graph_steps=4, xcd_mapping=0)
    a.set_vertex_form(2)
    a.set_step_form(3)
    a.set_geometry_form(2)
    a.save_tuning(path)
    a.save_tuning(path, updater=0)                 # a second key in the same file
    a.save_tuning(path)                            # replaces its own line
    lines = open(path).read().splitlines()
    assert len(lines) == 2 and all(l.startswith("svm-schedule-v2 ") for l in lines)
    b = make_ctx(states, p0=3.6)
    for i in range(len(states)):
        b.set_active(i, dt=0.01, mu=1.0, v0=0.05, tau=1.0)
    assert b.load_tuning(path)
    assert b.tuning() == dict(cluster_blocks=512, graph_steps=4, xcd_mapping=0, lanes=2, vertex_form=2,
                              step_form=3, geometry_form=2)
    # a line of the previous format (no step or geometry form) reads them as 0
    old = str(tmp_path / "schedules_v1.txt")
    with open(old, "w") as f:
        for l in lines:
            w = l.split()
            f.write(" ".join(["svm-schedule-v1"] + w[1:10] + w[12:]) + "\n")
    c = make_ctx(states, p0=3.6)
    c.set_step_form(1)
    c.set_geometry_form(1)
    assert c.load_tuning(old)
    assert c.tuning() == dict(cluster_blocks=512, graph_steps=4, xcd_mapping=0, lanes=2, vertex_form=2,
                              step_form=0, geometry_form=0)
    b.set_force_precision(1)                       # another key: FP32 forces
    assert not b.load_tuning(path)
    b.set_force_precision(0)
    b.run_graph(1, 0, 10)
    _, ref = run(states, 40)
    for i in range(len(states)):
        np.testing.assert_array_equal(b.positions(i), ref[i])
