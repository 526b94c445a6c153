"""Lanes: the batch cut into sphere ranges whose steps run on streams of
their own.  Spheres are independent, so every lane count must give the
one-lane trajectory bit for bit -- positions, directors, topology and the T1
counters -- for eager steps and for graph replays, with T1 flips, group
runs and sphere sizes that differ between the lanes."""
import numpy as np
import pytest

import svmtest
import curvedspacevertexmodel_amd as svm

pytestmark = pytest.mark.gpu


def params(n_sys=14, seed=5151):
    g = np.random.Generator(np.random.Philox(seed))
    return [(int(round(np.exp(np.log(200) + u[0] * np.log(25)))), 3.7 + 0.25 * u[1], 0.05 + 0.4 * u[2])
            for u in g.random((n_sys, 3))]


def build(ps, lanes):
    ctx = svm.Context([dict(n_cells=n, a0=1.0, p0=p0, kr=1.0, t1_threshold=0.1, seed=3100 + i)
                       for i, (n, p0, v0) in enumerate(ps)])
    for i, (n, p0, v0) in enumerate(ps):
        pts = svm.generate_cell_points(n, 3100 + i)
        ctx.build_from_cell_points(i, pts[svm.spatial_order(pts)])
        ctx.set_directors(i, svm.generate_directors(ctx.positions(i), 3100 + i))
        ctx.set_preferred_parameters(i, 1.0, 1.0)
        ctx.set_brownian(i, dt=0.01, mu=1.0, T=0.0)
    if isinstance(lanes, int):
        ctx.set_lanes(lanes)
    else:
        ctx.set_lane_spheres(lanes)
    return ctx


def state(ctx, n):
    return ([ctx.positions(i) for i in range(n)], [ctx.directors(i) for i in range(n)],
            [ctx.get_topology(i, svmtest.RING_CAP) for i in range(n)], ctx.counters())


def assert_same(a, b):
    for x, y in zip(a[0], b[0]):
        np.testing.assert_array_equal(x, y)
    for x, y in zip(a[1], b[1]):
        np.testing.assert_array_equal(x, y)
    for tx, ty in zip(a[2], b[2]):
        for x, y in zip(tx, ty):
            np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(a[3], b[3])


def trajectory(ps, lanes, updater):
    ctx = build(ps, lanes)
    ctx.brownian_steps(30)          # eager, then a graph of the Brownian updater
    ctx.run_graph(0, 10, 2)
    for i, (n, p0, v0) in enumerate(ps):
        ctx.set_preferred_parameters(i, 1.0, p0)
        if updater == 2:
            ctx.set_vicsek(i, dt=0.01, mu=1.0, v0=v0, tau=1.0, eta=0.2)
        else:
            ctx.set_active(i, dt=0.01, mu=1.0, v0=v0, tau=1.0)
    steps = {1: ctx.active_steps, 2: ctx.vicsek_steps}[updater]
    steps(25)                       # odd: the graphs below start from the other buffers
    ctx.run_graph(updater, 6, 5)
    steps(7)
    ctx.run_graph(updater, 4, 3)
    ctx.synchronize()
    return ctx, state(ctx, len(ps))


@pytest.mark.parametrize("updater", [1, 2])
def test_lanes_give_the_one_lane_trajectory(updater):
    ps = params()
    _, ref = trajectory(ps, 1, updater)
    c = ref[3]
    print(f"updater {updater}: flips {int(c[:, 0].sum())}, group sphere-steps {int(c[:, 5].sum())}")
    assert c[:, 0].sum() > 200      # T1 flips in the run
    for lanes in (2, 3, 4):
        ctx, got = trajectory(ps, lanes, updater)
        n, first = ctx.lanes()
        assert n == lanes and first[0] == 0 and (np.diff(first) > 0).all()
        assert_same(got, ref)


def test_explicit_lane_ranges():
    """Lanes over caller-chosen sphere ranges (svm_set_lane_spheres) give the
    one-lane trajectory too; malformed ranges are refused."""
    ps = params()
    _, ref = trajectory(ps, 1, 1)
    for starts in ([0, 1], [0, 5, 6, 13], [0, 12]):
        ctx, got = trajectory(ps, starts, 1)
        n, first = ctx.lanes()
        assert n == len(starts) and list(first) == starts
        assert_same(got, ref)
    ctx = build(ps[:5], 1)
    for bad in ([1, 3], [0, 3, 3], [0, 5], [0, 1, 2, 3, 4], []):
        with pytest.raises(svm.SVMError):
            ctx.set_lane_spheres(bad)
    assert ctx.lanes()[0] == 1


def test_lane_partition_and_limits():
    ps = params(5, 77)
    ctx = build(ps, 1)
    assert ctx.lanes()[0] == 1
    ctx.set_lanes(4)
    n, first = ctx.lanes()
    assert n == 4 and first[0] == 0 and (np.diff(first) > 0).all() and first[-1] < 5
    small = build(ps[:3], 1)
    small.set_lanes(4)             # capped at one lane per sphere
    assert small.lanes()[0] == 3 and list(small.lanes()[1]) == [0, 1, 2]
    with pytest.raises(svm.SVMError):
        ctx.set_lanes(0)
    with pytest.raises(svm.SVMError):
        ctx.set_lanes(svm.MAX_LANES + 1)
    # about equal vertex counts per lane
    ctx.set_lanes(2)
    n, first = ctx.lanes()
    nv = np.array([2 * p[0] - 4 for p in ps])
    a, b = nv[:first[1]].sum(), nv[first[1]:].sum()
    assert max(a, b) <= nv.sum() - min(a, b)


def test_lanes_then_whole_batch_calls():
    """Calls that cover the whole batch (forces, FIRE) see the lanes' steps."""
    ps = params(6, 91)
    a, b = build(ps, 1), build(ps, 3)
    for ctx in (a, b):
        for i, (n, p0, v0) in enumerate(ps):
            ctx.set_preferred_parameters(i, 1.0, p0)
            ctx.set_active(i, dt=0.01, mu=1.0, v0=v0, tau=1.0)
        ctx.active_steps(12)
        ctx.run_graph(1, 4, 2)
        ctx.compute_forces()
    for i in range(len(ps)):
        np.testing.assert_array_equal(a.forces(i), b.forces(i))
        np.testing.assert_array_equal(a.area_perimeter(i), b.area_perimeter(i))


@pytest.mark.parametrize("check_every", [40, 7])
def test_minimisers_with_lanes(check_every):
    """FIRE and Adam over lanes (graph replays for an even check interval,
    eager iterations for an odd one): the one-lane result bit for bit, with
    spheres converging at different iterations."""
    ps = params(9, 313)
    out = {}
    for lanes in (1, 3):
        ctx = build(ps, lanes)
        ctx.brownian_steps(20)
        for i, (n, p0, v0) in enumerate(ps):
            ctx.set_preferred_parameters(i, 1.0, p0)
            ctx.fire_set_parameters(i, dt=0.01, max_iterations=400 + 50 * i, force_cutoff=1e-9)
        ctx.fire_minimize(check_every)
        fire = ([ctx.positions(i) for i in range(len(ps))], [ctx.fire_state(i) for i in range(len(ps))])
        for i in range(len(ps)):
            ctx.adam_set_parameters(i, dt=1e-3, max_iterations=150 + 20 * i)
        ctx.adam_minimize(check_every)
        adam = ([ctx.positions(i) for i in range(len(ps))], [ctx.adam_state(i) for i in range(len(ps))])
        out[lanes] = (fire, adam, ctx.counters())
    for k in (0, 1):
        for a, b in zip(out[1][k][0], out[3][k][0]):
            np.testing.assert_array_equal(a, b)
        assert out[1][k][1] == out[3][k][1]
    its = [st["iterations"] for st in out[1][0][1]]
    assert len(set(its)) > 1        # the spheres stop at different iterations
    np.testing.assert_array_equal(out[1][2], out[3][2])


def test_lanes_on_a_caller_stream():
    """Lanes forked from and joined into a stream the caller owns (torch's):
    the trajectory is the one-lane one, and a read that synchronises only the
    caller's stream already sees every lane's last step (the join)."""
    import torch
    ps = params(8, 4040)
    ref = build(ps, 1)
    got = build(ps, 3)
    st = torch.cuda.Stream()
    got.set_stream(st.cuda_stream)
    for ctx in (ref, got):
        for i, (n, p0, v0) in enumerate(ps):
            ctx.set_preferred_parameters(i, 1.0, p0)
            ctx.set_active(i, dt=0.01, mu=1.0, v0=v0, tau=1.0)
    ref.active_steps(9)
    ref.run_graph(1, 10, 3)
    got.active_steps(9)
    got.run_graph(1, 10, 3)
    for i in range(len(ps)):
        np.testing.assert_array_equal(got.positions(i), ref.positions(i))
        np.testing.assert_array_equal(got.directors(i), ref.directors(i))
    got.set_stream(0)


def test_lanes_at_the_bench_scale():
    """The bench's 256-sphere batch (982,401 cells): 3 lanes give the
    one-lane state of every sphere, bit for bit, after graph-replayed
    Brownian and active steps with T1 flips, and the same flips, fan
    fallbacks and skips.  Which T1 path a sphere's step took may differ: the
    candidate lists hold SVM_T1_MAXCL entries per lane, and in the first
    steps after a random start one lane's list overflows (those spheres take
    the sequential pass, which is exact too)."""
    import bench
    ps = bench.batch_params(256)
    states = []
    for lanes in (1, 3):
        ctx, _ = bench.build_batch(svm, ps)
        ctx.set_lanes(lanes)
        bench.stabilise(ctx, ps, 20)
        ctx.run_graph(1, 10, 4)
        ctx.synchronize()
        states.append(([ctx.positions(i) for i in range(len(ps))], [ctx.directors(i) for i in range(len(ps))],
                       ctx.counters()))
    (p1, d1, c1), (p3, d3, c3) = states
    for i in range(len(ps)):
        np.testing.assert_array_equal(p1[i], p3[i])
        np.testing.assert_array_equal(d1[i], d3[i])
    np.testing.assert_array_equal(c1[:, :4], c3[:, :4])
    assert c1[:, 0].sum() > 1000     # T1 flips in the run
