"""GPU parity: the HIP library against the CPU oracle on identical seeded inputs.

Bar (FP64 parity mode): topology bit-exact; areas, perimeters, cell centres,
forces, positions, directors and velocities bit-exact (assert_array_equal),
which is stricter than the 1e-10 relative the path promises.  Energies are
compared against the oracle's tree-ordered sum bit-exactly and against the
reference's sequential sum within 1e-12 relative.
"""
import numpy as np
import pytest

import svmtest
import curvedspacevertexmodel_amd as svm

pytestmark = pytest.mark.gpu


def make_ctx(topos, p0=3.8, kr=1.0, seeds=None):
    systems = []
    for i, t in enumerate(topos):
        systems.append(dict(n_cells=t["cnn"].shape[0], a0=1.0, p0=p0, kr=kr, t1_threshold=0.1,
                            seed=(seeds[i] if seeds else 1137 + i)))
    ctx = svm.Context(systems)
    for i, t in enumerate(topos):
        ctx.set_topology(i, t["vpos"], t["vn"], t["vcn"], t["ring"], t["cnn"])
        if "directors" in t:
            ctx.set_directors(i, t["directors"])
    return ctx


def assert_topology_equal(ctx, sys, om):
    vn, vcn, ring, cnn = ctx.get_topology(sys, svmtest.RING_CAP)
    ovn, ovcn, oring, ocnn = om.topology(svmtest.RING_CAP)
    np.testing.assert_array_equal(cnn, ocnn)
    np.testing.assert_array_equal(vn, ovn)
    np.testing.assert_array_equal(vcn, ovcn)
    for c in range(len(cnn)):
        np.testing.assert_array_equal(ring[c, :cnn[c]], oring[c, :ocnn[c]])


@pytest.fixture(scope="module")
def state512():
    return svmtest.seeded_state(512, 1137)


def test_geometry_forces_n512(state512):
    om = svmtest.OracleModel(state512)
    om.compute_forces()
    ctx = make_ctx([state512])
    ctx.compute_forces()
    np.testing.assert_array_equal(ctx.area_perimeter(0), om.area_perimeter())
    np.testing.assert_array_equal(ctx.cell_positions(0), om.cell_positions())
    np.testing.assert_array_equal(ctx.forces(0), om.forces())
    pe, ke = ctx.energies()
    assert pe[0] == om.energy(mode=1)
    assert abs(pe[0] - om.energy(mode=0)) <= 1e-12 * abs(om.energy(mode=0))
    # sum of areas is the sphere's area
    assert abs(ctx.area_perimeter(0)[:, 0].sum() - 512.0) < 1e-9


def folded_positions(state):
    """Positions of `state` with vertex 17 pushed across its neighbours by the
    first overshoot that breaks the angle-sum tiling (found on the CPU), and
    the oracle model at those positions after its geometry pass."""
    base = svmtest.OracleModel(state).positions()
    v = 17
    nb = state["vn"][v]
    for f in (3.0, 4.0, 5.0, 6.0):
        p = base.copy()
        q = p[v] + f * (p[nb].mean(axis=0) - p[v])
        p[v] = np.linalg.norm(p[v]) * q / np.linalg.norm(q)
        om = svmtest.OracleModel(state)
        om.set_positions(p)
        n0 = om.n_fan
        om.compute_geometry()
        if om.n_fan > n0:
            return p, om
    raise AssertionError("no overshoot folded a cell")


def test_fan_fallback_matches_oracle(state512):
    """A vertex pushed far across its neighbours folds its cells: the
    angle-sum areas no longer tile the sphere, and the oracle and the device
    both recompute every cell of that sphere as a triangle fan
    (recomputeAreasCPU), bit for bit, counting one fallback.  In a batch of
    three spheres where only the middle one is folded, the other two keep
    their angle-sum areas."""
    folded, om = folded_positions(state512)
    ctx = make_ctx([state512, state512, state512])
    ctx.set_positions(1, folded)
    c0 = ctx.counters()[:, 1].copy()
    ctx.compute_geometry()
    np.testing.assert_array_equal(ctx.counters()[:, 1] - c0, [0, 1, 0])
    np.testing.assert_array_equal(ctx.area_perimeter(1), om.area_perimeter())
    plain = svmtest.OracleModel(state512)
    plain.compute_geometry()
    for i in (0, 2):
        np.testing.assert_array_equal(ctx.area_perimeter(i), plain.area_perimeter())


def test_active_steps_n512(state512):
    om = svmtest.OracleModel(state512)
    ctx = make_ctx([state512])
    ctx.set_active(0, dt=0.01, mu=1.0, v0=0.1, tau=1.0)
    for chunk in range(4):
        om.active(50, dt=0.01, mu=1.0, v0=0.1, tau=1.0, seed=1137, step0=50 * chunk)
        ctx.active_steps(50)
        np.testing.assert_array_equal(ctx.positions(0), om.positions())
        np.testing.assert_array_equal(ctx.directors(0), om.directors())
        assert_topology_equal(ctx, 0, om)
    c = ctx.counters()[0]
    assert c[0] == om.n_t1
    assert om.n_t1 > 0, "the run should exercise T1 flips"
    # most steps' flips were applied from independent dry runs
    assert c[4] + c[5] < 200


def test_brownian_noise_steps_n512(state512):
    om = svmtest.OracleModel(state512, p0=3.9)
    ctx = make_ctx([state512], p0=3.9)
    ctx.set_brownian(0, dt=0.01, mu=1.0, T=0.05)
    om.brownian(60, dt=0.01, mu=1.0, T=0.05, seed=1137, step0=0)
    ctx.brownian_steps(60)
    np.testing.assert_array_equal(ctx.positions(0), om.positions())
    assert_topology_equal(ctx, 0, om)
    assert ctx.counters()[0, 0] == om.n_t1


def test_move_particles_triggers_t1(state512):
    rng = np.random.default_rng(7)
    om = svmtest.OracleModel(state512)
    ctx = make_ctx([state512])
    disp = rng.normal(scale=0.15, size=(state512["vpos"].shape[0], 3))
    om.move(disp, 1.0)
    ctx.move_particles(0, disp, 1.0)
    np.testing.assert_array_equal(ctx.positions(0), om.positions())
    assert_topology_equal(ctx, 0, om)
    om.move(disp, -0.5)
    ctx.move_particles(0, disp, -0.5)
    np.testing.assert_array_equal(ctx.positions(0), om.positions())
    assert_topology_equal(ctx, 0, om)
    assert om.n_t1 > 10
    c = ctx.counters()[0]
    assert c[4] + c[5] >= 1       # overlapping candidates: group run or sequential pass


def test_t1_paths_agree(state512):
    """Dry runs (+ group runs) and the always-sequential pass give the same trajectory."""
    runs = []
    for sequential in (False, True):
        ctx = make_ctx([state512], p0=3.9)
        ctx.set_t1_mode(sequential)
        ctx.set_active(0, dt=0.01, mu=1.0, v0=0.3, tau=1.0)
        ctx.active_steps(150)
        runs.append((ctx.positions(0), ctx.get_topology(0, svmtest.RING_CAP), ctx.counters()[0]))
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    for a, b in zip(runs[0][1], runs[1][1]):
        np.testing.assert_array_equal(a, b)
    assert runs[0][2][0] == runs[1][2][0] > 0
    # every step with candidates went through the sequential pass in mode 1,
    # far fewer in mode 0
    assert runs[1][2][4] > runs[0][2][4]


def _t1_batch_run(sequential, distance=None, step_form=0):
    """48 spheres of 300-6000 cells near the fluid transition: 50 Brownian
    stabilisation steps, then 200 active steps, on one T1 path."""
    g = np.random.Generator(np.random.Philox(4242))
    params = [(int(round(np.exp(np.log(300) + u[0] * np.log(20)))), 3.75 + 0.2 * u[1]) for u in g.random((48, 2))]
    ctx = svm.Context([dict(n_cells=n, a0=1.0, p0=p0, kr=1.0, t1_threshold=0.1, seed=2000 + i)
                       for i, (n, p0) in enumerate(params)])
    for i, (n, _) in enumerate(params):
        pts = svm.generate_cell_points(n, 2000 + i)
        ctx.build_from_cell_points(i, pts[svm.spatial_order(pts)])
        ctx.set_directors(i, svm.generate_directors(ctx.positions(i), 2000 + i))
        ctx.set_preferred_parameters(i, 1.0, 1.0)
        ctx.set_brownian(i, dt=0.01, mu=1.0, T=0.0)
    ctx.set_t1_mode(sequential)
    ctx.set_step_form(step_form)
    if distance is not None:
        ctx.set_t1_component_distance(distance)
        assert ctx.t1_component_distance() == distance
    ctx.brownian_steps(50)
    for i, (n, p0) in enumerate(params):
        ctx.set_preferred_parameters(i, 1.0, p0)
        ctx.set_active(i, dt=0.01, mu=1.0, v0=0.3, tau=1.0)
    ctx.active_steps(200)
    return ([ctx.positions(i) for i in range(len(params))],
            [ctx.get_topology(i, svmtest.RING_CAP) for i in range(len(params))], ctx.counters(),
            ctx.t1_diagnostics())


@pytest.fixture(scope="module")
def t1_batch_sequential():
    return _t1_batch_run(True)


@pytest.mark.parametrize("distance", [0.0, 3.0, 12.0])
def test_t1_paths_agree_on_a_batch(t1_batch_sequential, distance):
    """The same at batch scale: the dry runs over predicted components, group
    runs and their LDS fast paths (mode 0) and the always-sequential pass on
    global memory (mode 1) give bit-identical positions and topology.  The
    component distance moves work between the kinds of run -- 0: one run per
    candidate, overlaps resolved by group runs; 3 (the default): few group
    runs; 12: components too large for a run's cache, redone by group runs --
    and must change nothing.  Each run includes thousands of flips."""
    seq = t1_batch_sequential
    run = _t1_batch_run(False, distance)
    for i in range(len(seq[0])):
        np.testing.assert_array_equal(run[0][i], seq[0][i])
        for a, b in zip(run[1][i], seq[1][i]):
            np.testing.assert_array_equal(a, b)
    c0, c1 = run[2], seq[2]
    np.testing.assert_array_equal(c0[:, 0], c1[:, 0])
    dg = run[3].sum(axis=0)
    print(f"distance {distance}: flips {int(c0[:, 0].sum())}, group sphere-steps {int(c0[:, 5].sum())}, "
          f"group members {int(dg[6])}, sequential sphere-steps {int(c0[:, 4].sum())} (mode 0) vs "
          f"{int(c1[:, 4].sum())} (mode 1)")
    assert c0[:, 0].sum() > 2000
    assert c0[:, 5].sum() > 10


def test_fire_n512(state512):
    om = svmtest.OracleModel(state512, p0=3.6)
    ctx = make_ctx([state512], p0=3.6)
    fp = dict(dt=0.01, alpha_start=0.99, dt_max=0.5, dt_inc=1.1, dt_dec=0.95, alpha_dec=0.9, n_min=4,
              force_cutoff=1e-12, alpha_min=0.0, max_iterations=600)
    ores = om.fire(**fp)
    ctx.fire_set_parameters(0, **fp)
    ctx.fire_minimize(check_every=64)
    st = ctx.fire_state(0)
    assert st["iterations"] == ores["iterations"]
    assert st["force_max"] == ores["force_max"]
    assert st["dt"] == ores["dt"] and st["alpha"] == ores["alpha"]
    np.testing.assert_array_equal(ctx.positions(0), om.positions())
    np.testing.assert_array_equal(ctx.velocities(0), om.velocities())
    assert_topology_equal(ctx, 0, om)


def test_batch_equivalence_and_graph(state512):
    s2 = svmtest.seeded_state(300, 1138)
    s3 = svmtest.seeded_state(1200, 1139)
    batch = make_ctx([state512, s2, s3], seeds=[1137, 1138, 1139])
    for i in range(3):
        batch.set_active(i, dt=0.01, mu=1.0, v0=0.05 * (i + 1), tau=1.0)
    batch.run_graph(1, 10, 3)
    for i, s in enumerate([state512, s2, s3]):
        alone = make_ctx([s], seeds=[1137 + i])
        alone.set_active(0, dt=0.01, mu=1.0, v0=0.05 * (i + 1), tau=1.0)
        alone.active_steps(30)
        np.testing.assert_array_equal(batch.positions(i), alone.positions(0))
        np.testing.assert_array_equal(batch.directors(i), alone.directors(0))


def test_graph_replays_between_eager_work(state512):
    """Replays of one captured graph, with eager steps, state reads and
    synchronisations between them, each equal to the same steps taken
    eagerly (a replay must not depend on what ran before it), on 3 lanes."""
    s2 = svmtest.seeded_state(900, 1141)
    topos = [state512, s2, state512]

    def fresh():
        c = make_ctx(topos, seeds=[1137, 1141, 1139])
        for i in range(3):
            c.set_active(i, dt=0.01, mu=1.0, v0=0.1 * (i + 1), tau=1.0)
        return c

    g, e = fresh(), fresh()
    g.set_lanes(3)
    for eager_between in (0, 3, 1, 2, 0):
        g.run_graph(1, 8, 1)
        e.active_steps(8)
        for i in range(3):
            np.testing.assert_array_equal(g.positions(i), e.positions(i))
            np.testing.assert_array_equal(g.directors(i), e.directors(i))
        if eager_between:
            g.active_steps(eager_between)
            e.active_steps(eager_between)
        g.synchronize()
    np.testing.assert_array_equal(g.counters(), e.counters())


def test_topology_from_triangles_matches_oracle(state512):
    """The library's triangle -> topology builder follows the same conventions as the oracle's."""
    ctx = svm.Context([dict(n_cells=512, seed=1137)])
    tris = state512["tris"].copy()
    rng = np.random.default_rng(3)
    tris = tris[rng.permutation(len(tris))]          # any order and orientation is accepted
    flip = rng.random(len(tris)) < 0.5
    tris[flip] = tris[flip][:, [0, 2, 1]]
    ctx.set_topology_from_triangles(0, state512["points"], tris)
    np.testing.assert_array_equal(ctx.delaunay_triangles(0), state512["tris"])
    np.testing.assert_array_equal(ctx.positions(0), state512["vpos"])
    vn, vcn, ring, cnn = ctx.get_topology(0, svmtest.RING_CAP)
    np.testing.assert_array_equal(vn, state512["vn"])
    np.testing.assert_array_equal(vcn, state512["vcn"])
    np.testing.assert_array_equal(cnn, state512["cnn"])
    for c in range(512):
        np.testing.assert_array_equal(ring[c, :cnn[c]], state512["ring"][c, :cnn[c]])


def test_flip_without_fourth_cell_is_skipped():
    """Sphere 35 of the batched workload reaches, at step ~370, an edge whose two
    vertices share all three cells; that flip is skipped on both sides and the
    run stays bit-identical to the oracle with a valid topology."""
    n, p0, v0, dr, seed = 2842, 3.865664562197794, 0.4494973611929375, 1.0, 1137 + 35
    st = svmtest.seeded_state(n, seed)
    om = svmtest.OracleModel(st, p0=p0)
    ctx = make_ctx([st], p0=p0, seeds=[seed])
    ctx.set_active(0, dt=0.01, mu=1.0, v0=v0, tau=1.0 / dr)
    om.active(400, dt=0.01, mu=1.0, v0=v0, tau=1.0 / dr, seed=seed, step0=0)
    ctx.active_steps(400)
    ctx.synchronize()
    np.testing.assert_array_equal(ctx.positions(0), om.positions())
    assert_topology_equal(ctx, 0, om)
    skipped = om.L.or_num_t1_skipped(om.h)
    assert skipped >= 1
    assert ctx.counters()[0, 3] == skipped
    assert ctx.counters()[0, 0] == om.n_t1
