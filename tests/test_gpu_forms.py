"""Kernel forms of the vertex and geometry passes give identical trajectories.

svm_set_vertex_form selects how the Brownian and active steps evaluate the
chain-rule forces: a thread per vertex (form 0, form 3 compiled for three
wavefronts per SIMD, form 4 with the tile's cell and vertex windows staged in
LDS) or the corner form, one wavefront per cell slot (forms 1 and 2, compiled
for four and five wavefronts per SIMD).
svm_set_geometry_form gives the geometry pass's threads 1, 2 or 4 cells each
(forms 0, 1, 2; form 0 reads the first eight ring slots of its cell at once).
Every form must give the oracle's trajectory bit for bit, over steps with T1
flips, NaN-free and with noise.
"""
import numpy as np
import pytest

import svmtest
import curvedspacevertexmodel_amd as svm

pytestmark = pytest.mark.gpu


def batch(params, form, lanes=1, gform=0):
    ctx = svm.Context([dict(n_cells=n, a0=1.0, p0=p0, kr=1.0, t1_threshold=0.1, seed=1137 + i)
                       for i, (n, p0) in enumerate(params)])
    for i, (n, p0) in enumerate(params):
        pts = svm.generate_cell_points(n, 1137 + i)
        ctx.build_from_cell_points(i, pts[svm.spatial_order(pts)])
        ctx.set_directors(i, svm.generate_directors(ctx.positions(i), 1137 + i))
    ctx.set_vertex_form(form)
    ctx.set_geometry_form(gform)
    ctx.set_lanes(lanes)
    assert ctx.vertex_form() == form and ctx.geometry_form() == gform
    return ctx


def test_forms_agree_on_a_batch():
    rng = np.random.default_rng(5)
    params = [(int(n), float(p)) for n, p in zip(np.exp(rng.uniform(np.log(200), np.log(6000), 24)).round(),
                                                  rng.uniform(3.5, 4.0, 24))]
    out = []
    for form, lanes, gform in ((0, 1, 0), (1, 1, 0), (2, 3, 0), (1, 3, 0), (3, 1, 0), (0, 1, 1), (3, 3, 2), (0, 2, 2),
                               (4, 1, 0), (4, 3, 1)):
        ctx = batch(params, form, lanes, gform)
        for i in range(len(params)):
            ctx.set_preferred_parameters(i, 1.0, 1.0)
            ctx.set_brownian(i, dt=0.01, mu=1.0, T=0.002)
        ctx.run_graph(0, 10, 6)                    # Brownian with noise, graphs
        for i, (n, p0) in enumerate(params):
            ctx.set_preferred_parameters(i, 1.0, p0)
            ctx.set_active(i, dt=0.01, mu=1.0, v0=0.3, tau=1.0)
        ctx.active_steps(40)                       # eager
        ctx.run_graph(1, 10, 4)                    # graphs
        out.append(([ctx.positions(i) for i in range(len(params))],
                    [ctx.directors(i) for i in range(len(params))], ctx.counters()))
    assert out[0][2][:, 0].sum() > 0              # T1 flips happened
    for pos, dirs, cnt in out[1:]:
        for i in range(len(params)):
            np.testing.assert_array_equal(pos[i], out[0][0][i])
            np.testing.assert_array_equal(dirs[i], out[0][1][i])
        np.testing.assert_array_equal(cnt, out[0][2])


@pytest.mark.parametrize("form", [1, 2, 3, 4])
def test_vertex_forms_match_the_oracle(form):
    st = svmtest.seeded_state(512, 1137)
    om = svmtest.OracleModel(st)
    ctx = svm.Context([dict(n_cells=512, a0=1.0, p0=3.8, kr=1.0, t1_threshold=0.1, seed=1137)])
    ctx.set_topology(0, st["vpos"], st["vn"], st["vcn"], st["ring"], st["cnn"])
    ctx.set_directors(0, st["directors"])
    ctx.set_vertex_form(form)
    ctx.set_active(0, dt=0.01, mu=1.0, v0=0.1, tau=1.0)
    ctx.active_steps(100)
    om.active(100, dt=0.01, mu=1.0, v0=0.1, tau=1.0, seed=1137, step0=0)
    np.testing.assert_array_equal(ctx.positions(0), om.positions())
    np.testing.assert_array_equal(ctx.directors(0), om.directors())
    assert ctx.counters()[0, 2] == om.L.or_num_nan(om.h)
    ctx.set_brownian(0, dt=0.01, mu=1.0, T=0.01)
    ctx.brownian_steps(50)
    om.brownian(50, dt=0.01, mu=1.0, T=0.01, seed=1137, step0=100)
    np.testing.assert_array_equal(ctx.positions(0), om.positions())


@pytest.mark.parametrize("gform", [0, 1, 2])
def test_geometry_forms_match_the_oracle(gform):
    """Areas, perimeters and cell centres of the geometry pass with 1, 2 and
    4 cells per thread, then active steps, against the oracle."""
    st = svmtest.seeded_state(700, 1140)
    om = svmtest.OracleModel(st)
    ctx = svm.Context([dict(n_cells=700, a0=1.0, p0=3.8, kr=1.0, t1_threshold=0.1, seed=1140)])
    ctx.set_topology(0, st["vpos"], st["vn"], st["vcn"], st["ring"], st["cnn"])
    ctx.set_directors(0, st["directors"])
    ctx.set_geometry_form(gform)
    ctx.compute_forces()
    om.compute_forces()
    np.testing.assert_array_equal(ctx.area_perimeter(0), om.area_perimeter())
    np.testing.assert_array_equal(ctx.cell_positions(0), om.cell_positions())
    np.testing.assert_array_equal(ctx.forces(0), om.forces())
    assert ctx.energies()[0][0] == om.energy(mode=1)
    ctx.set_active(0, dt=0.01, mu=1.0, v0=0.1, tau=1.0)
    ctx.active_steps(60)
    om.active(60, dt=0.01, mu=1.0, v0=0.1, tau=1.0, seed=1140, step0=0)
    np.testing.assert_array_equal(ctx.positions(0), om.positions())
    np.testing.assert_array_equal(ctx.directors(0), om.directors())
