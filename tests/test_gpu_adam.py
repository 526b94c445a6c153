"""Adam minimiser (energyMinimizerAdam, src/updaters/energyMinimizerAdam.cpp;
the CPU step is the reference) on the GPU against the oracle.

Bar: bit-exact positions, topology, iteration counts, force statistic and
bias-correction powers against the oracle with the device's fixed-tree force
sum (reduce_mode 1), over runs with hundreds of T1 flips inside the
minimisation; a batch of spheres with their own parameters equals each sphere
run alone, whatever the host's check interval; two calls with a raised
iteration limit equal one call.
"""
import numpy as np
import pytest

import svmtest
from test_gpu_parity import assert_topology_equal, make_ctx

pytestmark = pytest.mark.gpu


def test_adam_matches_oracle():
    st = svmtest.seeded_state(512, 1137)
    ctx = make_ctx([st], p0=3.9)
    ctx.adam_set_parameters(0, dt=5e-3, max_iterations=300)
    ctx.adam_minimize(check_every=40)
    s = ctx.adam_state(0)
    om = svmtest.OracleModel(st, p0=3.9)
    r = om.adam(dt=5e-3, max_iterations=300, reduce_mode=1)
    assert s["iterations"] == r["iterations"] == 300
    assert s["force_max"] == r["force_max"]
    assert s["beta1t"] == r["beta1t"] and s["beta2t"] == r["beta2t"]
    np.testing.assert_array_equal(ctx.positions(0), om.positions())
    assert_topology_equal(ctx, 0, om)
    assert om.n_t1 > 100
    assert int(ctx.counters()[0, 0]) == om.n_t1


def test_adam_stops_at_the_force_cutoff():
    st = svmtest.seeded_state(200, 1139)
    ctx = make_ctx([st], p0=3.7)
    ctx.adam_set_parameters(0, dt=1e-3, force_cutoff=2e-3, max_iterations=5000)
    ctx.adam_minimize(check_every=64)
    s = ctx.adam_state(0)
    om = svmtest.OracleModel(st, p0=3.7)
    r = om.adam(dt=1e-3, force_cutoff=2e-3, max_iterations=5000, reduce_mode=1)
    assert s["iterations"] == r["iterations"] < 5000
    assert s["converged"] and s["force_max"] == r["force_max"] <= 2e-3
    np.testing.assert_array_equal(ctx.positions(0), om.positions())


def test_adam_batch_equals_single_spheres_and_continues():
    states = [svmtest.seeded_state(n, 1137 + i) for i, n in enumerate((300, 700, 1200))]
    pars = [dict(dt=2e-3, max_iterations=150), dict(dt=5e-3, beta2=0.999, max_iterations=220),
            dict(dt=1e-3, max_iterations=90)]
    ctx = make_ctx(states, p0=3.85)
    for i, p in enumerate(pars):
        ctx.adam_set_parameters(i, **p)
    ctx.adam_minimize(check_every=24)       # even: replayed from a captured graph
    for i, st in enumerate(states):
        one = make_ctx([st], p0=3.85, seeds=[1137 + i])
        one.adam_set_parameters(0, **pars[i])
        one.adam_minimize(check_every=7)    # odd: eager launches
        np.testing.assert_array_equal(ctx.positions(i), one.positions(0))
        assert ctx.adam_state(i) == one.adam_state(0)
        assert ctx.adam_state(i)["iterations"] == pars[i]["max_iterations"]
    # 90 iterations, then the limit raised to 150: the same as 150 at once
    two = make_ctx([states[0]], p0=3.85)
    two.adam_set_parameters(0, dt=2e-3, max_iterations=90)
    two.adam_minimize()
    two.adam_set_max_iterations(0, 150)
    two.adam_minimize()
    np.testing.assert_array_equal(two.positions(0), ctx.positions(0))
    assert two.adam_state(0) == ctx.adam_state(0)
