"""CPU tests that pin the oracle's updaters and their random draws.

The active updater (sphericalSelfPropelledParticle::integrateEOMCPU,
src/updaters/sphericalSelfPropelledParticle.cpp:16-47) and the Brownian
updater (brownianDynamics::integrateEOMCPU, src/updaters/brownianDynamics.cpp:29-48)
are checked against what their single steps must be, recovered from the
positions before and after a step:

  * a move p -> p' = R (p + d_t) / |p + d_t| with d_t tangent at p gives back
    d_t = lambda p' - p with lambda = R^2 / (p . p') (p . (p + d_t) = R^2);
  * the active step with mu = 0 is the tangent move by dt v0 n;
  * its director is the unit tangent of n at p' rotated by Rodrigues' formula
    about the UNNORMALISED axis p' (|p'| = R, inc/functions.h:175-181), so with
    n' the unit tangent director and theta the angle, the result is
    cos(theta) n' + sin(theta) p' x n' with squared norm cos^2 + R^2 sin^2, and
    theta = U(-1/2, 1/2) sqrt(2 dt tau) is read back as
    atan2(v . (p' x n') / R^2, v . n');
  * the Brownian step is dt mu F + sqrt(2 dt mu T) xi per component.

A wrong amplitude (sqrt(2 dt / tau), sqrt(dt mu T)), Gaussian angles, a
normalised rotation axis or a missing mobility fails one of them.  The draws
themselves (uniform on [0, 1), standard normals by Box-Muller) are checked for
their distributions, so none of this retypes the oracle's code.
"""
import ctypes as C
import math

import numpy as np
import pytest

import svmtest

STREAM_BROWNIAN = 0x42524e31
STREAM_ACTIVE = 0x41435431


@pytest.fixture(scope="module")
def L():
    lib = svmtest.oracle_lib()
    lib.or_uniform2.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint32, C.POINTER(C.c_double)]
    lib.or_normal3.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.POINTER(C.c_double)]
    return lib


@pytest.fixture(scope="module")
def state():
    return svmtest.seeded_state(512, 1137)


def uniforms(L, seed, stream, step, n):
    out = np.zeros((n, 2))
    buf = (C.c_double * 2)()
    for i in range(n):
        L.or_uniform2(C.c_uint64(seed), C.c_uint32(stream), C.c_uint64(step), C.c_uint32(i), buf)
        out[i] = buf[0], buf[1]
    return out


def normals3(L, seed, step, n):
    out = np.zeros((n, 3))
    buf = (C.c_double * 3)()
    for i in range(n):
        L.or_normal3(C.c_uint64(seed), C.c_uint64(step), C.c_uint32(i), buf)
        out[i] = buf[0], buf[1], buf[2]
    return out


def tangent_displacement(p, q, R):
    """d_t with q = R (p + d_t) / |p + d_t| and d_t . p = 0."""
    lam = R * R / np.sum(p * q, 1)
    return lam[:, None] * q - p


def ks_uniform(u):
    x = np.sort(u)
    n = len(x)
    return max(np.max(np.arange(1, n + 1) / n - x), np.max(x - np.arange(n) / n))


def ks_normal(z):
    x = np.sort(z)
    n = len(x)
    cdf = 0.5 * (1.0 + np.vectorize(math.erf)(x / math.sqrt(2.0)))
    return max(np.max(np.arange(1, n + 1) / n - cdf), np.max(cdf - np.arange(n) / n))


# ------------------------------------------------------------ draws -------
def test_uniform_draws_are_uniform_on_the_unit_interval(L):
    u = uniforms(L, 1137, STREAM_ACTIVE, 3, 20000).ravel()
    assert u.min() >= 0.0 and u.max() < 1.0
    # Kolmogorov-Smirnov: the 0.1% critical value for n = 40000 is 1.95 / sqrt(n)
    assert ks_uniform(u) < 1.95 / math.sqrt(u.size)
    # every draw is a multiple of 2^-53 (53 random bits per double)
    assert np.all(u * 2.0 ** 53 == np.floor(u * 2.0 ** 53))
    # streams, steps and seeds give different draws
    assert not np.array_equal(u[:100], uniforms(L, 1137, STREAM_ACTIVE, 4, 50).ravel())
    assert not np.array_equal(u[:100], uniforms(L, 1138, STREAM_ACTIVE, 3, 50).ravel())


def test_normal_draws_are_standard_normals_by_box_muller(L):
    z = normals3(L, 1137, 5, 20000)
    for k in range(3):
        assert abs(z[:, k].mean()) < 4.0 / math.sqrt(len(z))
        assert abs(z[:, k].var() - 1.0) < 4.0 * math.sqrt(2.0 / len(z))
        assert ks_normal(z[:, k]) < 1.95 / math.sqrt(len(z))
    c = np.corrcoef(z.T)
    assert np.all(np.abs(c[np.triu_indices(3, 1)]) < 4.0 / math.sqrt(len(z)))
    # Box-Muller of the uniform pair (a, b) of the Brownian stream at draw 2v:
    # z0 = sqrt(-2 ln(1 - a)) cos(2 pi b), z1 = ... sin(2 pi b) (z1 is component y)
    ab = uniforms(L, 1137, STREAM_BROWNIAN, 5, 2 * 200)[0::2]
    r = np.sqrt(-2.0 * np.log(1.0 - ab[:, 0]))
    np.testing.assert_allclose(z[:200, 0], r * np.cos(2 * np.pi * ab[:, 1]), rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(z[:200, 1], r * np.sin(2 * np.pi * ab[:, 1]), rtol=1e-13, atol=1e-13)


# ------------------------------------------------------------ active ------
def test_active_step_without_mobility_is_the_tangent_move_along_the_director(state):
    R = state["R"]
    dt, v0, tau, seed = 0.01, 0.3, 4.0, 1137
    out = []
    for p0 in (3.5, 3.9):   # different forces must not matter when mu = 0
        om = svmtest.OracleModel(state, p0=p0, t1=0.0)
        p = om.positions()
        n = om.directors()
        om.active(1, dt=dt, mu=0.0, v0=v0, tau=tau, seed=seed, step0=0)
        out.append(om.positions())
    np.testing.assert_array_equal(out[0], out[1])
    # the seeded directors are unit tangents, so the move is exactly dt v0 n
    np.testing.assert_allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-15)
    d = tangent_displacement(p, out[0], R)
    np.testing.assert_allclose(d, dt * v0 * n, rtol=0, atol=1e-13)
    np.testing.assert_allclose(np.linalg.norm(out[0], axis=1), R, rtol=1e-15)


def test_active_director_rotation_closed_form(L, state):
    R = state["R"]
    assert R > 5   # the norm test separates an axis of length R from a unit axis
    dt, tau, seed, step = 0.01, 4.0, 1137, 7
    om = svmtest.OracleModel(state, p0=3.8, t1=0.0)
    n0 = om.directors()
    om.active(1, dt=dt, mu=1.0, v0=0.1, tau=tau, seed=seed, step0=step)
    p1 = om.positions()
    v = om.directors()
    # unit tangent of the old director at the new position
    nt = n0 - (np.sum(n0 * p1, 1) / np.sum(p1 * p1, 1))[:, None] * p1
    nt /= np.linalg.norm(nt, axis=1)[:, None]
    k = np.cross(p1, nt)
    theta = np.arctan2(np.sum(v * k, 1) / R ** 2, np.sum(v * nt, 1))
    u = uniforms(L, seed, STREAM_ACTIVE, step, len(v))[:, 0]
    expect = (u - 0.5) * np.sqrt(2.0 * dt * tau)
    np.testing.assert_allclose(theta, expect, rtol=0, atol=1e-12)
    np.testing.assert_allclose(np.sum(v * v, 1), np.cos(theta) ** 2 + R ** 2 * np.sin(theta) ** 2, rtol=1e-12)
    np.testing.assert_allclose(np.sum(v * p1, 1) / R, 0.0, atol=1e-12)
    # the angles fill the uniform interval of half-width sqrt(2 dt tau) / 2
    half = 0.5 * np.sqrt(2.0 * dt * tau)
    assert np.abs(theta).max() <= half * (1 + 1e-12) and np.abs(theta).max() > 0.98 * half
    assert abs(theta.var() / (2.0 * dt * tau / 12.0) - 1.0) < 0.15


def test_active_step_moves_by_the_director_before_its_renormalisation(state):
    """The displacement uses the director as stored (not renormalised), so a
    director of length 2 moves a vertex twice as far (the rotation that
    follows renormalises it)."""
    R = state["R"]
    om = svmtest.OracleModel(state, p0=3.8, t1=0.0)
    p = om.positions()
    n = om.directors() * 2.0
    om.L.or_set_directors(om.h, svmtest._dp(np.ascontiguousarray(n)))
    om.active(1, dt=0.01, mu=0.0, v0=0.2, tau=1.0, seed=3, step0=0)
    d = tangent_displacement(p, om.positions(), R)
    np.testing.assert_allclose(d, 0.01 * 0.2 * n, rtol=0, atol=1e-13)


# ------------------------------------------------------------ Brownian ----
def test_brownian_step_at_zero_temperature_is_mobility_times_force(state):
    R = state["R"]
    dt, mu = 0.01, 0.5
    om = svmtest.OracleModel(state, p0=3.8, t1=0.0)
    p = om.positions()
    om.compute_forces()
    F = om.forces()
    om.brownian(1, dt=dt, mu=mu, T=0.0, seed=1137, step0=0)
    d = tangent_displacement(p, om.positions(), R)
    np.testing.assert_allclose(d, dt * mu * F, rtol=0, atol=1e-13 * max(1.0, np.abs(F).max()))


def test_brownian_noise_amplitude_is_sqrt_2_dt_mu_T(L, state):
    R = state["R"]
    dt, mu, T, seed, step = 1e-3, 0.5, 2.0, 1137, 11
    om0 = svmtest.OracleModel(state, p0=3.8, t1=0.0)
    p = om0.positions()
    om0.brownian(1, dt=dt, mu=mu, T=0.0, seed=seed, step0=step)
    omT = svmtest.OracleModel(state, p0=3.8, t1=0.0)
    omT.brownian(1, dt=dt, mu=mu, T=T, seed=seed, step0=step)
    noise = tangent_displacement(p, omT.positions(), R) - tangent_displacement(p, om0.positions(), R)
    xi = normals3(L, seed, step, len(p))
    xt = xi - (np.sum(xi * p, 1) / np.sum(p * p, 1))[:, None] * p
    np.testing.assert_allclose(noise, np.sqrt(2.0 * dt * mu * T) * xt, rtol=0, atol=1e-12)
    # one-step mean squared tangent displacement: two tangent components of variance 2 dt mu T
    msd = np.mean(np.sum(noise * noise, 1))
    assert abs(msd / (2 * 2.0 * dt * mu * T) - 1.0) < 4.0 / np.sqrt(len(p))
