"""Warm-start network on the GPU against its CPU oracle (oracle/mlp_oracle.py).

Tolerances: the GPU accumulates bf16 products in FP32 on MFMA and uses a fast
exp in SiLU; the oracle sums in FP64 with exact exp.  Activations rounded to
bf16 can then land one bf16 step (2^-8 relative) apart on a few entries, so
outputs are compared norm-wise at 5e-3 and gradients at 2e-2; the bulk of
entries agree to ~1e-5.
"""
import numpy as np
import pytest

import svmtest
import curvedspacevertexmodel_amd as svm
import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def random_net(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    th = mo.init_theta(seed)
    P = mo.unpack(np.arange(mo.NP))
    for b in ("b1", "b2", "b3"):
        th[P[b]] = rng.normal(0, 0.1, 64)
    th[P["b4"][:2]] = rng.normal(0, 0.1, 2)
    mu = rng.normal(0, 1, 12).astype(np.float32)
    sg = rng.uniform(0.5, 2, 12).astype(np.float32)
    return th, mu, sg, scale


def test_forward_matches_oracle():
    th, mu, sg, sc = random_net(11, 0.8)
    net = svm.MLP()
    net.set_parameters(th, mu, sg, sc)
    rng = np.random.default_rng(2)
    X = np.zeros((4099, 16), np.float32)     # not a multiple of the 16-cell tile
    X[:, :12] = rng.normal(0, 1.5, (4099, 12))
    y = net.forward(X)
    yo, _ = mo.forward(th, mu, sg, sc, X)
    err = np.linalg.norm(y - yo) / np.linalg.norm(yo)
    close = np.mean(np.abs(y - yo) <= 1e-3 * (np.abs(yo) + 1e-2))
    print(f"forward: rel norm error {err:.2e}, entries within 1e-3: {close:.4f}")
    assert err < 5e-3
    assert close > 0.97
    # parameters round-trip
    th2, mu2, sg2, sc2 = net.get_parameters()
    np.testing.assert_array_equal(th2, th)
    np.testing.assert_array_equal(mu2, mu)
    assert sc2 == np.float32(sc)


def test_gradient_matches_oracle():
    th, mu, sg, sc = random_net(12, 1.3)
    net = svm.MLP()
    net.set_parameters(th, mu, sg, sc)
    rng = np.random.default_rng(4)
    n = 3000
    X = np.zeros((n, 16), np.float32)
    X[:, :12] = rng.normal(0, 1, (n, 12))
    T = rng.normal(0, 1, (n, 2)).astype(np.float32)
    g, loss = net.gradient(X, T)
    yo, cache = mo.forward(th, mu, sg, sc, X)
    go, lo = mo.backward(cache, yo, T)
    err = np.linalg.norm(g - go) / np.linalg.norm(go)
    print(f"gradient: rel norm error {err:.2e}; loss {loss:.6f} vs {lo:.6f}")
    assert err < 2e-2
    assert abs(loss - lo) <= 1e-3 * lo
    # the structural zeros get no gradient
    assert not g[~mo.structural_mask()].any()
    # per-layer agreement, not only overall
    P = mo.unpack(np.arange(mo.NP))
    for k in P:
        sl = P[k].ravel()
        m = mo.structural_mask()[sl]
        e = np.linalg.norm(g[sl][m] - go[sl][m]) / max(np.linalg.norm(go[sl][m]), 1e-30)
        assert e < 5e-2, (k, e)


def test_training_fits_a_smooth_map():
    rng = np.random.default_rng(7)
    n = 20000
    X = np.zeros((n, 16), np.float32)
    X[:, :12] = rng.normal(0, 1, (n, 12))
    A = rng.normal(0, 0.5, (2, 12))
    T = np.tanh(X[:, :12] @ A.T).astype(np.float32)
    net = svm.MLP()
    net.set_parameters(mo.init_theta(3), np.zeros(12), np.ones(12), 1.0)
    losses = net.train(X, T, epochs=12, batch=2048, lr=2e-3, seed=5)
    print("training losses", np.round(losses, 5))
    assert losses[-1] < 0.1 * losses[0]
    # and the trained network predicts held-out data as well as the training loss says
    Xh = np.zeros((4000, 16), np.float32)
    Xh[:, :12] = rng.normal(0, 1, (4000, 12))
    Th = np.tanh(Xh[:, :12] @ A.T)
    assert np.mean((net.forward(Xh) - Th) ** 2) < 3 * losses[-1] + 1e-3


def test_descriptors_against_numpy():
    st = svmtest.seeded_state(600, 1301)
    ctx = svm.Context([dict(n_cells=600, a0=1.0, p0=3.75, kr=1.0, t1_threshold=0.1, seed=1301)])
    ctx.set_topology(0, st["vpos"], st["vn"], st["vcn"], st["ring"], st["cnn"])
    X, F = ctx.cell_descriptors(0)
    ap = ctx.area_perimeter(0)
    C = ctx.cell_positions(0)
    f = ctx.forces(0)
    vn, vcn, ring, cnn = ctx.get_topology(0, 32)
    A, P = ap[:, 0], ap[:, 1]
    for c in range(0, 600, 37):
        n = cnn[c]
        r = ring[c, :n]
        nh = C[c] / np.linalg.norm(C[c])
        Fc = f[r].mean(0)
        Ft = Fc - (Fc @ nh) * nh
        e1 = Ft / np.linalg.norm(Ft)
        e2 = np.cross(nh, e1)
        nbs = []
        for k in range(n):
            a, b = set(vcn[r[k]]), set(vcn[r[(k + 1) % n]])
            nbs.append(((a & b) - {c}).pop())
        nbs = np.array(nbs)
        q = P[c] / np.sqrt(A[c])
        dv = C[nbs] - C[c]
        dv -= np.outer(dv @ nh, nh)
        u = dv / np.linalg.norm(dv, axis=1)[:, None]
        ref = [A[c] - 1.0, P[c] - 3.75, q, n - 6, np.linalg.norm(Ft), np.linalg.norm(f[r], axis=1).mean(),
               A[nbs].mean() - A[c], A[nbs].std(), ((A[nbs] - A[c]) * (u @ e1)).mean(),
               ((A[nbs] - A[c]) * (u @ e2)).mean(), (P[nbs] / np.sqrt(A[nbs])).mean() - q, (P[nbs] - 3.75).mean()]
        np.testing.assert_allclose(X[c, :12], ref, rtol=1e-4, atol=1e-5)
        np.testing.assert_allclose(F[c], np.concatenate([e1, e2]), atol=1e-6)
        assert not X[c, 12:].any()


def test_warm_start_moves_vertices_by_the_mean_cell_displacement():
    st = svmtest.seeded_state(500, 1302)
    ctx = svm.Context([dict(n_cells=500, a0=1.0, p0=3.7, kr=1.0, t1_threshold=1e-6, seed=1302)])
    ctx.set_topology(0, st["vpos"], st["vn"], st["vcn"], st["ring"], st["cnn"])
    p0 = ctx.positions(0)
    _, F = ctx.cell_descriptors(0)
    net = svm.MLP()
    th = np.zeros(mo.NP, np.float32)
    th[mo.unpack(np.arange(mo.NP))["b4"][:2]] = [0.01, -0.02]    # constant output
    net.set_parameters(th, np.zeros(12), np.ones(12), 1.0)
    ctx.warm_start(net, scale=1.0)
    p1 = ctx.positions(0)
    disp = np.zeros_like(p0)
    for v in range(len(p0)):
        cs = st["vcn"][v]
        disp[v] = np.mean([0.01 * F[c, :3] + (-0.02) * F[c, 3:] for c in cs], axis=0)
        # capped at 0.3 x the shortest edge
        emin = min(np.linalg.norm(p0[w] - p0[v]) for w in st["vn"][v])
        ln = np.linalg.norm(disp[v])
        if ln > 0.3 * emin:
            disp[v] *= 0.3 * emin / ln
    R = np.linalg.norm(p0[0])
    exp = p0 + disp - np.sum(disp * p0, 1)[:, None] * p0 / R ** 2
    exp *= R / np.linalg.norm(exp, axis=1)[:, None]
    np.testing.assert_allclose(p1, exp, atol=1e-7)


def test_mlp_forward_throughput():
    net = svm.MLP()
    th, mu, sg, sc = random_net(13)
    net.set_parameters(th, mu, sg, sc)
    n = 1 << 21
    ms = net.time_forward(n, reps=10)
    tflops = 18176.0 * n / (ms * 1e-3) / 1e12
    print(f"MLP forward: {n} cells in {ms:.3f} ms, {tflops:.1f} TFLOP/s (algorithmic)")
    assert ms > 0
