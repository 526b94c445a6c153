"""CPU checks of the warm-start network's oracle (oracle/mlp_oracle.py)."""
import numpy as np

import mlp_oracle as mo


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -7 + 2 ** -8, -2.5, 0.0, 1.0 + 2 ** -9], np.float32)
    r = mo.bf16(x)
    # bf16 keeps 7 fraction bits: 1 + 2^-8 is a tie -> even (1); 1 + 3*2^-9 -> 1 + 2^-7;
    # 1 + 2^-7 + 2^-8 is a tie -> even (1 + 2^-6); 1 + 2^-9 -> 1
    np.testing.assert_array_equal(r, [1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5, 0.0, 1.0])


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    th = mo.init_theta(5).astype(np.float64)
    th[mo.unpack(np.arange(mo.NP))["b1"]] = rng.normal(0, 0.1, 64)
    mu, sg = rng.normal(0, 1, 12), rng.uniform(0.5, 2, 12)
    X = rng.normal(0, 1, (37, 16))
    T = rng.normal(0, 1, (37, 2))
    y, cache = mo.forward(th, mu, sg, 0.7, X, rounding=False)
    G, L = mo.backward(cache, y, T)
    mask = mo.structural_mask()
    idx = np.flatnonzero(mask)
    for p in rng.choice(idx, 60, replace=False):
        h = 1e-6
        tp, tm = th.copy(), th.copy()
        tp[p] += h
        tm[p] -= h
        yp, cp = mo.forward(tp, mu, sg, 0.7, X, rounding=False)
        ym, cm = mo.forward(tm, mu, sg, 0.7, X, rounding=False)
        fd = (mo.backward(cp, yp, T)[1] - mo.backward(cm, ym, T)[1]) / (2 * h)
        assert abs(fd - G[p]) <= 1e-6 * max(1.0, abs(fd)), (p, fd, G[p])
    # parameters outside the structure get no gradient
    assert not G[~mask].any()
