"""CPU oracle of the warm-start network (TEST INFRASTRUCTURE ONLY).

The per-cell MLP has no counterpart in the reference; this is the plain
numpy statement of what the GPU kernels compute, used by the tests as the
checker:

  x' = bf16((x - mu) * (1 / sigma))           (float32 arithmetic, as on the GPU)
  z1 = W1 x' + b1,  a1 = bf16(silu(z1))
  z2 = W2 a1 + b2,  a2 = bf16(silu(z2))
  z3 = W3 a2 + b3,  a3 = bf16(silu(z3))
  y  = out_scale * (W4 a3 + b4)[:2]

with W = bf16(theta) and sums in float64 (the GPU accumulates in float32 on
MFMA).  `rounding=False` gives the same network with no bf16 rounding, in
float64 throughout: its gradient is checked against central finite
differences, which pins `backward`.  The gradient is that of
L = (1/n) sum |y - t|^2 with the roundings passed straight through, as the
GPU backward does.

Parameter layout (theta, 10448 floats): W1 [64][16] (inputs 12..15 unused),
b1 [64], W2 [64][64], b2, W3 [64][64], b3, W4 [16][64] (rows 2..15 unused),
b4 [16].
"""
import numpy as np

NP = 10448
SL = {
    "W1": (0, (64, 16)), "b1": (1024, (64,)), "W2": (1088, (64, 64)), "b2": (5184, (64,)),
    "W3": (5248, (64, 64)), "b3": (9344, (64,)), "W4": (9408, (16, 64)), "b4": (10432, (16,)),
}


def unpack(theta):
    th = np.asarray(theta)
    return {k: th[o:o + int(np.prod(s))].reshape(s) for k, (o, s) in SL.items()}


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even), as float64."""
    f = np.ascontiguousarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def structural_mask():
    m = np.ones(NP, bool)
    P = unpack(np.arange(NP))
    m[P["W1"][:, 12:].ravel()] = False
    m[P["W4"][2:, :].ravel()] = False
    m[P["b4"][2:]] = False
    return m


def forward(theta, mu, sigma, out_scale, X, rounding=True):
    """Returns y [n][2] and the cache for backward."""
    P = unpack(np.asarray(theta, np.float32 if rounding else np.float64).astype(np.float64))
    r = bf16 if rounding else (lambda v: np.asarray(v, np.float64))
    X = np.asarray(X)[:, :12]
    if rounding:
        inv = (np.float32(1.0) / np.asarray(sigma, np.float32)).astype(np.float32)
        x = (np.asarray(X, np.float32) - np.asarray(mu, np.float32)).astype(np.float32) * inv
        x = bf16(x.astype(np.float32))
    else:
        x = (np.asarray(X, np.float64) - np.asarray(mu, np.float64)) / np.asarray(sigma, np.float64)
    W = {k: r(P[k]) if k.startswith("W") else P[k] for k in P}
    z1 = x @ W["W1"][:, :12].T + W["b1"]
    a1 = r(silu(z1))
    z2 = a1 @ W["W2"].T + W["b2"]
    a2 = r(silu(z2))
    z3 = a2 @ W["W3"].T + W["b3"]
    a3 = r(silu(z3))
    y = out_scale * (a3 @ W["W4"][:2].T + W["b4"][:2])
    return y, dict(x=x, z1=z1, a1=a1, z2=z2, a2=a2, z3=z3, a3=a3, W=W, out_scale=out_scale)


def backward(cache, y, T):
    """Gradient (float64, length NP) of L = (1/n) sum |y - t|^2 and L."""
    n = y.shape[0]
    e = y - np.asarray(T, np.float64)
    loss = float((e ** 2).sum() / n)
    W = cache["W"]
    g4 = 2.0 * e / n * cache["out_scale"]                  # dL/d(W4 a3 + b4)
    d3 = (g4 @ W["W4"][:2]) * dsilu(cache["z3"])
    d2 = (d3 @ W["W3"]) * dsilu(cache["z2"])
    d1 = (d2 @ W["W2"]) * dsilu(cache["z1"])
    G = np.zeros(NP)
    P = unpack(np.arange(NP))
    G[P["W1"][:, :12].ravel()] = (d1.T @ cache["x"]).ravel()
    G[P["b1"]] = d1.sum(0)
    G[P["W2"].ravel()] = (d2.T @ cache["a1"]).ravel()
    G[P["b2"]] = d2.sum(0)
    G[P["W3"].ravel()] = (d3.T @ cache["a2"]).ravel()
    G[P["b3"]] = d3.sum(0)
    G[P["W4"][:2].ravel()] = (g4.T @ cache["a3"]).ravel()
    G[P["b4"][:2]] = g4.sum(0)
    return G, loss


def init_theta(seed):
    """He-style random initialisation, structural zeros kept."""
    rng = np.random.default_rng(seed)
    th = np.zeros(NP, np.float32)
    P = unpack(np.arange(NP))
    th[P["W1"][:, :12].ravel()] = rng.normal(0, np.sqrt(2.0 / 12), 64 * 12)
    th[P["W2"].ravel()] = rng.normal(0, np.sqrt(2.0 / 64), 4096)
    th[P["W3"].ravel()] = rng.normal(0, np.sqrt(2.0 / 64), 4096)
    th[P["W4"][:2].ravel()] = rng.normal(0, np.sqrt(1.0 / 64), 128)
    return th
