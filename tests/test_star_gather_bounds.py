"""Host check of the star kernel's candidate gather (k_dl_star, the GPU
Delaunay's neighbour search): every index it forms stays inside its array,
and the candidates it gathers are exactly the points within the search
radius.

The kernel's integer arithmetic is restated here step by step -- the cell
list of dl_cell_list (grid over [-Rs, Rs]^3, G bins per axis, counting sort),
the box of bins [p - r, p + r] per axis with the clamped bin index of
dl_axis, the 64-bin chunks with a scan of the bin sizes over the lanes, the
binary search that maps entry t of a chunk to its bin, and the slot of each
hit in the wave's candidate array -- with bounds assertions on every access,
for the 16-point fixture (the input on which an earlier, filter-only variant
of the star pass faulted) and for random spheres, at every search radius the
host uses (r0, 2 r0, 3 r0 and the whole sphere).
"""
import math

import numpy as np
import pytest

import svmtest

MAXC = 512     # SVM_DL_MAXC: candidates one star holds


def grid(P, Rs, h):
    lo = -Rs * (1.0 + 1e-9)
    G = max(1, min(256, math.ceil(2.0 * Rs * (1.0 + 1e-9) / h)))
    return lo, 2.0 * Rs * (1.0 + 1e-9) / G, G


def axis(x, lo, h, G):
    k = math.floor((x - lo) / h)
    return 0 if k < 0 else (G - 1 if k >= G else k)


def cell_list(P, lo, h, G):
    n = len(P)
    bin_of = [(axis(p[0], lo, h, G) * G + axis(p[1], lo, h, G)) * G + axis(p[2], lo, h, G) for p in P]
    count = np.zeros(G ** 3 + 1, np.int64)
    for b in bin_of:
        assert 0 <= b < G ** 3
        count[b] += 1
    start = np.concatenate([[0], np.cumsum(count)])[:G ** 3 + 1]   # exclusive scan, start[G^3] = n
    assert start[-1] == n
    lst = np.empty(n, np.int64)
    fill = np.zeros(G ** 3, np.int64)
    for i, b in enumerate(bin_of):            # each bin's points in index order
        lst[start[b] + fill[b]] = i
        fill[b] += 1
    return start, lst


def gather(P, i, r, lo, h, G, start, lst):
    """The candidates of star i as k_dl_star gathers them, with every index checked."""
    n, nbins = len(P), G ** 3
    pi = P[i]
    rb = r * (1.0 + 1e-9)
    x0, y0, z0 = (axis(pi[k] - rb, lo, h, G) for k in range(3))
    nx, ny, nz = (axis(pi[k] + rb, lo, h, G) - c + 1 for k, c in zip(range(3), (x0, y0, z0)))
    assert nx >= 1 and ny >= 1 and nz >= 1
    nbin = nx * ny * nz
    cand, nc = [], 0
    for c0 in range(0, nbin, 64):
        s0 = np.zeros(64, np.int64)
        cnt = np.zeros(64, np.int64)
        for lane in range(64):
            b = c0 + lane
            if b < nbin:
                x, y, z = x0 + b % nx, y0 + (b // nx) % ny, z0 + b // (nx * ny)
                assert 0 <= x < G and 0 <= y < G and 0 <= z < G
                binid = (x * G + y) * G + z
                assert 0 <= binid and binid + 1 <= nbins        # start[] has nbins + 1 entries
                s0[lane] = start[binid]
                cnt[lane] = start[binid + 1] - s0[lane]
        inc = np.cumsum(cnt)
        total = int(inc[-1])
        bp = inc - cnt
        for t in range(total):
            lo_, hi_ = 0, 63
            while lo_ < hi_:
                mid = (lo_ + hi_ + 1) >> 1
                if bp[mid] <= t:
                    lo_ = mid
                else:
                    hi_ = mid - 1
            assert c0 + lo_ < nbin and cnt[lo_] > 0           # a real bin that holds entries
            idx = s0[lo_] + (t - bp[lo_])
            assert 0 <= idx < n and start[0] <= idx < start[-1]
            q = int(lst[idx])
            assert 0 <= q < n
            if q != i and float(np.sum((P[q] - pi) ** 2)) < r * r:
                if nc < MAXC:
                    cand.append(q)
                nc += 1
    return cand, nc


def check_points(P):
    n = len(P)
    Rs = float(np.max(np.linalg.norm(P, axis=1)))
    r0 = 3.5 * math.sqrt(4.0 * math.pi * Rs * Rs / n)
    lo, h, G = grid(P, Rs, r0)
    start, lst = cell_list(P, lo, h, G)
    d2 = np.sum((P[:, None, :] - P[None, :, :]) ** 2, -1)
    for r in (r0, 2 * r0, 3 * r0, 2.0 * Rs * (1.0 + 1e-9)):
        for i in range(n):
            cand, nc = gather(P, i, r, lo, h, G, start, lst)
            brute = sorted(int(q) for q in np.flatnonzero(d2[i] < r * r) if q != i)
            assert nc == len(brute)
            if nc <= MAXC:
                assert sorted(cand) == brute          # nothing within r is missed, nothing else is taken
            else:
                assert len(cand) == MAXC              # the capacity guard: the star is reported crowded


def fixture_points(path):
    rows = [l.split() for l in open(path) if not l.startswith("#")]
    return np.array([[float(x) for x in r] for r in rows if len(r) == 3 and "." in r[0]])


def test_gather_on_the_16_point_fixture():
    P = fixture_points(svmtest.ROOT + "/tests/golden/delaunay_N16_s1137.txt")
    assert P.shape == (16, 3)
    check_points(P)


@pytest.mark.parametrize("n,seed", [(16, 1138), (64, 1137), (300, 7), (1000, 1137)])
def test_gather_on_random_spheres(n, seed):
    check_points(svmtest.cell_points(n, seed))


def test_gather_on_a_clustered_sphere():
    """A dense cap beside a sparse rest: some stars exceed the capacity at the
    larger radii, and the guard keeps their writes inside the array."""
    rng = np.random.default_rng(4)
    n = 700
    R = math.sqrt(n / (4 * math.pi))
    v = rng.normal(size=(n, 3))
    v[:500] = np.array([0.0, 0.0, 1.0]) + 0.08 * v[:500]       # 500 points crowded near the pole
    P = R * v / np.linalg.norm(v, axis=1)[:, None]
    check_points(P)
