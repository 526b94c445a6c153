"""Write the spherical Delaunay fixtures (run once; outputs are committed).

    python tests/golden/make_fixtures.py

For each (N, seed): the seeded cell centres (repr of every double) and the
Delaunay triangles from scipy's Qhull, in canonical form (counter-clockwise
seen from outside, smallest index first, sorted).  For N <= 64 the triangle
set is also confirmed by brute force: a triple (a, b, c) is a Delaunay
triangle iff every other point lies strictly on the origin's side of the
plane through a, b, c (exact rational arithmetic).
"""
import itertools
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import svmtest  # noqa: E402

CASES = [(16, 1137), (64, 1137), (64, 1138), (64, 1139), (512, 1137), (512, 1138), (2000, 1137)]


def brute_force(points):
    P = np.asarray(points)
    n = len(P)
    out = []
    for a, b, c in itertools.combinations(range(n), 3):
        if svmtest.orient_exact(P[a], P[b], P[c]) < 0:
            b, c = c, b
        ok = True
        for d in range(n):
            if d in (a, b, c):
                continue
            if svmtest.orient_exact(P[a], P[b], P[c], P[d]) >= 0:
                ok = False
                break
        if ok:
            k = int(np.argmin([a, b, c]))
            out.append(([a, b, c][k:] + [a, b, c][:k]))
    out.sort()
    return np.array(out, np.int32)


def main():
    for n, seed in CASES:
        pts = svmtest.cell_points(n, seed)
        tris = svmtest.qhull_triangles(pts)
        if n <= 64:
            bf = brute_force(pts)
            assert np.array_equal(bf, tris), f"Qhull and brute force disagree for N={n} seed={seed}"
        assert not svmtest.locally_delaunay_violations(pts, tris, exact=n <= 512)
        path = os.path.join(HERE, f"delaunay_N{n}_s{seed}.txt")
        with open(path, "w") as f:
            f.write(f"# N={n} seed={seed}: {n} cell centres (x y z), then {len(tris)} triangles (a b c)\n")
            for p in pts:
                f.write(" ".join(repr(float(x)) for x in p) + "\n")
            for t in tris:
                f.write(f"{t[0]} {t[1]} {t[2]}\n")
        print("wrote", path)


if __name__ == "__main__":
    main()
