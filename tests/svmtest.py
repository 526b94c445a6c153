"""Test helpers: the CPU oracle binding and seeded topologies.

The oracle (oracle/liboracle.so) is the checker.  Spherical Delaunay
triangulations for test inputs come from scipy's Qhull (an independent
implementation), are put in canonical form and verified with exact
predicates before use.
"""
import ctypes as C
import os
from fractions import Fraction

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
ORACLE_LIBM = os.path.join(ROOT, "oracle", "liboracle_libm.so")
RING_CAP = 32

_dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
_ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))


def oracle_lib(libm=False):
    path = ORACLE_LIBM if libm else ORACLE
    if not os.path.exists(path):
        import subprocess
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, capture_output=True)
    L = C.CDLL(path)
    L.or_create.restype = C.c_void_p
    for fn in ("or_energy", "or_kinetic_energy", "or_tree_sum", "or_acos", "or_sin", "or_cos", "or_log",
               "or_acos_hi", "or_acos_lo", "or_own_acos", "or_cos_quarter_turn"):
        getattr(L, fn).restype = C.c_double
    for fn in ("or_acos", "or_sin", "or_cos", "or_log", "or_acos_hi", "or_acos_lo", "or_own_acos"):
        getattr(L, fn).argtypes = [C.c_double]
    L.or_energy.argtypes = [C.c_void_p, C.c_int]
    L.or_kinetic_energy.argtypes = [C.c_void_p, C.c_int]
    for fn in ("or_destroy", "or_compute_geometry", "or_compute_forces", "or_enforce_topology", "or_num_vertices",
               "or_max_deg", "or_num_t1", "or_num_fan", "or_num_nan"):
        getattr(L, fn).argtypes = [C.c_void_p]
    return L


def oracle_neighbor_list(points, range_=1.0, box_half=None, L=None):
    """The oracle's metric neighbour lists (or_neighbor_list): (counts, idx[N, M], vec[N, M, 3],
    dist[N, M]) with M the largest count, neighbours in index order."""
    L = L or oracle_lib()
    pts = np.ascontiguousarray(points, dtype=np.float64)
    n = len(pts)
    if box_half is None:
        box_half = float(np.abs(pts).max()) + 2.5
    counts = np.zeros(n, np.int32)
    f = L.or_neighbor_list
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                  C.c_void_p]
    M = f(n, pts.ctypes.data, range_, box_half, 0, counts.ctypes.data, None, None, None)
    idx = np.zeros((n, M), np.int32)
    vec = np.zeros((n, M, 3), np.float64)
    dist = np.zeros((n, M), np.float64)
    f(n, pts.ctypes.data, range_, box_half, M, counts.ctypes.data, idx.ctypes.data, vec.ctypes.data,
      dist.ctypes.data)
    return counts, idx, vec, dist


# ------------------------------------------------------------ topology ----
def orient_exact(a, b, c, d=None):
    """Sign of det[b-a, c-a, d-a] (or det[a, b, c] when d is None) in exact rationals."""
    if d is None:
        M = [[Fraction(float(x)) for x in v] for v in (a, b, c)]
    else:
        A = [Fraction(float(x)) for x in a]
        M = [[Fraction(float(v[k])) - A[k] for k in range(3)] for v in (b, c, d)]
    det = (M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1])
           - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0])
           + M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]))
    return (det > 0) - (det < 0)


def canonical_triangles(points, simplices):
    """Triangles counter-clockwise seen from outside, smallest index first, sorted."""
    P = np.asarray(points, np.float64)
    tris = []
    for t in simplices:
        a, b, c = (int(x) for x in t)
        det = np.dot(P[a], np.cross(P[b], P[c]))
        if abs(det) < 1e-9 * np.linalg.norm(P[a]) ** 3:
            if orient_exact(P[a], P[b], P[c]) < 0:
                b, c = c, b
        elif det < 0:
            b, c = c, b
        k = int(np.argmin([a, b, c]))
        tri = [a, b, c][k:] + [a, b, c][:k]
        tris.append(tri)
    tris.sort()
    return np.array(tris, dtype=np.int32)


def qhull_triangles(points):
    from scipy.spatial import ConvexHull
    hull = ConvexHull(np.asarray(points, np.float64))
    return canonical_triangles(points, hull.simplices)


def locally_delaunay_violations(points, tris, exact=False, limit=None):
    """Edges (a, b) whose opposite vertex d lies inside the circumcircle of the
    triangle (a, b, c) on the other side.

    For points on a sphere the circumcircle of a triangle is the intersection
    of the sphere with the triangle's plane, and its inside is the cap beyond
    that plane (away from the origin).  With the triangle counter-clockwise
    seen from outside, d is inside exactly when det[b-a, c-a, d-a] > 0.  A
    triangulation of the sphere without such an edge is the Delaunay
    triangulation (unique for points in general position).
    """
    P = np.asarray(points, np.float64)
    edge_opp = {}
    for t in tris:
        for k in range(3):
            a, b, c = int(t[k]), int(t[(k + 1) % 3]), int(t[(k + 2) % 3])
            edge_opp[(a, b)] = (c, t)
    bad = []
    for (a, b), (c, t) in edge_opp.items():
        if (b, a) not in edge_opp:
            bad.append(("open edge", a, b))
            continue
        d = edge_opp[(b, a)][0]
        if exact:
            s = orient_exact(P[t[0]], P[t[1]], P[t[2]], P[d])
        else:
            s = np.sign(np.dot(np.cross(P[t[1]] - P[t[0]], P[t[2]] - P[t[0]]), P[d] - P[t[0]]))
        if s > 0:
            bad.append(("in-circle", a, b, d))
        if limit and len(bad) >= limit:
            break
    return bad


def topology_from_triangles(points, tris, L=None):
    """Vertex-model topology (reference layout) from canonical triangles via the oracle."""
    L = L or oracle_lib()
    P = np.ascontiguousarray(points, np.float64)
    T = np.ascontiguousarray(tris, np.int32)
    n, nt = P.shape[0], T.shape[0]
    vpos = np.zeros((nt, 3), np.float64)
    vn = np.zeros((nt, 3), np.int32)
    vcn = np.zeros((nt, 3), np.int32)
    ring = np.full((n, RING_CAP), -1, np.int32)
    cnn = np.zeros(n, np.int32)
    md = L.or_topology_from_triangles(C.c_int(n), _dp(P), C.c_int(nt), _ip(T), _dp(vpos), _ip(vn), _ip(vcn),
                                      _ip(ring), _ip(cnn), C.c_int(RING_CAP))
    assert md > 0, f"ring capacity too small ({-md})"
    return dict(vpos=vpos, vn=vn, vcn=vcn, ring=ring, cnn=cnn, maxdeg=md)


def cell_points(n, seed, L=None):
    L = L or oracle_lib()
    out = np.zeros((n, 3), np.float64)
    L.or_generate_cell_points(C.c_int(n), C.c_uint64(seed), _dp(out))
    return out


def directors(vpos, seed, L=None):
    L = L or oracle_lib()
    P = np.ascontiguousarray(vpos, np.float64)
    out = np.zeros_like(P)
    L.or_generate_directors(C.c_int(P.shape[0]), _dp(P), C.c_uint64(seed), _dp(out))
    return out


def seeded_state(n, seed=1137):
    """Seeded cell points, their Delaunay topology and vertex directors."""
    L = oracle_lib()
    pts = cell_points(n, seed, L)
    tris = qhull_triangles(pts)
    topo = topology_from_triangles(pts, tris, L)
    topo["points"] = pts
    topo["tris"] = tris
    topo["directors"] = directors(topo["vpos"], seed, L)
    topo["R"] = np.sqrt(n / (4.0 * np.pi))
    return topo


# ------------------------------------------------------------ oracle model --
class OracleModel:
    """The CPU oracle's vertex model (reference CPU path restated)."""

    def __init__(self, topo, a0=1.0, p0=3.8, kr=1.0, t1=0.1, reduce_mode=1, libm=False):
        self.L = oracle_lib(libm)
        n = topo["cnn"].shape[0]
        nv = topo["vpos"].shape[0]
        self.n, self.nv = n, nv
        R = np.sqrt(n / (4.0 * np.pi))
        self.h = C.c_void_p(self.L.or_create(C.c_int(n), C.c_int(nv), C.c_double(R), _dp(topo["vpos"]),
                                             _ip(topo["vn"]), _ip(topo["vcn"]),
                                             _ip(np.ascontiguousarray(topo["ring"])),
                                             C.c_int(topo["ring"].shape[1]), _ip(topo["cnn"])))
        prefs = np.tile(np.array([a0, p0], np.float64), n)
        self.L.or_set_preferences(self.h, _dp(prefs))
        self.L.or_set_kr(self.h, C.c_double(kr))
        self.L.or_set_t1_threshold(self.h, C.c_double(t1))
        self.L.or_set_reduce_mode(self.h, C.c_int(reduce_mode))
        if "directors" in topo:
            self.L.or_set_directors(self.h, _dp(np.ascontiguousarray(topo["directors"])))

    def __del__(self):
        try:
            self.L.or_destroy(self.h)
        except Exception:
            pass

    def _get(self, fn, rows, cols=3):
        out = np.zeros((rows, cols), np.float64)
        getattr(self.L, fn)(self.h, _dp(out))
        return out

    def positions(self):
        return self._get("or_get_positions", self.nv)

    def forces(self):
        return self._get("or_get_forces", self.nv)

    def velocities(self):
        return self._get("or_get_velocities", self.nv)

    def directors(self):
        return self._get("or_get_directors", self.nv)

    def cell_positions(self):
        return self._get("or_get_cell_positions", self.n)

    def area_perimeter(self):
        return self._get("or_get_area_perimeter", self.n, 2)

    def topology(self, stride=RING_CAP):
        vn = np.zeros((self.nv, 3), np.int32)
        vcn = np.zeros((self.nv, 3), np.int32)
        ring = np.zeros((self.n, stride), np.int32)
        cnn = np.zeros(self.n, np.int32)
        r = self.L.or_get_topology(self.h, _ip(vn), _ip(vcn), _ip(ring), C.c_int(stride), _ip(cnn))
        assert r == 0
        return vn, vcn, ring, cnn

    def set_positions(self, p):
        self.L.or_set_positions(self.h, _dp(np.ascontiguousarray(p, np.float64)))

    def set_velocities(self, v):
        self.L.or_set_velocities(self.h, _dp(np.ascontiguousarray(v, np.float64)))

    def set_preferences(self, prefs):
        self.L.or_set_preferences(self.h, _dp(np.ascontiguousarray(prefs, np.float64)))

    def compute_geometry(self):
        self.L.or_compute_geometry(self.h)

    def compute_forces(self):
        self.L.or_compute_forces(self.h)

    def energy(self, mode=1):
        return self.L.or_energy(self.h, C.c_int(mode))

    def kinetic_energy(self, mode=1):
        return self.L.or_kinetic_energy(self.h, C.c_int(mode))

    def move(self, disp, scale=1.0):
        self.L.or_move(self.h, _dp(np.ascontiguousarray(disp, np.float64)), C.c_double(scale))

    def enforce_topology(self):
        self.L.or_enforce_topology(self.h)

    def brownian(self, n, dt=0.01, mu=1.0, T=0.0, seed=1137, step0=0):
        self.L.or_brownian_steps(self.h, C.c_int(n), C.c_double(dt), C.c_double(mu), C.c_double(T),
                                 C.c_uint64(seed), C.c_uint64(step0))

    def active(self, n, dt=0.01, mu=1.0, v0=0.1, tau=1.0, seed=1137, step0=0):
        self.L.or_active_steps(self.h, C.c_int(n), C.c_double(dt), C.c_double(mu), C.c_double(v0),
                               C.c_double(tau), C.c_uint64(seed), C.c_uint64(step0))

    def vicsek(self, n, dt=0.01, mu=1.0, v0=0.01, tau=1.0, eta=1.0, seed=1137, step0=0):
        self.L.or_vicsek_steps(self.h, C.c_int(n), C.c_double(dt), C.c_double(mu), C.c_double(v0),
                               C.c_double(tau), C.c_double(eta), C.c_uint64(seed), C.c_uint64(step0))

    def fire(self, dt=0.01, alpha_start=0.99, dt_max=0.1, dt_inc=1.05, dt_dec=0.95, alpha_dec=0.9, n_min=5,
             force_cutoff=1e-7, alpha_min=0.0, max_iterations=1000, state=None, reduce_mode=1, stop_rms=0):
        params = np.array([dt, alpha_start, dt_max, dt_inc, dt_dec, alpha_dec, n_min, force_cutoff, alpha_min,
                           max_iterations], np.float64)
        st = np.array(state if state is not None else [0, dt, alpha_start, 0, 100.0, 0.0], np.float64)
        self.L.or_fire_minimize(self.h, _dp(params), _dp(st), C.c_int(reduce_mode), C.c_int(stop_rms))
        return dict(iterations=int(st[0]), dt=st[1], alpha=st[2], n_since_negative_power=int(st[3]),
                    force_max=st[4], power=st[5], state=st)

    def adam(self, beta1=0.9, beta2=0.99, epsilon=1e-8, dt=1e-4, force_cutoff=1e-12, max_iterations=1000,
             state=None, moments=None, reduce_mode=1):
        """Adam minimisation (oracle or_adam_minimize); returns the state and the moments."""
        params = np.array([beta1, beta2, epsilon, dt, force_cutoff, max_iterations], np.float64)
        st = np.array(state if state is not None else [0, beta1, beta2, 110.0], np.float64)
        n = self.L.or_num_vertices(self.h)
        m1, m2 = (np.zeros((n, 3)), np.zeros((n, 3))) if moments is None else (moments[0].copy(), moments[1].copy())
        self.L.or_adam_minimize(self.h, _dp(params), _dp(st), _dp(m1), _dp(m2), C.c_int(reduce_mode))
        return dict(iterations=int(st[0]), beta1t=st[1], beta2t=st[2], force_max=st[3], state=st, moments=(m1, m2))

    def msd_overlap(self, ref, fraction=1.0, cutoff=0.5, reduce_mode=1):
        """dynamicalFeatures on the oracle's current cell positions against ref [N][3]."""
        out = np.zeros(2)
        r = np.ascontiguousarray(ref, np.float64)
        self.L.or_msd_overlap(self.h, _dp(r), C.c_double(fraction), C.c_double(cutoff), C.c_int(reduce_mode),
                              _dp(out))
        return out[0], out[1]

    @property
    def n_t1(self):
        return self.L.or_num_t1(self.h)

    @property
    def n_fan(self):
        return self.L.or_num_fan(self.h)


# ------------------------------------------- well-conditioned FP64 forces --
def vector_forces(pos, cell_pos, area_perimeter, vn, vcn, ring, cnn, R, a0=1.0, p0=3.8, kr=1.0):
    """The model's forces from the vector-form gradients (the same formulas
    the FP32 force path uses, here in float64, vectorised over vertices).

    They equal the reference's trig-basis expressions mathematically, but
    stay accurate on large spheres, where the reference's Gram determinant of
    a small fan triangle (~theta^4, from O(1) terms) loses digits even in
    FP64: at N = 262144 the two differ by 2e-3 rms and high-precision finite
    differences of the energy side with this form."""
    X = np.asarray(pos, np.float64)
    C = np.asarray(cell_pos, np.float64)
    A, P = area_perimeter[:, 0], area_perimeter[:, 1]
    nv = len(X)
    a = X / R

    def garc(e):
        t = e - np.sum(a * e, -1)[:, None] * a
        return -t / np.linalg.norm(np.cross(a, e), axis=-1)[:, None]

    def gtri(eb, ec):
        bxc = np.cross(a, ec - eb) + np.cross(eb, ec)
        N = np.sum(a * np.cross(eb, ec), -1)
        dbc = eb - ec
        D = 4 - 0.5 * (np.sum(eb * eb, -1) + np.sum(ec * ec, -1) + np.sum(dbc * dbc, -1))
        tb = bxc - N[:, None] * a
        s = eb + ec
        ts = s - np.sum(a * s, -1)[:, None] * a
        return (2 * R / (N * N + D * D))[:, None] * (D[:, None] * tb - N[:, None] * ts)

    F = np.zeros_like(X)
    idx = np.arange(nv)
    for sl in range(3):
        c = vcn[:, sl]
        n = cnn[c]
        rg = ring[c]
        k = np.argmax(rg == idx[:, None], axis=1)
        last = rg[idx, (k - 1) % n]
        nxt = rg[idx, (k + 1) % n]
        el, en, ec = (X[last] - X) / R, (X[nxt] - X) / R, (C[c] - X) / R
        cp = 2 * (P[c] - p0)
        ca = 2 * kr * (A[c] - a0)
        F += -cp[:, None] * (garc(el) + garc(en)) - ca[:, None] * (gtri(el, ec) + gtri(ec, en))
    return F
