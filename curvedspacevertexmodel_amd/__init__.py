"""ctypes binding of libsvm.so, the MI355X spherical vertex-model library.

The library is the product; this module only moves numpy arrays across its C
ABI (include/svm.h).  Importing it never falls back to anything: if
libsvm.so is missing or cannot be loaded, `lib()` raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsvm.so")

SVM_OK = 0
STATUS = {
    0: "SVM_OK", 1: "SVM_ERR_ARG", 2: "SVM_ERR_HIP", 3: "SVM_ERR_NO_DEVICE",
    4: "SVM_ERR_RING_OVERFLOW", 5: "SVM_ERR_TOPOLOGY", 6: "SVM_ERR_DELAUNAY",
    7: "SVM_ERR_STATE", 8: "SVM_ERR_CAPACITY",
}

# Every symbol include/svm.h declares.
EXPORTS = [
    "svm_create", "svm_destroy", "svm_set_stream", "svm_synchronize", "svm_get_num_vertices",
    "svm_last_error", "svm_build_from_cell_points", "svm_generate_cell_points",
    "svm_generate_directors", "svm_set_topology_from_triangles", "svm_set_topology", "svm_get_topology", "svm_get_delaunay_triangles",
    "svm_set_positions", "svm_get_positions", "svm_set_directors", "svm_get_directors",
    "svm_set_velocities", "svm_get_velocities", "svm_get_forces", "svm_get_cell_positions",
    "svm_get_area_perimeter", "svm_set_preferred_parameters", "svm_set_preferences",
    "svm_set_scalar_model_parameter", "svm_set_t1_threshold", "svm_compute_geometry",
    "svm_compute_forces", "svm_move_particles", "svm_enforce_topology", "svm_get_energies",
    "svm_get_counters", "svm_set_brownian", "svm_brownian_steps", "svm_set_active",
    "svm_active_steps", "svm_run_graph", "svm_fire_set_parameters", "svm_fire_minimize",
    "svm_fire_get_state", "svm_set_t1_mode", "svm_set_t1_component_distance", "svm_get_t1_component_distance", "svm_get_t1_diagnostics", "svm_set_profiling", "svm_get_timing",
    "svm_copy_probe", "svm_spatial_order", "svm_time_vertex_pass",
    "svm_set_force_precision", "svm_get_force_error", "svm_delaunay_triangles",
    "svm_cell_list", "svm_mlp_create", "svm_mlp_destroy", "svm_mlp_set_parameters", "svm_mlp_get_parameters",
    "svm_mlp_forward", "svm_mlp_gradient", "svm_mlp_train", "svm_mlp_time_forward", "svm_cell_descriptors",
    "svm_warm_start", "svm_fire_set_max_iterations", "svm_set_tuning", "svm_get_tuning", "svm_autotune",
    "svm_set_lanes", "svm_get_lanes", "svm_set_lane_spheres",
    "svm_adam_set_parameters", "svm_adam_minimize", "svm_adam_get_state", "svm_adam_set_max_iterations",
    "svm_dyn_set_reference", "svm_dyn_measure", "svm_nc_open", "svm_nc_write", "svm_nc_num_records", "svm_nc_close",
    "svm_set_vicsek", "svm_vicsek_steps",
    "svm_build_voronoi_from_cell_points", "svm_voronoi_forces", "svm_get_voronoi_forces", "svm_voronoi_brownian_steps", "svm_get_generators",
    "svm_set_positions_device", "svm_get_positions_device", "svm_set_directors_device", "svm_get_directors_device",
    "svm_set_velocities_device", "svm_get_velocities_device", "svm_get_forces_device",
    "svm_delaunay_triangles_noalloc", "svm_set_vertex_form", "svm_get_vertex_form",
    "svm_tuning_save", "svm_tuning_load", "svm_set_step_form", "svm_get_step_form", "svm_phase_boundary_probe",
    "svm_set_geometry_form", "svm_get_geometry_form", "svm_neighbor_list", "svm_fast_ops_check",
]

NOALLOC_MAX_POINTS = 65536


class SystemDesc(C.Structure):
    _fields_ = [("n_cells", C.c_int32), ("ring_capacity", C.c_int32), ("a0", C.c_double),
                ("p0", C.c_double), ("kr", C.c_double), ("t1_threshold", C.c_double),
                ("seed", C.c_uint64)]


class FireParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("alpha_start", C.c_double), ("dt_max", C.c_double),
                ("dt_inc", C.c_double), ("dt_dec", C.c_double), ("alpha_dec", C.c_double),
                ("n_min", C.c_int32), ("max_iterations", C.c_int32), ("force_cutoff", C.c_double),
                ("alpha_min", C.c_double), ("stop_rms", C.c_int32), ("reserved", C.c_int32)]


class FireState(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("n_since_negative_power", C.c_int32),
                ("converged", C.c_int32), ("reserved", C.c_int32), ("dt", C.c_double),
                ("alpha", C.c_double), ("force_max", C.c_double), ("power", C.c_double)]


class AdamParams(C.Structure):
    _fields_ = [("beta1", C.c_double), ("beta2", C.c_double), ("epsilon", C.c_double), ("dt", C.c_double),
                ("force_cutoff", C.c_double), ("max_iterations", C.c_int32), ("reserved", C.c_int32)]


class AdamState(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32), ("force_max", C.c_double),
                ("beta1t", C.c_double), ("beta2t", C.c_double)]


_LIB = None


def lib():
    """Load libsvm.so (raises if it is not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run python -m curvedspacevertexmodel_amd.build")
        L = C.CDLL(LIB_PATH)
        L.svm_last_error.restype = C.c_char_p
        L.svm_last_error.argtypes = [C.c_void_p]
        _LIB = L
    return _LIB


MAX_LANES = 4   # SVM_MAX_LANES of include/svm.h
COUNTER_COLS = 6      # SVM_COUNTER_COLS: per-sphere counters
FORCE_ERR_COLS = 7    # SVM_FORCE_ERR_COLS: per-sphere force-error statistics
T1_DIAG_COLS = 12     # SVM_T1_DIAG_COLS: per-sphere T1 diagnostics
TUNE_COLS = 8         # SVM_TUNE_COLS: autotuner trial-table width


class SVMError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def generate_cell_points(n_cells, seed):
    out = np.zeros((n_cells, 3), dtype=np.float64)
    r = lib().svm_generate_cell_points(C.c_int32(n_cells), C.c_uint64(seed), _dp(out))
    if r:
        raise SVMError(r, "svm_generate_cell_points")
    return out


def generate_directors(positions, seed):
    pos = np.ascontiguousarray(positions, dtype=np.float64)
    out = np.zeros_like(pos)
    r = lib().svm_generate_directors(C.c_int32(pos.shape[0]), _dp(pos), C.c_uint64(seed), _dp(out))
    if r:
        raise SVMError(r, "svm_generate_directors")
    return out


def spatial_order(points):
    """Permutation that relabels points in Z-order (svm_spatial_order)."""
    pts = np.ascontiguousarray(points, dtype=np.float64)
    out = np.empty(len(pts), np.int32)
    rc = lib().svm_spatial_order(C.c_int32(len(pts)), _dp(pts), _ip(out))
    if rc != 0:
        raise SVMError(rc, "svm_spatial_order")
    return out


def delaunay_triangles(points):
    """Spherical Delaunay triangles of points (N x 3) on the GPU, int32 (2N-4) x 3,
    CCW seen from outside, smallest index first, sorted lexicographically."""
    pts = np.ascontiguousarray(points, dtype=np.float64)
    out = np.empty((2 * len(pts) - 4, 3), np.int32)
    rc = lib().svm_delaunay_triangles(C.c_int32(len(pts)), _dp(pts), _ip(out))
    if rc != 0:
        raise SVMError(rc, "svm_delaunay_triangles")
    return out


def delaunay_triangles_noalloc(points, stream=None):
    """delaunay_triangles through svm_delaunay_triangles_noalloc: no device
    allocation (the library's static work array), on `stream` (a hipStream_t
    handle as an int, e.g. torch.cuda.current_stream().cuda_stream; None = the
    null stream).  At most NOALLOC_MAX_POINTS points."""
    pts = np.ascontiguousarray(points, dtype=np.float64)
    out = np.empty((max(2 * len(pts) - 4, 0), 3), np.int32)
    rc = lib().svm_delaunay_triangles_noalloc(C.c_int32(len(pts)), _dp(pts), _ip(out), C.c_void_p(stream))
    if rc != 0:
        raise SVMError(rc, "svm_delaunay_triangles_noalloc")
    return out


def neighbor_list(points, range_=1.0, box_half=None, max_neighbors=4):
    """Metric neighbour lists (svm_neighbor_list; neighborList::computeNeighborLists): returns
    (counts int32[N], indices int32[N, M], vectors float64[N, M, 3], distances float64[N, M]) with M the
    largest count, each point's neighbours in index order.  box_half defaults to max|coordinate| + 2.5
    (sphericalModel's sphereRadius R + 2.5).  Like the reference, a list that overflows max_neighbors is
    recomputed with the needed width."""
    pts = np.ascontiguousarray(points, dtype=np.float64)
    n = len(pts)
    if box_half is None:
        box_half = float(np.abs(pts).max()) + 2.5
    counts = np.empty(n, np.int32)
    needed = C.c_int32(0)
    M = max(int(max_neighbors), 1)
    while True:
        idx = np.empty((n, M), np.int32)
        vec = np.empty((n, M, 3), np.float64)
        dist = np.empty((n, M), np.float64)
        rc = lib().svm_neighbor_list(C.c_int32(n), _dp(pts), C.c_double(range_), C.c_double(box_half),
                                     C.c_int32(M), _ip(counts), _ip(idx), _dp(vec), _dp(dist), C.byref(needed))
        if rc == 8 and needed.value > M:     # SVM_ERR_CAPACITY: recompute with a larger width
            M = needed.value
            continue
        if rc != 0:
            raise SVMError(rc, "svm_neighbor_list")
        w = max(int(needed.value), 0)
        return counts, idx[:, :w].copy(), vec[:, :w].copy(), dist[:, :w].copy()


def cell_list(points, cell_size):
    """GPU cell list: (G, bin_of, bin_start, sorted_points); see svm_cell_list."""
    pts = np.ascontiguousarray(points, dtype=np.float64)
    G = C.c_int32(0)
    rc = lib().svm_cell_list(C.c_int32(len(pts)), _dp(pts), C.c_double(cell_size), C.byref(G), None, None, None)
    if rc != 0:
        raise SVMError(rc, "svm_cell_list")
    bin_of = np.empty(len(pts), np.int32)
    start = np.empty(G.value ** 3 + 1, np.int32)
    order = np.empty(len(pts), np.int32)
    rc = lib().svm_cell_list(C.c_int32(len(pts)), _dp(pts), C.c_double(cell_size), C.byref(G), _ip(bin_of),
                             _ip(start), _ip(order))
    if rc != 0:
        raise SVMError(rc, "svm_cell_list")
    return G.value, bin_of, start, order


MLP_NPARAM = 10448
MLP_NDESC = 12


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class MLP:
    """The per-cell warm-start network (svm_mlp_*): 12 -> 64 -> 64 -> 64 -> 2,
    bf16 weights on MFMA, trained on the GPU."""

    def __init__(self):
        self.h = C.c_void_p()
        rc = lib().svm_mlp_create(C.byref(self.h))
        if rc != 0:
            raise SVMError(rc, "svm_mlp_create")

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            lib().svm_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def _ck(self, rc, what):
        if rc != 0:
            raise SVMError(rc, what)

    def set_parameters(self, theta, mu, sigma, out_scale=1.0):
        th = np.ascontiguousarray(theta, np.float32)
        mu = np.ascontiguousarray(mu, np.float32)
        sg = np.ascontiguousarray(sigma, np.float32)
        assert th.size == MLP_NPARAM and mu.size == MLP_NDESC and sg.size == MLP_NDESC
        self._ck(lib().svm_mlp_set_parameters(self.h, _fp(th), _fp(mu), _fp(sg), C.c_float(out_scale)),
                 "svm_mlp_set_parameters")

    def get_parameters(self):
        th = np.zeros(MLP_NPARAM, np.float32)
        mu = np.zeros(MLP_NDESC, np.float32)
        sg = np.zeros(MLP_NDESC, np.float32)
        sc = C.c_float(0)
        self._ck(lib().svm_mlp_get_parameters(self.h, _fp(th), _fp(mu), _fp(sg), C.byref(sc)),
                 "svm_mlp_get_parameters")
        return th, mu, sg, sc.value

    def forward(self, X):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.zeros((X.shape[0], 2), np.float32)
        self._ck(lib().svm_mlp_forward(self.h, _fp(X), C.c_int32(X.shape[0]), _fp(Y)), "svm_mlp_forward")
        return Y

    def gradient(self, X, T):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        g = np.zeros(MLP_NPARAM, np.float32)
        loss = C.c_double(0)
        self._ck(lib().svm_mlp_gradient(self.h, _fp(X), _fp(T), C.c_int32(X.shape[0]), _fp(g), C.byref(loss)),
                 "svm_mlp_gradient")
        return g, loss.value

    def train(self, X, T, epochs, batch=4096, lr=1e-3, seed=1):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        losses = np.zeros(epochs, np.float64)
        self._ck(lib().svm_mlp_train(self.h, _fp(X), _fp(T), C.c_int32(X.shape[0]), C.c_int32(epochs),
                                     C.c_int32(batch), C.c_float(lr), C.c_uint64(seed), _dp(losses)),
                 "svm_mlp_train")
        return losses

    def time_forward(self, n, reps=20):
        ms = C.c_double(0)
        self._ck(lib().svm_mlp_time_forward(self.h, C.c_int32(n), C.c_int32(reps), C.byref(ms)),
                 "svm_mlp_time_forward")
        return ms.value


def copy_probe(n_bytes, reps=3):
    """Mean milliseconds of one plain device copy of n_bytes (k_copy_probe)."""
    ms = C.c_double(0.0)
    rc = lib().svm_copy_probe(C.c_int64(n_bytes), C.c_int32(reps), C.byref(ms))
    if rc != 0:
        raise SVMError(rc, "svm_copy_probe")
    return ms.value


def phase_boundary_probe(workgroups, phases=200, bytes_per_phase=64 << 20):
    """Microseconds per phase with a kernel boundary (one launch per phase, in a HIP graph) and with a
    grid barrier inside one persistent cooperative launch, and the values either read wrong
    (svm_phase_boundary_probe)."""
    a, b, e = C.c_double(0.0), C.c_double(0.0), C.c_int64(0)
    rc = lib().svm_phase_boundary_probe(C.c_int32(workgroups), C.c_int32(phases), C.c_int64(bytes_per_phase),
                                        C.byref(a), C.byref(b), C.byref(e))
    if rc != 0:
        raise SVMError(rc, "svm_phase_boundary_probe")
    return a.value, b.value, e.value


def fast_ops_check(n=1 << 24, seed=1):
    """Counts of svm_fast_ops_check: (quotient samples in the precondition, accepted, differing from a / b,
    root samples, accepted, differing from sqrt, accepted quotients and roots not the nearest double)."""
    out = np.zeros(8, np.int64)
    rc = lib().svm_fast_ops_check(C.c_int64(n), C.c_uint64(seed), out.ctypes.data_as(C.POINTER(C.c_int64)))
    if rc != 0:
        raise SVMError(rc, "svm_fast_ops_check")
    return tuple(int(x) for x in out)


class Context:
    """A batch of spheres resident on the GPU (one svm_context)."""

    def __init__(self, systems):
        """systems: list of dicts with n_cells, a0, p0, kr, t1_threshold, seed, ring_capacity."""
        self.n = len(systems)
        arr = (SystemDesc * self.n)()
        for i, s in enumerate(systems):
            arr[i] = SystemDesc(int(s["n_cells"]), int(s.get("ring_capacity", 0)), float(s.get("a0", 1.0)),
                                float(s.get("p0", 3.8)), float(s.get("kr", 1.0)),
                                float(s.get("t1_threshold", 0.1)), int(s.get("seed", 1137 + i)))
        self.n_cells = [int(s["n_cells"]) for s in systems]
        self.n_vertices = [2 * n - 4 for n in self.n_cells]
        h = C.c_void_p()
        r = lib().svm_create(arr, C.c_int32(self.n), C.byref(h))
        if r:
            raise SVMError(r, "svm_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().svm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, r, what):
        if r:
            raise SVMError(r, f"{what}: {lib().svm_last_error(self.h).decode()}")

    # ---- topology ----
    def set_topology(self, sys, vpos, vn, vcn, ring, counts):
        vpos = np.ascontiguousarray(vpos, np.float64)
        vn = np.ascontiguousarray(vn, np.int32)
        vcn = np.ascontiguousarray(vcn, np.int32)
        ring = np.ascontiguousarray(ring, np.int32)
        counts = np.ascontiguousarray(counts, np.int32)
        self._ck(lib().svm_set_topology(self.h, C.c_int32(sys), C.c_int32(vpos.shape[0]), _dp(vpos), _ip(vn),
                                        _ip(vcn), _ip(ring), C.c_int32(ring.shape[1]), _ip(counts)),
                 "svm_set_topology")

    def set_topology_from_triangles(self, sys, points, triangles):
        pts = np.ascontiguousarray(points, np.float64)
        tris = np.ascontiguousarray(triangles, np.int32)
        self._ck(lib().svm_set_topology_from_triangles(self.h, C.c_int32(sys), _dp(pts), C.c_int32(tris.shape[0]),
                                                       _ip(tris)), "svm_set_topology_from_triangles")

    def get_topology(self, sys, stride=32):
        nv, nc = self.n_vertices[sys], self.n_cells[sys]
        vn = np.zeros((nv, 3), np.int32)
        vcn = np.zeros((nv, 3), np.int32)
        ring = np.zeros((nc, stride), np.int32)
        counts = np.zeros(nc, np.int32)
        self._ck(lib().svm_get_topology(self.h, C.c_int32(sys), _ip(vn), _ip(vcn), _ip(ring), C.c_int32(stride),
                                        _ip(counts)), "svm_get_topology")
        return vn, vcn, ring, counts

    def build_from_cell_points(self, sys, points):
        pts = np.ascontiguousarray(points, np.float64)
        self._ck(lib().svm_build_from_cell_points(self.h, C.c_int32(sys), _dp(pts)), "svm_build_from_cell_points")

    def build_voronoi_from_cell_points(self, sys, points):
        """Voronoi model: triangulate the generators, vertices at the circumcentres."""
        pts = np.ascontiguousarray(points, np.float64)
        self._ck(lib().svm_build_voronoi_from_cell_points(self.h, C.c_int32(sys), _dp(pts)),
                 "svm_build_voronoi_from_cell_points")

    def voronoi_forces(self):
        """Generator forces of every sphere (built as Voronoi models); returns a list of [N][3] arrays."""
        self._ck(lib().svm_voronoi_forces(self.h), "svm_voronoi_forces")
        out = []
        for s in range(self.n):
            f = np.zeros((self.n_cells[s], 3), np.float64)
            self._ck(lib().svm_get_voronoi_forces(self.h, C.c_int32(s), _dp(f)), "svm_get_voronoi_forces")
            out.append(f)
        return out

    # ---- state through caller-owned device buffers (e.g. torch tensors) ----
    def _dev(self, fn, sys, buf):
        ptr = buf.data_ptr() if hasattr(buf, "data_ptr") else int(buf)
        self._ck(getattr(lib(), fn)(self.h, C.c_int32(sys), C.c_void_p(ptr)), fn)

    def get_positions_device(self, sys, out):
        """Copy sphere sys's vertex positions into device buffer `out` ([V][3] float64,
        e.g. a contiguous torch tensor on the GPU), on the context's stream."""
        self._dev("svm_get_positions_device", sys, out)

    def set_positions_device(self, sys, src):
        self._dev("svm_set_positions_device", sys, src)

    def get_directors_device(self, sys, out):
        self._dev("svm_get_directors_device", sys, out)

    def set_directors_device(self, sys, src):
        self._dev("svm_set_directors_device", sys, src)

    def get_velocities_device(self, sys, out):
        self._dev("svm_get_velocities_device", sys, out)

    def set_velocities_device(self, sys, src):
        self._dev("svm_set_velocities_device", sys, src)

    def get_forces_device(self, sys, out):
        self._dev("svm_get_forces_device", sys, out)

    def voronoi_brownian_steps(self, n):
        """n Brownian steps of every (Voronoi-model) sphere's generators on the GPU,
        each followed by the GPU retriangulation; dt, mu, T from set_brownian."""
        self._ck(lib().svm_voronoi_brownian_steps(self.h, C.c_int32(n)), "svm_voronoi_brownian_steps")

    def generators(self, sys):
        return self._get3("svm_get_generators", sys, self.n_cells[sys])

    def delaunay_triangles(self, sys):
        out = np.zeros((2 * self.n_cells[sys] - 4, 3), np.int32)
        self._ck(lib().svm_get_delaunay_triangles(self.h, C.c_int32(sys), _ip(out)), "svm_get_delaunay_triangles")
        return out

    # ---- state ----
    def _get3(self, fn, sys, rows):
        out = np.zeros((rows, 3), np.float64)
        self._ck(getattr(lib(), fn)(self.h, C.c_int32(sys), _dp(out)), fn)
        return out

    def _set3(self, fn, sys, a):
        a = np.ascontiguousarray(a, np.float64)
        self._ck(getattr(lib(), fn)(self.h, C.c_int32(sys), _dp(a)), fn)

    def positions(self, sys):
        return self._get3("svm_get_positions", sys, self.n_vertices[sys])

    def set_positions(self, sys, a):
        self._set3("svm_set_positions", sys, a)

    def directors(self, sys):
        return self._get3("svm_get_directors", sys, self.n_vertices[sys])

    def set_directors(self, sys, a):
        self._set3("svm_set_directors", sys, a)

    def velocities(self, sys):
        return self._get3("svm_get_velocities", sys, self.n_vertices[sys])

    def set_velocities(self, sys, a):
        self._set3("svm_set_velocities", sys, a)

    def forces(self, sys):
        return self._get3("svm_get_forces", sys, self.n_vertices[sys])

    def cell_positions(self, sys):
        return self._get3("svm_get_cell_positions", sys, self.n_cells[sys])

    def area_perimeter(self, sys):
        out = np.zeros((self.n_cells[sys], 2), np.float64)
        self._ck(lib().svm_get_area_perimeter(self.h, C.c_int32(sys), _dp(out)), "svm_get_area_perimeter")
        return out

    def set_preferences(self, sys, prefs):
        prefs = np.ascontiguousarray(prefs, np.float64)
        self._ck(lib().svm_set_preferences(self.h, C.c_int32(sys), _dp(prefs)), "svm_set_preferences")

    def set_preferred_parameters(self, sys, a0, p0):
        self._ck(lib().svm_set_preferred_parameters(self.h, C.c_int32(sys), C.c_double(a0), C.c_double(p0)),
                 "svm_set_preferred_parameters")

    def set_kr(self, sys, kr):
        self._ck(lib().svm_set_scalar_model_parameter(self.h, C.c_int32(sys), C.c_double(kr)), "set Kr")

    def set_t1_threshold(self, sys, t):
        self._ck(lib().svm_set_t1_threshold(self.h, C.c_int32(sys), C.c_double(t)), "set threshold")

    # ---- operations ----
    def compute_geometry(self):
        self._ck(lib().svm_compute_geometry(self.h), "svm_compute_geometry")

    def compute_forces(self, precision=0):
        self._ck(lib().svm_compute_forces(self.h, C.c_int32(precision)), "svm_compute_forces")

    def move_particles(self, sys, disp, scale=1.0):
        disp = np.ascontiguousarray(disp, np.float64)
        self._ck(lib().svm_move_particles(self.h, C.c_int32(sys), _dp(disp), C.c_double(scale)),
                 "svm_move_particles")

    def enforce_topology(self):
        self._ck(lib().svm_enforce_topology(self.h), "svm_enforce_topology")

    def energies(self):
        pe = np.zeros(self.n, np.float64)
        ke = np.zeros(self.n, np.float64)
        self._ck(lib().svm_get_energies(self.h, _dp(pe), _dp(ke)), "svm_get_energies")
        return pe, ke

    def counters(self):
        """Per sphere (rows): T1 flips, fan-area fallbacks, skipped NaN force contributions,
        flips skipped for want of a fourth cell, sequential-pass T1 steps, group-run T1 steps."""
        out = np.zeros((self.n, COUNTER_COLS), np.int64)
        self._ck(lib().svm_get_counters(self.h, out.ctypes.data_as(C.POINTER(C.c_int64))), "svm_get_counters")
        return out

    def set_brownian(self, sys, dt=0.01, mu=1.0, T=0.0):
        self._ck(lib().svm_set_brownian(self.h, C.c_int32(sys), C.c_double(dt), C.c_double(mu), C.c_double(T)),
                 "svm_set_brownian")

    def set_active(self, sys, dt=0.01, mu=1.0, v0=0.01, tau=1.0):
        self._ck(lib().svm_set_active(self.h, C.c_int32(sys), C.c_double(dt), C.c_double(mu), C.c_double(v0),
                                      C.c_double(tau)), "svm_set_active")

    def brownian_steps(self, n):
        self._ck(lib().svm_brownian_steps(self.h, C.c_int32(n)), "svm_brownian_steps")

    def set_vicsek(self, sys, dt=0.01, mu=1.0, v0=0.01, tau=1.0, eta=1.0):
        """sphericalVectorialVicsek parameters (defaults as sphericalVectorialVicsek.h:11, dt as the drivers')."""
        self._ck(lib().svm_set_vicsek(self.h, C.c_int32(sys), C.c_double(dt), C.c_double(mu), C.c_double(v0),
                                      C.c_double(tau), C.c_double(eta)), "svm_set_vicsek")

    def vicsek_steps(self, n):
        self._ck(lib().svm_vicsek_steps(self.h, C.c_int32(n)), "svm_vicsek_steps")

    def active_steps(self, n):
        self._ck(lib().svm_active_steps(self.h, C.c_int32(n)), "svm_active_steps")

    def run_graph(self, updater, steps_per_graph, replays):
        self._ck(lib().svm_run_graph(self.h, C.c_int32(updater), C.c_int32(steps_per_graph), C.c_int32(replays)),
                 "svm_run_graph")

    def fire_set_parameters(self, sys, dt=0.01, alpha_start=0.99, dt_max=0.1, dt_inc=1.05, dt_dec=0.95,
                            alpha_dec=0.9, n_min=5, max_iterations=1000, force_cutoff=1e-7, alpha_min=0.0,
                            stop_rms=0):
        p = FireParams(dt, alpha_start, dt_max, dt_inc, dt_dec, alpha_dec, n_min, max_iterations, force_cutoff,
                       alpha_min, stop_rms, 0)
        self._ck(lib().svm_fire_set_parameters(self.h, C.c_int32(sys), C.byref(p)), "svm_fire_set_parameters")

    def fire_set_max_iterations(self, sys, max_iterations):
        self._ck(lib().svm_fire_set_max_iterations(self.h, C.c_int32(sys), C.c_int32(max_iterations)),
                 "svm_fire_set_max_iterations")

    def fire_minimize(self, check_every=50):
        self._ck(lib().svm_fire_minimize(self.h, C.c_int32(check_every)), "svm_fire_minimize")

    def fire_state(self, sys):
        st = FireState()
        self._ck(lib().svm_fire_get_state(self.h, C.c_int32(sys), C.byref(st)), "svm_fire_get_state")
        return {k: getattr(st, k) for k, _ in FireState._fields_ if k != "reserved"}

    def adam_set_parameters(self, sys, beta1=0.9, beta2=0.99, epsilon=1e-8, dt=1e-4, force_cutoff=1e-12,
                            max_iterations=1000):
        """setAdamParameters + setMaximumIterations (defaults as energyMinimizerAdam.h:19); resets the moments."""
        p = AdamParams(beta1, beta2, epsilon, dt, force_cutoff, max_iterations, 0)
        self._ck(lib().svm_adam_set_parameters(self.h, C.c_int32(sys), C.byref(p)), "svm_adam_set_parameters")

    def adam_minimize(self, check_every=50):
        self._ck(lib().svm_adam_minimize(self.h, C.c_int32(check_every)), "svm_adam_minimize")

    def adam_set_max_iterations(self, sys, max_iterations):
        self._ck(lib().svm_adam_set_max_iterations(self.h, C.c_int32(sys), C.c_int32(max_iterations)),
                 "svm_adam_set_max_iterations")

    def adam_state(self, sys):
        st = AdamState()
        self._ck(lib().svm_adam_get_state(self.h, C.c_int32(sys), C.byref(st)), "svm_adam_get_state")
        return {k: getattr(st, k) for k, _ in AdamState._fields_}

    def dyn_set_reference(self):
        """dynamicalFeatures(initialPos): snapshot every sphere's current cell positions."""
        self._ck(lib().svm_dyn_set_reference(self.h), "svm_dyn_set_reference")

    def dyn_measure(self, fraction=1.0, cutoff=0.5):
        """(msd, overlap) per sphere against the snapshot: computeMSD / computeOverlapFunction."""
        msd = np.zeros(self.n, np.float64)
        ov = np.zeros(self.n, np.float64)
        self._ck(lib().svm_dyn_measure(self.h, C.c_double(fraction), C.c_double(cutoff), _dp(msd), _dp(ov)),
                 "svm_dyn_measure")
        return msd, ov

    def synchronize(self):
        self._ck(lib().svm_synchronize(self.h), "svm_synchronize")

    def set_stream(self, stream_ptr):
        self._ck(lib().svm_set_stream(self.h, C.c_void_p(stream_ptr)), "svm_set_stream")

    def set_tuning(self, cluster_blocks=256, graph_steps=10, xcd_mapping=1):
        self._ck(lib().svm_set_tuning(self.h, C.c_int32(cluster_blocks), C.c_int32(graph_steps),
                                      C.c_int32(xcd_mapping)), "svm_set_tuning")

    def tuning(self):
        a, b, c = C.c_int32(0), C.c_int32(0), C.c_int32(0)
        self._ck(lib().svm_get_tuning(self.h, C.byref(a), C.byref(b), C.byref(c)), "svm_get_tuning")
        return dict(cluster_blocks=a.value, graph_steps=b.value, xcd_mapping=c.value, lanes=self.lanes()[0],
                    vertex_form=self.vertex_form(), step_form=self.step_form(), geometry_form=self.geometry_form())

    def set_vertex_form(self, form):
        """0: a thread per vertex; 1, 2: corner form (three wavefronts per 64 vertices) compiled for
        4 or 5 wavefronts per SIMD; 3: form 0 compiled for 3 wavefronts per SIMD.  Same results, bit for bit."""
        self._ck(lib().svm_set_vertex_form(self.h, C.c_int32(form)), "svm_set_vertex_form")

    def vertex_form(self):
        f = C.c_int32(0)
        self._ck(lib().svm_get_vertex_form(self.h, C.byref(f)), "svm_get_vertex_form")
        return f.value

    def set_geometry_form(self, form):
        """Cells per thread of the geometry pass: 1 (form 0), 2 (form 1) or 4 (form 2) per 256-row tile's
        workgroup.  Same results, bit for bit."""
        self._ck(lib().svm_set_geometry_form(self.h, C.c_int32(form)), "svm_set_geometry_form")

    def geometry_form(self):
        f = C.c_int32(0)
        self._ck(lib().svm_get_geometry_form(self.h, C.byref(f)), "svm_get_geometry_form")
        return f.value

    def set_step_form(self, form):
        """Launch structure of a step, a sum of flags: 1 the area check and fan fallback inside the
        vertex pass (no k_fan_fix launch), 2 the blocked-candidate check inside the T1 commit (no
        k_t1_group launch; overlapping dry runs take the sequential pass).  Same results, bit for bit."""
        self._ck(lib().svm_set_step_form(self.h, C.c_int32(form)), "svm_set_step_form")

    def step_form(self):
        f = C.c_int32(0)
        self._ck(lib().svm_get_step_form(self.h, C.byref(f)), "svm_get_step_form")
        return f.value

    def set_lanes(self, n_lanes):
        """Cut the batch into n_lanes sphere ranges whose steps run on streams of their own."""
        self._ck(lib().svm_set_lanes(self.h, C.c_int32(n_lanes)), "svm_set_lanes")

    def set_lane_spheres(self, first_sphere):
        """Lanes over the given sphere ranges: lane k starts at sphere first_sphere[k]."""
        first = np.ascontiguousarray(first_sphere, np.int32)
        self._ck(lib().svm_set_lane_spheres(self.h, C.c_int32(len(first)), _ip(first)), "svm_set_lane_spheres")

    def lanes(self):
        """(number of lanes, first sphere of each lane)."""
        n = C.c_int32(0)
        first = np.zeros(8, np.int32)
        self._ck(lib().svm_get_lanes(self.h, C.byref(n), _ip(first)), "svm_get_lanes")
        return n.value, first[:n.value].copy()

    def autotune(self, updater=1, steps_per_trial=60):
        """Returns the trial table [(lanes, cluster_blocks, graph_steps, xcd_mapping, vertex_form, step_form,
        geometry_form, ms/step)]."""
        tab = np.zeros((32, TUNE_COLS), np.float64)
        n = C.c_int32(0)
        self._ck(lib().svm_autotune(self.h, C.c_int32(updater), C.c_int32(steps_per_trial), _dp(tab),
                                    C.c_int32(32), C.byref(n)), "svm_autotune")
        return tab[:n.value]

    def save_tuning(self, path, updater=1):
        """Store the current schedule under this batch's key (svm_tuning_save)."""
        self._ck(lib().svm_tuning_save(self.h, C.c_int32(updater), os.fsencode(path)), "svm_tuning_save")

    def load_tuning(self, path, updater=1):
        """Apply the schedule stored under this batch's key; returns whether one was found."""
        found = C.c_int32(0)
        self._ck(lib().svm_tuning_load(self.h, C.c_int32(updater), os.fsencode(path), C.byref(found)),
                 "svm_tuning_load")
        return bool(found.value)

    def cell_descriptors(self, sys):
        """(X float32 [N][16], frames float32 [N][6]) of sphere sys."""
        n = self.n_cells[sys]
        X = np.zeros((n, 16), np.float32)
        F = np.zeros((n, 6), np.float32)
        self._ck(lib().svm_cell_descriptors(self.h, C.c_int32(sys), _fp(X), _fp(F)), "svm_cell_descriptors")
        return X, F

    def warm_start(self, mlp, scale=1.0, max_step=0.3):
        """Move by the network's displacements, each capped at max_step x the vertex's shortest edge
        (<= 0: no cap).  Returns the number of spheres whose move was undone (a cell folded)."""
        rev = C.c_int32(0)
        self._ck(lib().svm_warm_start(self.h, mlp.h, C.c_double(scale), C.c_double(max_step), C.byref(rev)),
                 "svm_warm_start")
        return rev.value

    def set_force_precision(self, precision):
        """0: FP64 reference expressions (parity), 1: FP32, 2: FP64 vector form."""
        self._ck(lib().svm_set_force_precision(self.h, C.c_int32(precision)), "svm_set_force_precision")

    def force_error(self):
        """Per sphere: FP32 vs FP64 reference expressions (max|dF|/max|F|, rms|dF|/rms|F|, max|dF|,
        rms|F|), FP32 vs FP64 vector form (max, rms), FP64 vector form vs reference (rms)."""
        out = np.zeros((self.n, FORCE_ERR_COLS), np.float64)
        self._ck(lib().svm_get_force_error(self.h, _dp(out)), "svm_get_force_error")
        return out

    def time_vertex_pass(self, updater, reps=20):
        """Mean ms of one launch of the fused vertex pass (0 Brownian, 1 active)."""
        ms = C.c_double(0.0)
        self._ck(lib().svm_time_vertex_pass(self.h, C.c_int32(updater), C.c_int32(reps), C.byref(ms)),
                 "svm_time_vertex_pass")
        return ms.value

    def t1_diagnostics(self):
        out = np.zeros((self.n, T1_DIAG_COLS), np.int64)
        self._ck(lib().svm_get_t1_diagnostics(self.h, out.ctypes.data_as(C.POINTER(C.c_int64))), "diagnostics")
        return out

    def set_t1_mode(self, sequential):
        self._ck(lib().svm_set_t1_mode(self.h, C.c_int32(1 if sequential else 0)), "svm_set_t1_mode")

    def set_t1_component_distance(self, cells):
        """Flippable candidates closer than `cells` cell sizes share one T1 dry run (0: one run each).
        Same trajectory for every value."""
        self._ck(lib().svm_set_t1_component_distance(self.h, C.c_double(cells)), "svm_set_t1_component_distance")

    def t1_component_distance(self):
        x = C.c_double(0.0)
        self._ck(lib().svm_get_t1_component_distance(self.h, C.byref(x)), "svm_get_t1_component_distance")
        return x.value

    def set_profiling(self, on):
        self._ck(lib().svm_set_profiling(self.h, C.c_int32(1 if on else 0)), "svm_set_profiling")

    def timing(self):
        n = 32
        names = (C.c_char_p * n)()
        ms = np.zeros(n, np.float64)
        k = C.c_int32(0)
        self._ck(lib().svm_get_timing(self.h, names, _dp(ms), C.c_int32(n), C.byref(k)), "svm_get_timing")
        return {names[i].decode(): float(ms[i]) for i in range(k.value)}


class VectorValueNetCDF:
    """vectorValueNetCDF(fn, N, NcFile::Replace) (src/databases/vectorValueNetCDF.cpp:3-36):
    writeState(vec, value) appends one record of a NetCDF classic file with
    value(record, unit) and vector(record, N)."""

    def __init__(self, path, n):
        self.h = C.c_void_p()
        r = lib().svm_nc_open(os.fsencode(path), C.c_int32(n), C.byref(self.h))
        if r:
            raise SVMError(f"svm_nc_open({path!r}, {n}) failed with status {r}")
        self.n = n

    def writeState(self, vec, value):
        v = np.ascontiguousarray(vec, np.float64).ravel()
        if v.size != self.n:
            raise ValueError(f"vector of {v.size} values for N = {self.n}")
        r = lib().svm_nc_write(self.h, _dp(v), C.c_double(value))
        if r:
            raise SVMError(f"svm_nc_write failed with status {r}")

    @property
    def num_records(self):
        n = C.c_int32(0)
        lib().svm_nc_num_records(self.h, C.byref(n))
        return n.value

    def close(self):
        if self.h:
            lib().svm_nc_close(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

