/* svm.h -- C ABI of the MI355X spherical vertex-model library (libsvm.so).
 *
 * One context holds a batch of independent spheres ("systems") resident in
 * HBM.  Every entry point takes plain pointers and sizes, returns an int
 * status (SVM_OK == 0) and never throws.  Work is queued on the context's
 * HIP stream; entry points that return values to the host synchronise that
 * stream (their names say "get" or they document it).
 *
 * Host arrays use the reference's layouts:
 *   positions / directors / velocities / forces / displacements : double[3*V]
 *     (one dVec of three doubles per vertex, inc/dDimensionalVectorTypes.h:19-34)
 *   vertex_neighbors, vertex_cell_neighbors : int32[3*V] (stride 3,
 *     src/model/convexHullCGAL.cpp:131-150)
 *   cell_vertices : int32[n_cells*ring_stride], entry (k, c) at c*ring_stride+k
 *     (Index2D, inc/indexer.h:41-44; src/model/convexHullCGAL.cpp:163-185)
 *   cell_vertex_counts : int32[n_cells]
 *   area_perimeter, preferences : double[2*n_cells] as (x, y) pairs (scalar2)
 * All indices are local to their system (0-based).
 *
 * The reference's kernel-caller layer each entry point replaces is cited
 * beside it (paths relative to the reference repository root).
 */
#ifndef SVM_H
#define SVM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct svm_context svm_context;

enum svm_status {
    SVM_OK = 0,
    SVM_ERR_ARG = 1,            /* bad argument (null pointer, index, size) */
    SVM_ERR_HIP = 2,            /* a HIP runtime call failed */
    SVM_ERR_NO_DEVICE = 3,      /* no gfx950 device visible */
    SVM_ERR_RING_OVERFLOW = 4,  /* a T1 would grow a cell beyond ring_capacity */
    SVM_ERR_TOPOLOGY = 5,       /* inconsistent topology (ring/vertex lists disagree) */
    SVM_ERR_DELAUNAY = 6,       /* spherical Delaunay construction failed */
    SVM_ERR_STATE = 7,          /* call out of order (e.g. no topology yet) */
    SVM_ERR_CAPACITY = 8,       /* a fixed-size work buffer would overflow */
    SVM_ERR_IO = 9              /* a file could not be opened or written */
};

/* One sphere of the batch.  Mirrors the constructor arguments of
 * sphericalVertexModel(n, noise, a0, p0, ...) (src/model/sphericalVertexModel.h:13)
 * plus the values its constructor fixes (src/model/sphericalVertexModel.cpp:12-13,79-80). */
typedef struct {
    int32_t n_cells;        /* N, number of cells */
    int32_t ring_capacity;  /* max vertices per cell the ring storage holds (0 -> 32) */
    double a0, p0;          /* preferred area and perimeter for every cell */
    double kr;              /* K_A (Kr); K_P is 1 */
    double t1_threshold;    /* T1 edge-length threshold (reference default 0.1) */
    uint64_t seed;          /* key of this sphere's Philox noise streams */
} svm_system_desc;

/* Per-sphere FIRE parameters: the arguments of
 * energyMinimizerFIRE::setFIREParameters (src/updaters/energyMinimizerFIRE.cpp:262-273)
 * and setMaximumIterations.  stop_rms = 0 keeps the reference stop test
 * sqrt(sum F^2)/V <= force_cutoff (energyMinimizerFIRE.cpp:199); 1 uses the
 * rms force sqrt(sum F^2 / V). */
typedef struct {
    double dt, alpha_start, dt_max, dt_inc, dt_dec, alpha_dec;
    int32_t n_min;
    int32_t max_iterations;
    double force_cutoff, alpha_min;
    int32_t stop_rms;
    int32_t reserved;
} svm_fire_params;

/* FIRE state of one sphere (energyMinimizerFIRE.h:71-96). */
typedef struct {
    int32_t iterations, n_since_negative_power, converged, reserved;
    double dt, alpha, force_max, power;
} svm_fire_state;

/* Adam minimiser parameters (energyMinimizerAdam::setAdamParameters and
 * setMaximumIterations, src/updaters/energyMinimizerAdam.h:19-31; the
 * reference's defaults are beta1 0.9, beta2 0.99, epsilon 1e-8, dt 1e-4,
 * force_cutoff 1e-12). */
typedef struct {
    double beta1, beta2, epsilon, dt, force_cutoff;
    int32_t max_iterations, reserved;
} svm_adam_params;

/* Adam state of one sphere (energyMinimizerAdam.h:33-62). */
typedef struct {
    int32_t iterations, converged;
    double force_max, beta1t, beta2t;
} svm_adam_state;

/* ---- lifetime ---------------------------------------------------------- */
/* Allocate device storage for n_systems spheres on the current HIP device. */
int svm_create(const svm_system_desc* systems, int32_t n_systems, svm_context** out);
int svm_destroy(svm_context* ctx);
/* Use `stream` (a hipStream_t, NULL = the context's own stream) for all work. */
int svm_set_stream(svm_context* ctx, void* stream);
int svm_synchronize(svm_context* ctx);
/* Vertex count (2N-4) of system `sys` once its topology is set. */
int svm_get_num_vertices(svm_context* ctx, int32_t sys, int32_t* n_vertices);
/* Text of the last error on this context (static storage, never NULL). */
const char* svm_last_error(svm_context* ctx);

/* ---- initial topology -------------------------------------------------- */
/* Spherical Delaunay triangulation of the n_cells cell centres on the GPU
 * (exact in-circle predicates), then the vertex model's topology and vertex
 * positions (normalised triangle centroids) from it.  Replaces
 * convexHullCGALInterface::sphericalConvexHullForVertexModel
 * (src/model/convexHullCGAL.cpp:82-186) as called by the constructor
 * (src/model/sphericalVertexModel.cpp:14-18).  cell_points: double[3N]. */
int svm_build_from_cell_points(svm_context* ctx, int32_t sys, const double* cell_points);
/* Spherical Delaunay triangulation of n_points points on the GPU, with no
 * model attached: the per-move retriangulation of sphericalVoronoi::convexHull
 * (src/model/sphericalVoronoi.cpp:8-12, src/model/convexHullCGAL.cpp:15-65)
 * and its neighbour search (hyperrectangularCellList,
 * src/utilities/hyperrectangularCellList.cpp:14-51, 165-342).  Exact
 * predicates.  Input whose triangulation is not unique (four co-circular
 * points on an empty circle) is rejected (SVM_ERR_TOPOLOGY); so may be input
 * with other exactly co-circular quadruples near a point, when the exact
 * tournament compares them.  Each point searches its neighbours within a
 * radius of its own: 3.5 mean spacings first, doubled where the star cannot
 * be proven within it (voids), bisected down where it holds more than 512
 * candidates (dense clusters); input for which no radius does both (a point
 * with more than 512 others inside the smallest ball that proves its star)
 * is rejected with SVM_ERR_TOPOLOGY.  triangles: int32[3*(2N-4)],
 * counter-clockwise seen from outside, smallest index first, sorted
 * lexicographically. */
int svm_delaunay_triangles(int32_t n_points, const double* points, int32_t* triangles);
/* svm_delaunay_triangles without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, same triangles and the same status codes as
 * svm_delaunay_triangles, plus `stream` (a hipStream_t; NULL = the null
 * stream), on which the work runs; the call synchronises it.  Its device work
 * memory is one static array of SVM_NOALLOC_SCRATCH_BYTES per GPU, held until
 * the process ends, so n_points is limited to SVM_NOALLOC_MAX_POINTS (larger
 * input: SVM_ERR_CAPACITY).  Every call on a GPU uses that GPU's one array:
 * calls that overlap on one GPU (from several threads or streams) race. */
#define SVM_NOALLOC_MAX_POINTS 65536
#define SVM_NOALLOC_SCRATCH_BYTES (40u << 20)
int svm_delaunay_triangles_noalloc(int32_t n_points, const double* points, int32_t* triangles, void* stream);
/* Cell list of points on the GPU (the binning of hyperrectangularCellList,
 * src/utilities/hyperrectangularCellList.cpp:165-342, on a cube grid
 * [-L, L]^3 with L = (1 + 1e-9) * max |coordinate|): grid_dim G bins per axis,
 * G = min(256, ceil(2L / cell_size)), bin side 2L/G; bin of point i is
 * (gx*G + gy)*G + gz.  bin_start: int32[G^3 + 1], sorted_points: the points of
 * bin b at sorted_points[bin_start[b] .. bin_start[b+1]), in index order.
 * With bin_of, bin_start or sorted_points NULL only grid_dim is returned. */
int svm_cell_list(int32_t n_points, const double* points, double cell_size, int32_t* grid_dim, int32_t* bin_of,
                  int32_t* bin_start, int32_t* sorted_points);
/* Metric neighbour lists (neighborList::computeNeighborLists,
 * src/utilities/neighborList.cpp:102-232 and neighborList.cu:108-355; the
 * list sphericalModel::setRadius builds with range 1 and sphereRadius R + 2.5,
 * src/model/sphericalModel.cpp:21): for each point i the points j != i with
 * |d| < range, d = p_i - p_j and |d| = sqrt(dx*dx + dy*dy + dz*dz), as the
 * reference's minDist and norm form them.  box_half is the reference's
 * sphereRadius, the half side of its periodic box; only boxes where no
 * periodic image changes the result are accepted (box_half >= range and
 * 2 box_half - 2 max|coordinate| >= range, which R + 2.5 with range 1 meets),
 * otherwise SVM_ERR_ARG.  counts: int32[n]; with max_neighbors >= the largest
 * count, indices int32[n * max_neighbors], vectors double[3 * n *
 * max_neighbors] (d) and distances double[n * max_neighbors] (|d|), entry k of
 * point i at i * max_neighbors + k, each point's neighbours in increasing
 * index order (the reference's order is that of its atomics or cell scan),
 * unused entries zero.  When some count exceeds max_neighbors only counts
 * are written and SVM_ERR_CAPACITY is returned; *needed is the largest count
 * either way (the reference's recompute with a larger Nmax). */
int svm_neighbor_list(int32_t n_points, const double* points, double range, double box_half, int32_t max_neighbors,
                      int32_t* counts, int32_t* indices, double* vectors, double* distances, int32_t* needed);
/* Seeded cell centres: sphericalModel::setParticlePositionsRandomly + setRadius
 * (src/model/sphericalModel.cpp:11-51) with Philox draws keyed by `seed`. */
int svm_generate_cell_points(int32_t n_cells, uint64_t seed, double* cell_points);
/* A spatially coherent labelling of points (no counterpart in the reference,
 * whose labels are the generator's draw order): order[i] is the index of the
 * point that should get label i, sorting by the Z-order (Morton) key of the
 * points in their bounding cube, ties by index.  Relabelling the cell points
 * this way before building the topology puts the cells of a 256-row tile, and
 * (through the canonical triangle order) their vertices, close together in
 * memory.  The model itself is unchanged; only labels move. */
int svm_spatial_order(int32_t n_points, const double* points, int32_t* order);
/* Seeded unit tangent directors (src/model/sphericalVertexModel.cpp:58-73). */
int svm_generate_directors(int32_t n_vertices, const double* positions, uint64_t seed, double* directors);

/* Topology from a spherical Delaunay triangulation computed elsewhere:
 * triangles int32[3*n_triangles] (any orientation and order; n_triangles must
 * be 2N-4), cell_points double[3N].  Builds vertex positions and lists in the
 * conventions of sphericalConvexHullForVertexModel
 * (src/model/convexHullCGAL.cpp:82-186). */
int svm_set_topology_from_triangles(svm_context* ctx, int32_t sys, const double* cell_points,
                                    int32_t n_triangles, const int32_t* triangles);
/* Upload a complete topology in the reference layout (see file header). */
int svm_set_topology(svm_context* ctx, int32_t sys, int32_t n_vertices,
                     const double* vertex_positions, const int32_t* vertex_neighbors,
                     const int32_t* vertex_cell_neighbors, const int32_t* cell_vertices,
                     int32_t ring_stride, const int32_t* cell_vertex_counts);
int svm_get_topology(svm_context* ctx, int32_t sys, int32_t* vertex_neighbors,
                     int32_t* vertex_cell_neighbors, int32_t* cell_vertices,
                     int32_t ring_stride, int32_t* cell_vertex_counts);
/* Spherical Delaunay triangles of the last svm_build_from_cell_points call:
 * int32[3*(2N-4)], each counter-clockwise seen from outside, rotated so the
 * smallest index is first, sorted lexicographically. */
int svm_get_delaunay_triangles(svm_context* ctx, int32_t sys, int32_t* triangles);

/* ---- state ------------------------------------------------------------- */
int svm_set_positions(svm_context* ctx, int32_t sys, const double* positions);
int svm_get_positions(svm_context* ctx, int32_t sys, double* positions);
int svm_set_directors(svm_context* ctx, int32_t sys, const double* directors);
int svm_get_directors(svm_context* ctx, int32_t sys, double* directors);
int svm_set_velocities(svm_context* ctx, int32_t sys, const double* velocities);
int svm_get_velocities(svm_context* ctx, int32_t sys, double* velocities);
int svm_get_forces(svm_context* ctx, int32_t sys, double* forces);
/* The same state through caller-owned device buffers (double[3*V], device
 * memory of this process's HIP runtime, e.g. a torch tensor's data_ptr()
 * when torch was loaded first): a copy kernel enqueued on the context's
 * stream, ordered with the context's other work and not synchronised --
 * synchronise that stream (svm_synchronize) or work on it before the buffer
 * is read elsewhere.  The reference's kernel callers take device pointers of
 * GPUArray buffers (src/model/sphericalVertexModel.cuh:14-92); these let a
 * caller keep the state on the device.  A pointer that is not device memory
 * of this device, or whose allocation ends before 3*V doubles, is refused
 * with SVM_ERR_ARG; a sub-buffer of a larger allocation (a tensor of a
 * caching allocator) is the caller's to size. */
int svm_set_positions_device(svm_context* ctx, int32_t sys, const double* positions);
int svm_get_positions_device(svm_context* ctx, int32_t sys, double* positions);
int svm_set_directors_device(svm_context* ctx, int32_t sys, const double* directors);
int svm_get_directors_device(svm_context* ctx, int32_t sys, double* directors);
int svm_set_velocities_device(svm_context* ctx, int32_t sys, const double* velocities);
int svm_get_velocities_device(svm_context* ctx, int32_t sys, double* velocities);
int svm_get_forces_device(svm_context* ctx, int32_t sys, double* forces);
int svm_get_cell_positions(svm_context* ctx, int32_t sys, double* cell_positions);
int svm_get_area_perimeter(svm_context* ctx, int32_t sys, double* area_perimeter);
/* sphericalVertexModel::setPreferredParameters (src/model/sphericalVertexModel.cpp:103-109). */
int svm_set_preferred_parameters(svm_context* ctx, int32_t sys, double a0, double p0);
/* Per-cell (a0, p0) pairs, double[2N]. */
int svm_set_preferences(svm_context* ctx, int32_t sys, const double* preferences);
/* setScalarModelParameter (Kr) and setT1Threshold (src/model/sphericalVertexModel.h:82-85). */
int svm_set_scalar_model_parameter(svm_context* ctx, int32_t sys, double kr);
int svm_set_t1_threshold(svm_context* ctx, int32_t sys, double threshold);

/* ---- model operations (all systems of the context) ---------------------- */
/* computeGeometry: gpu_spherical_vertex_model_geometry
 * (src/model/sphericalVertexModel.cuh:34-48), CPU semantics of
 * computeGeometryCPU (src/model/sphericalVertexModel.cpp:209-311). */
int svm_compute_geometry(svm_context* ctx);
/* computeForces: geometry then gpu_quadratic_spherical_cellular_force
 * (src/model/sphericalVertexModel.cuh:50-63; CPU semantics
 * src/model/sphericalVertexModel.cpp:356-420).  precision: 0 = FP64 (parity),
 * 1 = FP32 forces from FP64 positions and cell centres. */
int svm_compute_forces(svm_context* ctx, int32_t precision);
/* Force arithmetic of every later force evaluation, the updaters included
 * (the SCALARFLOAT build switch, inc/std_include.h:66-78, made a run-time
 * choice): 0 = FP64 with the reference's expressions (the parity path),
 * 1 = FP32, 2 = FP64; 1 and 2 use a vector form of the same gradients that
 * stays well conditioned where the reference's expressions cancel (small
 * fan triangles on large spheres).  Positions, cell centres, areas,
 * perimeters, T1 tests and updates stay FP64. */
int svm_set_force_precision(svm_context* ctx, int32_t precision);
/* Error of the FP32 forces on the current state, per sphere
 * (SVM_FORCE_ERR_COLS doubles each): against the FP64 reference expressions
 * max|dF|/max|F|, rms|dF|/rms|F|, max|dF|, rms|F|; against the FP64 vector
 * form max and rms (relative); and the FP64 vector form against the FP64
 * reference expressions, rms (relative).  Leaves the FP32 forces in the
 * force buffer. */
#define SVM_FORCE_ERR_COLS 7
int svm_get_force_error(svm_context* ctx, double* stats);
/* moveParticles(displacements, scale) then enforceTopology:
 * gpu_move_particles_on_sphere + the T1 callers
 * (src/model/sphericalVertexModel.cuh:14-92, sphericalVertexModel.cpp:111-148).
 * displacements: host double[3V] for system `sys`; other systems stay still. */
int svm_move_particles(svm_context* ctx, int32_t sys, const double* displacements, double scale);
/* enforceTopology alone (src/model/sphericalVertexModel.cpp:422-433). */
int svm_enforce_topology(svm_context* ctx);
/* Energies after the last geometry: potential sum Kr(A-A0)^2+(P-P0)^2
 * (computeEnergy, sphericalVertexModel.cpp:150-160) and kinetic
 * sum m v^2/2 (simpleModel.cpp:63-73); double[n_systems] each (may be NULL). */
int svm_get_energies(svm_context* ctx, double* potential, double* kinetic);
/* Counters since creation, int64[SVM_COUNTER_COLS*n_systems]: T1 flips, fan-area
 * fallbacks, skipped NaN force contributions, T1 flips skipped because the
 * edge's two vertices share all three cells (no fourth cell exists), steps
 * whose T1 flips were applied by the sequential pass, and steps whose
 * overlapping candidates were resolved together by one group dry run. */
#define SVM_COUNTER_COLS 6
int svm_get_counters(svm_context* ctx, int64_t* counters);

/* ---- updaters ---------------------------------------------------------- */
/* brownianDynamics (src/updaters/brownianDynamics.cpp:29-48; kernel caller
 * gpu_brownian_eom_integration, src/updaters/brownianDynamics.cuh:15-22). */
int svm_set_brownian(svm_context* ctx, int32_t sys, double dt, double mu, double temperature);
int svm_brownian_steps(svm_context* ctx, int32_t n_steps);
/* sphericalSelfPropelledParticle (src/updaters/sphericalSelfPropelledParticle.cpp:16-47,
 * parameters src/updaters/vectorialVicsek.h:188-191; tau = 1/Dr). */
int svm_set_active(svm_context* ctx, int32_t sys, double dt, double mu, double v0, double tau);
int svm_active_steps(svm_context* ctx, int32_t n_steps);
/* sphericalVectorialVicsek (src/updaters/sphericalVectorialVicsek.cpp:5-66;
 * the reference's defaults mu 1, Eta 1, tau 1, v0 0.01, sphericalVectorialVicsek.h:11)
 * on the vertex model: vertices move by dt (v0 n + mu F) and the topology is
 * enforced; then each director relaxes at rate dt / tau towards the
 * normalised tangent mean of itself and its three vertex neighbours plus eta
 * times a random unit tangent direction.  The reference draws that direction
 * from its own generator; here it is a Philox stream keyed by (seed, step,
 * vertex). */
int svm_set_vicsek(svm_context* ctx, int32_t sys, double dt, double mu, double v0, double tau, double eta);
int svm_vicsek_steps(svm_context* ctx, int32_t n_steps);
/* Capture `steps_per_graph` steps of the given updater (0 Brownian, 1 active,
 * 2 Vicsek) into a HIP graph and replay it n_replays times. */
int svm_run_graph(svm_context* ctx, int32_t updater, int32_t steps_per_graph, int32_t n_replays);

/* energyMinimizerFIRE (src/updaters/energyMinimizerFIRE.cpp:181-259,
 * velocityVerlet.cpp:6-38; kernel callers gpu_update_velocity_FIRE,
 * gpu_dVec_dot_products, gpu_displacement_velocity_verlet,
 * gpu_update_velocity). */
int svm_fire_set_parameters(svm_context* ctx, int32_t sys, const svm_fire_params* params);
/* Minimise every sphere until its stop test passes or its iteration limit is
 * reached; the host checks convergence every `check_every` iterations. */
int svm_fire_minimize(svm_context* ctx, int32_t check_every);
/* updater::setMaximumIterations (src/updaters/baseUpdater.h:78): the
 * iteration limit only; the minimiser's state (iterations done, dt, alpha,
 * velocities) is kept, so a later svm_fire_minimize continues. */
int svm_fire_set_max_iterations(svm_context* ctx, int32_t sys, int32_t max_iterations);
int svm_fire_get_state(svm_context* ctx, int32_t sys, svm_fire_state* state);

/* Adam (energyMinimizerAdam, src/updaters/energyMinimizerAdam.cpp:7-103; the
 * CPU step is the reference).  svm_adam_set_parameters is setAdamParameters +
 * setMaximumIterations + initializeFromModel for sphere `sys`: moments
 * zeroed, iterations 0, beta1t = beta1, beta2t = beta2.  svm_adam_minimize
 * runs minimize() on every sphere: each iteration computes the forces, moves
 * every vertex by -dt mc / sqrt(vc) per component and enforces the topology,
 * until iterations reach the limit or forceMax = sqrt(sum F^2) / N drops to
 * the cutoff; the host checks every `check_every` iterations.  Iterations and
 * moments persist across calls, as in the reference.  FIRE and Adam share
 * the minimiser's freeze flag: run one of them at a time. */
int svm_adam_set_parameters(svm_context* ctx, int32_t sys, const svm_adam_params* params);
int svm_adam_minimize(svm_context* ctx, int32_t check_every);
/* setMaximumIterations: the limit only; moments and iterations are kept. */
int svm_adam_set_max_iterations(svm_context* ctx, int32_t sys, int32_t max_iterations);
int svm_adam_get_state(svm_context* ctx, int32_t sys, svm_adam_state* state);

/* T1 path: 0 (default) resolves each step's flips by per-candidate dry runs,
 * a group run for overlapping candidates, and the sequential pass only when
 * those cannot complete; 1 always uses the sequential pass.  Both give the
 * reference's sequential order exactly; 1 exists to show that. */
int svm_set_t1_mode(svm_context* ctx, int32_t mode);
/* Dry runs over predicted components (mode 0): flippable candidates of a
 * sphere closer than `cells` cell sizes (sqrt of the mean cell area) to each
 * other, transitively, share one dry run; 0 gives one run per candidate.
 * Default 3.  Candidates that still interact are caught and redone by the
 * group run, so every value gives the same trajectory; the distance only
 * moves work between the dry-run and group kernels.  No counterpart in the
 * reference (whose T1 loop is sequential): a tuning control of this
 * library's parallel T1 path. */
int svm_set_t1_component_distance(svm_context* ctx, double cells);
int svm_get_t1_component_distance(svm_context* ctx, double* cells);
/* T1 diagnostics, int64[6*n_systems]: steps whose sequential pass ran
 * because the flippable-candidate list overflowed, the overlapping candidates
 * were too many for one group, the group run ran out of room, the
 * blocked-candidate list overflowed; then the totals of flippable and of
 * blocked candidates seen; then, for group dry runs, the totals of members,
 * attempts and vertex visits and the longest walk; then, for single-candidate
 * dry runs, the total of visits and the longest walk.
 * `reasons` holds SVM_T1_DIAG_COLS int64 per sphere. */
#define SVM_T1_DIAG_COLS 12
int svm_get_t1_diagnostics(svm_context* ctx, int64_t* reasons);
/* Kernel timing: when on, every eagerly launched kernel is bracketed by HIP
 * events on the context stream (this serialises launches; graphs are never
 * timed this way).  svm_get_timing fills up to n (name, average ms) pairs. */
int svm_set_profiling(svm_context* ctx, int32_t on);
int svm_get_timing(svm_context* ctx, const char** names, double* ms, int32_t n, int32_t* n_written);

/* ---- learned warm start for FIRE ----------------------------------------
 * No counterpart in the reference: a per-cell MLP (12 shape descriptors ->
 * 3 x 64 SiLU -> 2-D tangent displacement of the cell), bf16 weights on MFMA
 * with FP32 accumulation, trained on the GPU (FP32 backward, Adam) on FIRE
 * trajectories.  svm_warm_start moves every vertex by the mean of its three
 * cells' predicted displacements before FIRE (energyMinimizerFIRE, unchanged,
 * remains the fallback and does the convergence).  Parameter vector layout
 * (SVM_MLP_NP floats): W1 [64][16] (inputs 12..15 zero), b1 [64],
 * W2 [64][64], b2 [64], W3 [64][64], b3 [64], W4 [16][64] (rows 2..15 zero),
 * b4 [16] (entries 2..15 zero); a layer computes z = W a + b. */
#define SVM_MLP_NPARAM 10448
#define SVM_MLP_NDESC 12
typedef struct svm_mlp svm_mlp;
int svm_mlp_create(svm_mlp** out);
int svm_mlp_destroy(svm_mlp* mlp);
/* theta: float[SVM_MLP_NPARAM]; mu, sigma: float[12] input normalisation
 * x' = (x - mu) / sigma; the output is out_scale * (W4 a3 + b4).  Resets Adam. */
int svm_mlp_set_parameters(svm_mlp* mlp, const float* theta, const float* mu, const float* sigma, float out_scale);
int svm_mlp_get_parameters(svm_mlp* mlp, float* theta, float* mu, float* sigma, float* out_scale);
/* X: float[n][16] raw descriptors (12 used) -> Y: float[n][2]. */
int svm_mlp_forward(svm_mlp* mlp, const float* X, int32_t n, float* Y);
/* Gradient of L = (1/n) sum_i |Y_i - T_i|^2 with respect to theta (as used
 * in the bf16 forward pass), and L. */
int svm_mlp_gradient(svm_mlp* mlp, const float* X, const float* T, int32_t n, float* grad, double* loss);
/* `epochs` passes of Adam (lr, beta 0.9 / 0.999) over minibatches of `batch`
 * rows in an order shuffled per epoch from `seed`; losses[epoch] (optional)
 * is the mean minibatch loss of that epoch. */
int svm_mlp_train(svm_mlp* mlp, const float* X, const float* T, int32_t n, int32_t epochs, int32_t batch, float lr,
                  uint64_t seed, double* losses);
/* Mean time of the forward kernel over n synthetic rows (device resident). */
int svm_mlp_time_forward(svm_mlp* mlp, int32_t n, int32_t reps, double* ms_per_launch);
/* The 12 descriptors of every cell of sphere `sys` (float[N][16]) after a
 * force evaluation, and the cells' tangent frames (float[N][6]: e1, e2) in
 * which the network's displacements are expressed. */
int svm_cell_descriptors(svm_context* ctx, int32_t sys, float* X, float* frames);
/* Move every vertex of every sphere by scale x (mean of its three cells'
 * predicted displacements), each move capped at max_step x the vertex's
 * shortest edge (max_step <= 0: no cap), then enforce the topology (T1).  A sphere whose cell areas tile it
 * worse afterwards (|sum A / 4 pi R^2 - 1| grown by more than 1e-8: a cell
 * folded) gets its positions and topology back; `reverted` (optional) counts
 * those spheres.
 * The warm start never leaves a sphere worse than plain FIRE would start.
 * Spheres whose minimiser has converged (FIRE or Adam force statistic at or
 * below its cutoff after at least one iteration) are not moved, so the call
 * can also correct the unconverged spheres between minimiser calls. */
int svm_warm_start(svm_context* ctx, svm_mlp* mlp, double scale, double max_step, int32_t* reverted);

/* ---- Voronoi model -------------------------------------------------------
 * The reference's sphericalVoronoi (src/model/sphericalVoronoi.cpp:8-12) has
 * only the topology: the convex hull of the cell points after every move.
 * This adds its forces.  svm_build_voronoi_from_cell_points triangulates the
 * generators on the GPU as svm_build_from_cell_points does, but places each
 * vertex at its triangle's circumcentre on the sphere, R n / |n| with
 * n = (b - a) x (c - a), and keeps the generators; svm_voronoi_forces then
 * computes, for every sphere (all must be built this way), the geometry, the
 * vertex forces F_v = -dE/dv of E = sum Kr (A - A0)^2 + (P - P0)^2 (in the
 * context's force precision) and their chain rule through the circumcentres
 * to the generators, projected on the tangent plane; svm_get_voronoi_forces
 * copies sphere `sys`'s generator forces out ([N][3]).  Call the builder
 * again after moving the generators (the per-move retriangulation). */
int svm_build_voronoi_from_cell_points(svm_context* ctx, int32_t sys, const double* cell_points);
int svm_voronoi_forces(svm_context* ctx);
int svm_get_voronoi_forces(svm_context* ctx, int32_t sys, double* forces);
/* n_steps Brownian steps of the Voronoi model on the GPU, the generators
 * never leaving the device: brownianDynamics::integrateEOMCPU
 * (src/updaters/brownianDynamics.cpp:29-48) on the cells, d = sqrt(2 dt mu T)
 * xi + dt mu F with F the generator forces of svm_voronoi_forces and xi
 * Philox normals keyed by (sphere seed, step, cell), the move of
 * sphericalDomain::move (inc/sphericalDomain.h:203-209) on the model's
 * radius, then the retriangulation of every sphere that
 * sphericalModel::moveParticles does after each move
 * (src/model/sphericalModel.cpp:96-121), on the GPU.  dt, mu and T are the
 * sphere's svm_set_brownian values.  Every sphere must have been built with
 * svm_build_voronoi_from_cell_points.  svm_get_generators copies sphere
 * `sys`'s generators out ([N][3]). */
int svm_voronoi_brownian_steps(svm_context* ctx, int32_t n_steps);
int svm_get_generators(svm_context* ctx, int32_t sys, double* points);

/* ---- trajectory analysis and output -------------------------------------
 * dynamicalFeatures (src/analysis/dynamicalFeatures.cpp:7-53) on the GPU.
 * svm_dyn_set_reference snapshots every sphere's current cell positions (the
 * constructor's initialPos; the geometry pass computes them).
 * svm_dyn_measure gives per sphere (arrays of n_systems) computeMSD and
 * computeOverlapFunction of the current cell positions against it, over the
 * first floor(N * fraction) cells (all for fraction >= 1): the mean squared
 * geodesic displacement (NaN distances add nothing) and the fraction moved
 * less than `cutoff` (the reference's default 0.5). */
int svm_dyn_set_reference(svm_context* ctx);
int svm_dyn_measure(svm_context* ctx, double fraction, double cutoff, double* msd, double* overlap);

/* vectorValueNetCDF (src/databases/vectorValueNetCDF.cpp:3-36): a NetCDF
 * classic file with dimensions record (unlimited), unit (1) and N (n), and
 * double variables value(record, unit) and vector(record, N); opened with
 * replace semantics.  svm_nc_write is writeState(vec[n], value): one record
 * appended and the file flushed.  Host-only (no device is needed). */
typedef struct svm_nc_file svm_nc_file;
int svm_nc_open(const char* path, int32_t n, svm_nc_file** out);
int svm_nc_write(svm_nc_file* f, const double* vec, double value);
int svm_nc_num_records(svm_nc_file* f, int32_t* n_records);
int svm_nc_close(svm_nc_file* f);

/* ---- autotuning ----------------------------------------------------------
 * The launch schedule of a step: the grid of the T1 single-candidate dry runs
 * (cluster_blocks), the number of steps captured per HIP graph replayed by
 * svm_run_graph with steps_per_graph = 0 (graph_steps, even), and whether
 * 256-row tiles are dealt to the chip's 8 XCDs in contiguous ranges
 * (xcd_mapping).  None of them changes any result, bit for bit. */
int svm_set_tuning(svm_context* ctx, int32_t cluster_blocks, int32_t graph_steps, int32_t xcd_mapping);
int svm_get_tuning(svm_context* ctx, int32_t* cluster_blocks, int32_t* graph_steps, int32_t* xcd_mapping);
/* Kernel form of the vertex pass of Brownian and active steps in the FP64
 * parity arithmetic (no counterpart in the reference): 0 = a thread per
 * vertex evaluating its three cells (k_vertex, 212 VGPRs: two wavefronts
 * per SIMD); 1 and 2 = corner form, three wavefronts per 64 vertices,
 * wavefront k evaluating the vertex's neighbour slot k and then its cell slot
 * k, the contributions added in slot order (k_vertex_corner), compiled for
 * four (1) or five (2) wavefronts per SIMD; 3 = form 0 compiled for three
 * wavefronts per SIMD (168 VGPRs, the rest spilled); 4 = form 0 with each
 * workgroup's inputs staged in LDS (a window of 512 cell rows from the least
 * cell of its 256 vertices and its vertex rows with 64 on either side, loaded
 * once with coalesced loads; references outside the windows read the
 * arrays).  All give identical results, bit for bit; they differ in register
 * use, in how many wavefronts a SIMD keeps in flight to hide memory latency,
 * and in how the inputs are gathered. */
int svm_set_vertex_form(svm_context* ctx, int32_t form);
int svm_get_vertex_form(svm_context* ctx, int32_t* form);
/* Kernel form of the geometry pass (computeGeometryCPU,
 * src/model/sphericalVertexModel.cpp:209-311; no counterpart in the
 * reference for the form itself): each 256-row cell tile is one workgroup
 * whose threads take 1 (form 0), 2 (form 1) or 4 (form 2) cells each, i.e.
 * 256, 128 or 64 threads.  A thread walks its cells' rings one after the
 * other, so a wavefront waits on the longest sum of its lanes' ring lengths
 * rather than on the longest single ring.  The tile's area sum keeps the
 * 256-row halving tree.  All forms give identical results, bit for bit. */
int svm_set_geometry_form(svm_context* ctx, int32_t form);
int svm_get_geometry_form(svm_context* ctx, int32_t* form);
/* Launch structure of the Brownian, active and Vicsek steps (no counterpart
 * in the reference, whose step is a fixed sequence of kernels and host
 * synchronisations, src/model/sphericalVertexModel.cpp:795-892): a sum of
 * flags.
 *   1: the total-area check of the geometry and its triangle-fan fallback
 *      (computeGeometryCPU :304-310, recomputeAreasCPU :162-207) run inside
 *      the vertex pass, which folds each sphere's area partials itself, instead
 *      of a launch of their own.  Applies with the thread-per-vertex form
 *      (vertex form 0, or either FP32/vector precision) to lanes whose spheres
 *      have at most 65,536 cells.
 *   2: the T1 step closes without the group kernel: the check of blocked
 *      candidates against the dry runs runs in the commit kernel, and a sphere
 *      whose dry runs overlap takes the sequential pass instead of a group run
 *      (counted in the sequential column of svm_get_counters, with no reason
 *      column).  Applies to lanes whose spheres have fewer than 65,536
 *      vertices; also to FIRE, Adam and warm-start T1 passes.
 * Every form gives identical results, bit for bit; form 3 is 5 launches per
 * lane per step against 7 for form 0. */
int svm_set_step_form(svm_context* ctx, int32_t form);
int svm_get_step_form(svm_context* ctx, int32_t* form);
/* Lanes of the dynamics step (no counterpart in the reference, which steps
 * one sphere at a time): the batch is cut into n_lanes contiguous runs of
 * spheres with about equal numbers of vertices (at most SVM_MAX_LANES, at
 * most one lane per sphere), and the steps of svm_brownian_steps,
 * svm_active_steps, svm_vicsek_steps and svm_run_graph enqueue each lane's
 * launches on a stream of its own, forked from the context's stream and
 * joined back into it before those calls return.  One lane's T1 dry runs
 * (long dependent chains on a few compute units) then overlap the other
 * lanes' force passes.  Spheres are independent, so results are identical
 * bit for bit for every n_lanes.  Each lane above the first holds its own
 * streams and T1 candidate lists (about 65 MB of device memory); those are
 * kept until svm_destroy.  svm_get_lanes returns the number of lanes and,
 * when first_sphere is not null, the first sphere of each lane. */
#define SVM_MAX_LANES 4
int svm_set_lanes(svm_context* ctx, int32_t n_lanes);
int svm_get_lanes(svm_context* ctx, int32_t* n_lanes, int32_t* first_sphere);
/* svm_set_lanes with the sphere ranges given: lane k runs spheres
 * [first_sphere[k], first_sphere[k+1]) (the last lane up to the sphere
 * count).  first_sphere[0] must be 0 and the starts must increase strictly;
 * SVM_ERR_ARG otherwise.  For callers that weigh spheres by more than their
 * size (T1 activity differs by an order of magnitude between spheres of one
 * batch).  No counterpart in the reference, which steps one sphere at a
 * time. */
int svm_set_lane_spheres(svm_context* ctx, int32_t n_lanes, const int32_t* first_sphere);
/* Project autotuner, in the manner of kernelTuner's median-of-samples scan
 * (src/utilities/kernelTuner.cpp:35-126) but over the whole step: times
 * real steps of `updater` (0 Brownian, 1 active) under each candidate
 * schedule -- first lanes in {1 .. min(SVM_MAX_LANES, spheres)} x
 * cluster_blocks in {256, 1024} with vertex, step and geometry forms 0; at
 * the best lane count the step forms 1, 2 and 3; then the geometry forms 1
 * and 2; then the vertex forms 1 to 4 (FP64 parity forces only); then the
 * rest of cluster_blocks in {128, 256, 512, 1024} x xcd_mapping in {0, 1},
 * all at 10 steps per graph; then graph_steps in {4, 20, 40} -- each stage
 * keeping the best of the stages before -- keeps the fastest and reports
 * every trial in table (rows of SVM_TUNE_COLS doubles: lanes, cluster_blocks,
 * graph_steps, xcd_mapping, vertex_form, step_form, geometry_form, ms per
 * step).  The steps advance the model (about steps_per_trial per candidate),
 * exactly as untuned steps would. */
#define SVM_TUNE_COLS 8
int svm_autotune(svm_context* ctx, int32_t updater, int32_t steps_per_trial, double* table, int32_t table_rows,
                 int32_t* rows_written);
/* Persisted schedules (no counterpart in the reference, whose kernelTuner
 * keeps its choice in memory for one run): a text file with one line per
 * key, "svm-schedule-v2 updater precision sphere_bucket size_bucket lanes
 * cluster_blocks graph_steps xcd_mapping vertex_form step_form geometry_form
 * ms_per_step" (a "svm-schedule-v1" line, without the step and geometry
 * forms, reads them as 0), where the
 * key is the updater, the force precision, floor(log2(number of spheres)) and
 * floor(log2(mean cells per sphere)) of the context's batch.
 * svm_tuning_save writes the context's current schedule (and the ms per step
 * its last svm_autotune measured) under the batch's key, replacing an older
 * line with that key.  svm_tuning_load applies the schedule stored under the
 * batch's key, if any (found = 1), before any graph is captured with it; a
 * missing file or key leaves the schedule unchanged (found = 0, SVM_OK). */
int svm_tuning_save(svm_context* ctx, int32_t updater, const char* path);
int svm_tuning_load(svm_context* ctx, int32_t updater, const char* path, int32_t* found);

/* Kernel timing without launch gaps (no counterpart in the reference): the
 * fused vertex pass of `updater` (0 Brownian, 1 active; force + update +
 * move, k_vertex) launched `reps` times back to back between two HIP events
 * on the context's stream.  The pass is a pure function of the current state
 * (it writes the spare position buffer, which the next step overwrites), so
 * the model state is unchanged; its counters are restored. */
int svm_time_vertex_pass(svm_context* ctx, int32_t updater, int32_t reps, double* ms_per_launch);

/* Streaming-copy yardstick (no counterpart in the reference): allocates two
 * n_bytes buffers on the current device, copies one into the other `reps`
 * times with a plain HIP kernel (k_copy_probe) and returns the mean time per
 * copy.  Used to calibrate profiler byte counters (FETCH_SIZE / WRITE_SIZE)
 * and as the achievable-bandwidth reference. */
int svm_copy_probe(int64_t n_bytes, int32_t reps, double* ms_per_copy);
/* Phase-boundary yardstick (no counterpart in the reference): what a
 * dependency between the phases of a step costs as a kernel boundary in a
 * HIP graph (the library's form) and as a grid barrier inside one persistent
 * launch.  `workgroups` workgroups of 256 threads (at most what the device
 * keeps resident at once) run `phases` phases; in each, every workgroup
 * writes bytes_per_phase / workgroups bytes and reads what the next
 * workgroup wrote in the phase before.  Form 1: one launch per phase,
 * captured in one graph.  Form 2: one cooperative launch whose phases are
 * separated by a device-scope release, an arrival counter, a bounded wait
 * and an acquire.  Returns the microseconds per phase of each and the
 * number of values either form read wrong (0 when both boundaries order the
 * data).  A wait that runs out returns SVM_ERR_HIP. */
int svm_phase_boundary_probe(int32_t workgroups, int32_t phases, int64_t bytes_per_phase, double* us_per_launch,
                             double* us_per_barrier, int64_t* errors);

/* Self-check of the vertex pass's short root and quotient sequences (no
 * counterpart in the reference; the reference uses the operators): n random
 * samples under `seed` of quotients a / b (divisors across and beyond
 * [2^-384, 2^500), quotients across and beyond [2^-400, 2^400), zero
 * numerators, all-ones and all-zero mantissas, all-ones divisors over powers
 * of two) and of roots (arguments across [2^-800, 2^1024), negative, zero,
 * infinite, NaN and subnormal ones, all-ones mantissas), each evaluated by
 * the sequences and by the operators.  counts[8]: quotient samples in the
 * sequences' precondition, accepted quotients, accepted quotients that
 * differ from a / b in any bit; root samples, accepted roots, accepted roots
 * that differ from sqrt(x); accepted quotients and accepted roots that are
 * not the double nearest to the exact value (checked with exact fma
 * remainders). */
int svm_fast_ops_check(int64_t n, uint64_t seed, int64_t* counts);

#ifdef __cplusplus
}
#endif
#endif /* SVM_H */
