// svm_capi.hip -- host side of the C ABI declared in include/svm.h.
//
// A context owns every device buffer of a batch of spheres (layout in
// svm_device.h) and one HIP stream.  Host arrays cross the boundary in the
// reference's layouts and are converted here; nothing on the per-step path
// touches the host: step loops only enqueue kernels (or replay a captured
// graph), and errors raised on the device are collected at the next
// synchronising call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

// Grid size for kernels that map workgroups to tiles with xcd_tile().
static inline int xcd_grid(int ntiles) { return (ntiles + 7) / 8 * 8; }


#include "../../include/svm.h"
#include "svm_kernels.h"
#include "svm_delaunay.h"
#include "svm_voronoi.h"
#include "svm_neighbors.h"
#include "svm_mlp.h"
#include "svm_netcdf.h"

using svm::Dev;
using svm::SysParams;

struct svm_context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t own_stream = nullptr;
    int nsys = 0;
    std::vector<svm_system_desc> desc;
    std::vector<SysParams> layout;   // sizes/offsets (host copy of the constant part)
    std::vector<int> have_topology;
    std::vector<int> have_adam;       // svm_adam_set_parameters called for the sphere
    std::vector<int> voronoi;         // sphere built by svm_build_voronoi_from_cell_points
    int Cs = 0, Vs = 0, cap = 0, nct = 0, nvt = 0;
    int cluster_blocks = 256;        // T1 dry-run grid (persistent loop over candidates)
    int graph_steps = 10;            // steps per captured graph when the caller passes 0
    int t1_sequential = 0;           // 1: skip the dry runs, always use the sequential pass
    double tuned_ms = 0.0;           // ms per step of the schedule the autotuner kept (0: not tuned)
    int vertex_corner = 0;           // Brownian/active vertex pass: 0 k_vertex; 1, 2 corner form at 4, 5 waves per SIMD
    int step_form = 0;               // svm_set_step_form bits: 1 fan check in the vertex pass, 2 no group kernel
    int geometry_form = 0;           // k_geometry: 1 (ring prefetched), 2 or 4 cells per thread (forms 0, 1, 2)
    Dev d{};
    int pcur = 0, ncur = 0;          // current position / director buffer
    std::string err;
    std::vector<void*> allocs;
    // captured step graphs keyed by (updater, steps)
    std::map<std::pair<int, int>, hipGraphExec_t> graphs;
    std::map<std::pair<int, int>, hipGraph_t> graph_defs;
    // Delaunay triangles per sphere (from svm_build_from_cell_points)
    std::vector<std::vector<int32_t>> triangles;
    // force arithmetic of every force evaluation: 0 FP64 reference expressions
    // (parity), 1 FP32 vector form, 2 FP64 vector form
    int force_fp32 = 0;
    // Lanes of the dynamics step (svm_set_lanes): contiguous sphere ranges
    // whose steps are enqueued on streams of their own, forked from `stream`
    // and joined back into it.  Lane 0 runs on `stream`; lane k > 0 on
    // lane_streams[k-1] with its own T1 candidate lists lane_t1[k-1].
    int nlanes = 1;
    std::vector<int> lane_s0;                 // first sphere of each lane, then nsys
    std::vector<hipStream_t> lane_streams;
    std::vector<hipEvent_t> lane_joins;
    hipEvent_t lane_fork = nullptr;
    std::vector<Dev> lane_t1;                 // only the t1_* buffers are used
    std::map<std::pair<int, int>, std::vector<hipGraphExec_t>> lane_graphs;
    std::map<std::pair<int, int>, std::vector<hipGraph_t>> lane_graph_defs;
    // profiling
    bool profiling = false;
    std::map<std::string, std::pair<double, long>> timing;
    std::vector<std::string> timing_names;
};

namespace {

const char* k_ok = "ok";

int fail(svm_context* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

#define HIPCHECK(ctx, call)                                                                      \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return fail(ctx, SVM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));    \
    } while (0)

template <typename T>
int dalloc(svm_context* ctx, T** p, size_t n) {
    void* q = nullptr;
    size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return fail(ctx, SVM_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    // the fill is complete before the buffer is handed out, so no later work
    // on any stream can be overwritten by the zeros (buffers allocated
    // lazily, mid-run, would otherwise race).  The library's own stream, not
    // the null stream: the library leaves every hardware queue of the
    // process (a handful) to its own streams and the lanes'.
    e = hipMemsetAsync(q, 0, bytes, ctx->own_stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->own_stream);
    if (e != hipSuccess) return fail(ctx, SVM_ERR_HIP, std::string("hipMemset: ") + hipGetErrorString(e));
    ctx->allocs.push_back(q);
    *p = (T*)q;
    return SVM_OK;
}

int round_up(int x, int m) { return (x + m - 1) / m * m; }

bool valid_sys(svm_context* ctx, int sys) { return ctx && sys >= 0 && sys < ctx->nsys; }

// Field update of one sphere's device parameter block.
template <typename T>
int set_field(svm_context* ctx, int sys, size_t offset, T value) {
    char* base = (char*)(ctx->d.sys + sys) + offset;
    HIPCHECK(ctx, hipMemcpyAsync(base, &value, sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    return SVM_OK;
}

int download_sys(svm_context* ctx, std::vector<SysParams>& out) {
    out.resize(ctx->nsys);
    HIPCHECK(ctx, hipMemcpyAsync(out.data(), ctx->d.sys, sizeof(SysParams) * ctx->nsys, hipMemcpyDeviceToHost,
                                 ctx->stream));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    return SVM_OK;
}

int check_device_errors(svm_context* ctx) {
    std::vector<SysParams> s;
    int r = download_sys(ctx, s);
    if (r) return r;
    for (int i = 0; i < ctx->nsys; ++i) {
        if (s[i].err == SVM_ERR_RING_OVERFLOW)
            return fail(ctx, SVM_ERR_RING_OVERFLOW, "sphere " + std::to_string(i) + ": a T1 flip would exceed ring_capacity");
        if (s[i].err == SVM_ERR_TOPOLOGY)
            return fail(ctx, SVM_ERR_TOPOLOGY, "sphere " + std::to_string(i) + ": rings and vertex lists disagree");
        if (s[i].err) return fail(ctx, s[i].err, "sphere " + std::to_string(i) + ": device error " + std::to_string(s[i].err));
    }
    return SVM_OK;
}

// Scatter a host AoS [3*nv] array into the SoA rows of sphere `sys`.
int upload_vec3(svm_context* ctx, int sys, const double* src, double* X, double* Y, double* Z) {
    const SysParams& L = ctx->layout[sys];
    std::vector<double> t(3 * (size_t)L.nv);
    for (int i = 0; i < L.nv; ++i) {
        t[i] = src[3 * (size_t)i];
        t[L.nv + i] = src[3 * (size_t)i + 1];
        t[2 * (size_t)L.nv + i] = src[3 * (size_t)i + 2];
    }
    HIPCHECK(ctx, hipMemcpyAsync(X + L.vo, t.data(), sizeof(double) * L.nv, hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(ctx, hipMemcpyAsync(Y + L.vo, t.data() + L.nv, sizeof(double) * L.nv, hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(ctx, hipMemcpyAsync(Z + L.vo, t.data() + 2 * (size_t)L.nv, sizeof(double) * L.nv, hipMemcpyHostToDevice,
                                 ctx->stream));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    return SVM_OK;
}

int download_rows(svm_context* ctx, int n, int off, const double* X, const double* Y, const double* Z, double* dst) {
    std::vector<double> t(3 * (size_t)n);
    HIPCHECK(ctx, hipMemcpyAsync(t.data(), X + off, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(ctx, hipMemcpyAsync(t.data() + n, Y + off, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(ctx, hipMemcpyAsync(t.data() + 2 * (size_t)n, Z + off, sizeof(double) * n, hipMemcpyDeviceToHost,
                                 ctx->stream));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < n; ++i) {
        dst[3 * (size_t)i] = t[i];
        dst[3 * (size_t)i + 1] = t[n + i];
        dst[3 * (size_t)i + 2] = t[2 * (size_t)n + i];
    }
    return SVM_OK;
}

// Kernel launch with optional event timing (profiling mode, eager only).
struct Timer {
    svm_context* ctx;
    const char* name;
    hipEvent_t a = nullptr, b = nullptr;
    Timer(svm_context* c, const char* n) : ctx(c), name(n) {
        if (ctx->profiling) {
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            (void)hipEventRecord(a, ctx->stream);
        }
    }
    ~Timer() {
        if (ctx->profiling) {
            (void)hipEventRecord(b, ctx->stream);
            (void)hipEventSynchronize(b);
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, a, b);
            auto& e = ctx->timing[name];
            e.first += ms;
            e.second += 1;
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    }
};

// Captured graphs hold kernel arguments and launch shapes: drop them when a
// setting they baked in changes.
void drop_graphs(svm_context* ctx) {
    for (auto& g : ctx->graphs) (void)hipGraphExecDestroy(g.second);
    for (auto& g : ctx->graph_defs) (void)hipGraphDestroy(g.second);
    for (auto& g : ctx->lane_graphs)
        for (auto e : g.second) (void)hipGraphExecDestroy(e);
    for (auto& g : ctx->lane_graph_defs)
        for (auto e : g.second) (void)hipGraphDestroy(e);
    ctx->graphs.clear();
    ctx->graph_defs.clear();
    ctx->lane_graphs.clear();
    ctx->lane_graph_defs.clear();
}

int last_launch(svm_context* ctx) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(ctx, SVM_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return SVM_OK;
}

// ---- step building blocks ----
// What a launch covers and where it is enqueued: the whole batch on the
// context's stream, or one lane of the dynamics step on the lane's stream.
struct Target {
    Dev d;
    hipStream_t st;
};
Target whole(svm_context* ctx) { return Target{ctx->d, ctx->stream}; }

// with_fan = false leaves the area check and the fan fallback to the vertex
// pass that follows (fan_check of k_vertex).
int launch_geometry(svm_context* ctx, const Target& w, int skip_frozen, bool with_fan = true) {
    {
        Timer t(ctx, "k_geometry");
        const dim3 grid(xcd_grid(w.d.ctn));
        if (ctx->geometry_form == 2)
            hipLaunchKernelGGL(svm::k_geometry<4>, grid, dim3(SVM_TILE / 4), 0, w.st, w.d, ctx->pcur, skip_frozen);
        else if (ctx->geometry_form == 1)
            hipLaunchKernelGGL(svm::k_geometry<2>, grid, dim3(SVM_TILE / 2), 0, w.st, w.d, ctx->pcur, skip_frozen);
        else
            hipLaunchKernelGGL((svm::k_geometry<1, SVM_GEO_PF>), grid, dim3(SVM_TILE), 0, w.st, w.d, ctx->pcur,
                               skip_frozen);
        int r = last_launch(ctx);
        if (r) return r;
    }
    if (!with_fan) return SVM_OK;
    Timer t(ctx, "k_fan_fix");
    hipLaunchKernelGGL(svm::k_fan_fix, dim3(xcd_grid(w.d.ctn)), dim3(SVM_TILE), 0, w.st, w.d, ctx->pcur,
                       skip_frozen);
    return last_launch(ctx);
}
int launch_geometry(svm_context* ctx, int skip_frozen) { return launch_geometry(ctx, whole(ctx), skip_frozen); }

// The active updater also writes the rotated directors, into the other
// director buffer (1 - ncur); no other mode touches directors.
// FAN: the vertex pass also does the geometry's area check and fan fallback
// (step form flag 1; Brownian and active passes only).
template <int MODE, bool FAN>
int launch_vertex_as(svm_context* ctx, const Target& w, int pout, int store_forces) {
    const dim3 grid(xcd_grid(w.d.vtn));
    const int nout = 1 - ctx->ncur;
    if (ctx->force_fp32 == 1)
        hipLaunchKernelGGL((svm::k_vertex<MODE, float, FAN>), grid, dim3(SVM_TILE), 0, w.st, w.d, ctx->pcur, pout,
                           ctx->ncur, nout, store_forces);
    else if (ctx->force_fp32 == 2)
        hipLaunchKernelGGL((svm::k_vertex<MODE, svm::VecF64, FAN>), grid, dim3(SVM_TILE), 0, w.st, w.d, ctx->pcur,
                           pout, ctx->ncur, nout, store_forces);
    else if ((MODE == svm::VM_BROWNIAN || MODE == svm::VM_ACTIVE) && ctx->vertex_corner == 4 && !store_forces) {
        if (FAN) return fail(ctx, SVM_ERR_ARG, "the staged vertex form has no fan check");
        constexpr int M = MODE == svm::VM_ACTIVE ? svm::VM_ACTIVE : svm::VM_BROWNIAN;
        hipLaunchKernelGGL((svm::k_vertex_staged<M>), grid, dim3(SVM_TILE), 0, w.st, w.d, ctx->pcur, pout,
                           ctx->ncur, nout);
    }
    else if ((MODE == svm::VM_BROWNIAN || MODE == svm::VM_ACTIVE) && ctx->vertex_corner == 3 && !store_forces) {
        constexpr int M = MODE == svm::VM_ACTIVE ? svm::VM_ACTIVE : svm::VM_BROWNIAN;
        hipLaunchKernelGGL((svm::k_vertex_w3<M, FAN>), grid, dim3(SVM_TILE), 0, w.st, w.d, ctx->pcur, pout,
                           ctx->ncur, nout);
    }
    else if ((MODE == svm::VM_BROWNIAN || MODE == svm::VM_ACTIVE) && ctx->vertex_corner && !store_forces) {
        if (FAN) return fail(ctx, SVM_ERR_ARG, "the corner vertex form has no fan check");
        constexpr int M = MODE == svm::VM_ACTIVE ? svm::VM_ACTIVE : svm::VM_BROWNIAN;
        const dim3 cgrid(xcd_grid(w.d.vtn * (SVM_TILE / SVM_CORNER_V)));
        if (ctx->vertex_corner == 2)
            hipLaunchKernelGGL((svm::k_vertex_corner<M, 5>), cgrid, dim3(3 * SVM_CORNER_V), 0, w.st, w.d, ctx->pcur,
                               pout, ctx->ncur, nout);
        else
            hipLaunchKernelGGL((svm::k_vertex_corner<M, 4>), cgrid, dim3(3 * SVM_CORNER_V), 0, w.st, w.d, ctx->pcur,
                               pout, ctx->ncur, nout);
    }
    else
        hipLaunchKernelGGL((svm::k_vertex<MODE, double, FAN>), grid, dim3(SVM_TILE), 0, w.st, w.d, ctx->pcur, pout,
                           ctx->ncur, nout, store_forces);
    return last_launch(ctx);
}

template <int MODE>
int launch_vertex(svm_context* ctx, const Target& w, int pout, int store_forces, const char* name,
                  bool fan_check = false) {
    Timer t(ctx, name);
    if constexpr (MODE == svm::VM_BROWNIAN || MODE == svm::VM_ACTIVE)
        if (fan_check) return launch_vertex_as<MODE, true>(ctx, w, pout, store_forces);
    if (fan_check) return fail(ctx, SVM_ERR_ARG, "fan check only in Brownian and active passes");
    return launch_vertex_as<MODE, false>(ctx, w, pout, store_forces);
}
template <int MODE>
int launch_vertex(svm_context* ctx, int pout, int store_forces, const char* name) {
    return launch_vertex<MODE>(ctx, whole(ctx), pout, store_forces, name);
}

// detect_skip: spheres the detection leaves out (k_t1_detect's `skip`; -1:
// as fire_mode, i.e. the spheres FIRE has stopped when in FIRE mode).
int launch_t1(svm_context* ctx, const Target& w, int pcur, int nin, int nout, int dir, int advance_step,
              int fire_mode, int detect_skip = -1) {
    {
        Timer t(ctx, "k_t1_detect");
        hipLaunchKernelGGL(svm::k_t1_detect, dim3(xcd_grid(w.d.vtn)), dim3(SVM_TILE), 0, w.st, w.d, pcur,
                           detect_skip < 0 ? fire_mode : detect_skip);
        int r = last_launch(ctx);
        if (r) return r;
    }
    if (!ctx->t1_sequential) {
        Timer t(ctx, "k_t1_cluster");
        hipLaunchKernelGGL(svm::k_t1_cluster, dim3(ctx->cluster_blocks), dim3(64 * SVM_T1_WAVES), 0, w.st, w.d,
                           pcur);
        int r = last_launch(ctx);
        if (r) return r;
    }
    bool large = false;
    for (int s = w.d.s0; s < w.d.s0 + w.d.ns; ++s) large |= ctx->layout[s].nv >= SVM_COMMIT_LARGE;
    // step form flag 2: the blocked-candidate check in k_t1_resolve instead of
    // a k_t1_group launch (not with the large-sphere commit, which decides
    // before k_t1_resolve)
    const int check_blocked = (ctx->step_form & 2) && !large && !ctx->t1_sequential;
    if (!ctx->t1_sequential && !check_blocked) {
        Timer t(ctx, "k_t1_group");
        hipLaunchKernelGGL(svm::k_t1_group, dim3(w.d.ns), dim3(SVM_T1_GT), 0, w.st, w.d, pcur);
        int r = last_launch(ctx);
        if (r) return r;
    }
    if (large && !ctx->t1_sequential) {
        Timer t(ctx, "k_t1_commit_large");
        hipLaunchKernelGGL(svm::k_t1_commit_large, dim3(w.d.ns), dim3(1024), 0, w.st, w.d, pcur, nin, nout, dir,
                           fire_mode, ctx->t1_sequential);
        int r = last_launch(ctx);
        if (r) return r;
    }
    Timer t(ctx, "k_t1_resolve");
    hipLaunchKernelGGL(svm::k_t1_resolve, dim3(w.d.ns), dim3(SVM_TILE), 0, w.st, w.d, pcur, nin, nout, dir,
                       advance_step, fire_mode, ctx->t1_sequential, check_blocked);
    return last_launch(ctx);
}
int launch_t1(svm_context* ctx, int pcur, int nin, int nout, int dir, int advance_step, int fire_mode,
              int detect_skip = -1) {
    return launch_t1(ctx, whole(ctx), pcur, nin, nout, dir, advance_step, fire_mode, detect_skip);
}

// One dynamics step of the target's spheres: geometry, force+updater+move,
// T1.  updater 0 = Brownian, 1 = active, 2 = Vicsek.  The ping-pong buffer
// indices advance in ctx->pcur / ctx->ncur.
int enqueue_step(svm_context* ctx, const Target& w, int updater) {
    // step form flag 1: the area check and fan fallback in the vertex pass (the
    // thread-per-vertex form, spheres whose tile partials one wavefront folds)
    bool fold = (ctx->step_form & 1) && (ctx->vertex_corner == 0 || ctx->vertex_corner == 3 || ctx->force_fp32 != 0);
    for (int s = w.d.s0; s < w.d.s0 + w.d.ns && fold; ++s) fold = ctx->layout[s].nctiles <= 64 * SVM_FOLD_K;
    int r = launch_geometry(ctx, w, 0, !fold);
    if (r) return r;
    int pout = 1 - ctx->pcur;
    if (updater == 0) r = launch_vertex<svm::VM_BROWNIAN>(ctx, w, pout, 0, "k_vertex_brownian", fold);
    else r = launch_vertex<svm::VM_ACTIVE>(ctx, w, pout, 0, "k_vertex_active", fold);
    if (r) return r;
    if (updater == 2) {
        // Vicsek: the same move as the active updater, T1, then the alignment
        // update of the directors on the new positions and topology
        r = launch_t1(ctx, w, pout, ctx->ncur, ctx->ncur, 0, 1, 0);
        if (r) return r;
        ctx->pcur = pout;
        const int nout = 1 - ctx->ncur;
        Timer t(ctx, "k_vicsek_dir");
        hipLaunchKernelGGL(svm::k_vicsek_dir, dim3(xcd_grid(w.d.vtn)), dim3(SVM_TILE), 0, w.st, w.d, ctx->pcur,
                           ctx->ncur, nout);
        r = last_launch(ctx);
        if (r) return r;
        ctx->ncur = nout;
        return SVM_OK;
    }
    int nout = updater == 1 ? 1 - ctx->ncur : ctx->ncur;
    r = launch_t1(ctx, w, pout, ctx->ncur, nout, updater == 1, 1, 0);
    if (r) return r;
    ctx->pcur = pout;
    ctx->ncur = nout;
    return SVM_OK;
}

// Lane k of the batch: its spheres and tiles, its T1 candidate lists, its stream.
Target lane_target(svm_context* ctx, int k) {
    Dev d = ctx->d;
    const int a = ctx->lane_s0[k], b = ctx->lane_s0[k + 1];
    const SysParams& A = ctx->layout[a];
    const SysParams& B = ctx->layout[b - 1];
    d.s0 = a;
    d.ns = b - a;
    d.vt0 = A.vtile0;
    d.vtn = B.vtile0 + B.nvtiles - A.vtile0;
    d.ct0 = A.ctile0;
    d.ctn = B.ctile0 + B.nctiles - A.ctile0;
    if (k == 0) return Target{d, ctx->stream};
    const Dev& L = ctx->lane_t1[k - 1];
    d.t1_count = L.t1_count;
    d.t1_list = L.t1_list;
    d.t1_flagged = L.t1_flagged;
    d.t1_logged = L.t1_logged;
    d.t1_lead = L.t1_lead;
    d.t1_total = L.t1_total;
    d.t1_btotal = L.t1_btotal;
    d.t1_bcount = L.t1_bcount;
    d.t1_blist = L.t1_blist;
    d.t1_bflag = L.t1_bflag;
    d.t1_bskip = L.t1_bskip;
    d.t1_log = L.t1_log;
    return Target{d, ctx->lane_streams[k - 1]};
}

// Runs `body` (which enqueues work of every lane on the lanes' streams)
// between a fork from the context's stream and a join back into it.  The
// join is enqueued even when `body` fails, so that later work on the
// context's stream stays ordered after whatever the lanes received.
template <class F>
int over_lanes(svm_context* ctx, F body) {
    HIPCHECK(ctx, hipEventRecord(ctx->lane_fork, ctx->stream));
    for (int k = 1; k < ctx->nlanes; ++k) HIPCHECK(ctx, hipStreamWaitEvent(ctx->lane_streams[k - 1], ctx->lane_fork, 0));
    int r = body();
    for (int k = 1; k < ctx->nlanes; ++k) {
        if (hipEventRecord(ctx->lane_joins[k - 1], ctx->lane_streams[k - 1]) != hipSuccess ||
            hipStreamWaitEvent(ctx->stream, ctx->lane_joins[k - 1], 0) != hipSuccess)
            r = r ? r : fail(ctx, SVM_ERR_HIP, "lane join failed");
    }
    return r;
}

// n iterations of `enqueue` (one step or minimiser iteration of a target)
// for every lane, interleaved iteration by iteration across the lanes'
// streams.  All lanes share the ping-pong buffer indices, which advance once
// per iteration.
template <class F>
int enqueue_lane_iterations(svm_context* ctx, int n, F enqueue) {
    return over_lanes(ctx, [&]() -> int {
        for (int i = 0; i < n; ++i) {
            const int p = ctx->pcur, q = ctx->ncur;
            int np = p, nq = q;
            for (int k = 0; k < ctx->nlanes; ++k) {
                ctx->pcur = p;
                ctx->ncur = q;
                int r = enqueue(lane_target(ctx, k));
                if (r) {
                    ctx->pcur = p;
                    ctx->ncur = q;
                    return r;
                }
                np = ctx->pcur;
                nq = ctx->ncur;
            }
            ctx->pcur = np;
            ctx->ncur = nq;
        }
        return SVM_OK;
    });
}

// One FIRE iteration: mixing + VV first half + move, T1, geometry, force +
// VV second half + reductions + controller.
int enqueue_fire_iteration(svm_context* ctx, const Target& w) {
    int pout = 1 - ctx->pcur;
    {
        Timer t(ctx, "k_fire_vv1");
        hipLaunchKernelGGL(svm::k_fire_vv1, dim3(xcd_grid(w.d.vtn)), dim3(SVM_TILE), 0, w.st, w.d, ctx->pcur, pout);
        int r = last_launch(ctx);
        if (r) return r;
    }
    int r = launch_t1(ctx, w, pout, ctx->ncur, ctx->ncur, 0, 0, 1);
    if (r) return r;
    ctx->pcur = pout;
    r = launch_geometry(ctx, w, 1);
    if (r) return r;
    r = launch_vertex<svm::VM_FIRE>(ctx, w, ctx->pcur, 1, "k_vertex_fire");
    if (r) return r;
    Timer t(ctx, "k_fire_control");
    hipLaunchKernelGGL(svm::k_fire_control, dim3(w.d.ns), dim3(SVM_TILE), 0, w.st, w.d);
    return last_launch(ctx);
}

// One Adam iteration: geometry, forces + moment update + move + reductions,
// T1, then the per-sphere controller.
int enqueue_adam_iteration(svm_context* ctx, const Target& w) {
    int r = launch_geometry(ctx, w, 1);
    if (r) return r;
    const int pout = 1 - ctx->pcur;
    r = launch_vertex<svm::VM_ADAM>(ctx, w, pout, 0, "k_vertex_adam");
    if (r) return r;
    r = launch_t1(ctx, w, pout, ctx->ncur, ctx->ncur, 0, 0, 1);
    if (r) return r;
    ctx->pcur = pout;
    Timer t(ctx, "k_adam_control");
    hipLaunchKernelGGL(svm::k_adam_control, dim3(w.d.ns), dim3(SVM_TILE), 0, w.st, w.d);
    return last_launch(ctx);
}

// `replays` x `k` iterations of `enqueue(target)` -- dynamics steps (kind =
// updater 0-2), FIRE (3) or Adam (4) iterations -- over the whole batch, or
// over every lane when the context has lanes.  With `graph` and k even they
// are replayed from captured HIP graphs: the ping-pong buffers then end where
// they started, so a graph captured from buffer state (pcur, ncur) replays
// whenever the context is in that state again.  Otherwise, or when
// profiling, they run eagerly.  Same kernels in the same order either way.
template <class F>
int iterate(svm_context* ctx, int kind, int k, int replays, bool graph, F enqueue) {
    const bool lanes = ctx->nlanes > 1 && !ctx->profiling;   // per-kernel timing is on the context's stream
    if (!graph || (k & 1) || ctx->profiling) {
        for (int rep = 0; rep < replays; ++rep) {
            if (lanes) {
                int r = enqueue_lane_iterations(ctx, k, enqueue);
                if (r) return r;
                continue;
            }
            for (int i = 0; i < k; ++i) {
                int r = enqueue(whole(ctx));
                if (r) return r;
            }
        }
        return SVM_OK;
    }
    auto key = std::make_pair(kind, k * 4 + ctx->pcur * 2 + ctx->ncur);
    const int nl = lanes ? ctx->nlanes : 1;
    std::vector<hipGraphExec_t>* execs = nullptr;
    hipGraphExec_t whole_exec = nullptr;
    if (lanes) {
        auto lt = ctx->lane_graphs.find(key);
        if (lt != ctx->lane_graphs.end()) execs = &lt->second;
    } else {
        auto it = ctx->graphs.find(key);
        if (it != ctx->graphs.end()) whole_exec = it->second;
    }
    if (!execs && !whole_exec) {
        // capture: one graph per lane, each on the lane's stream (the whole
        // batch is the one lane of the context's stream)
        const int p = ctx->pcur, q = ctx->ncur;
        std::vector<hipGraphExec_t> ex;
        std::vector<hipGraph_t> defs;
        auto discard = [&]() {
            for (auto e : ex) (void)hipGraphExecDestroy(e);
            for (auto g : defs) (void)hipGraphDestroy(g);
            ctx->pcur = p;
            ctx->ncur = q;
        };
        for (int l = 0; l < nl; ++l) {
            const Target w = lanes ? lane_target(ctx, l) : whole(ctx);
            ctx->pcur = p;
            ctx->ncur = q;
            if (hipStreamBeginCapture(w.st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
                discard();
                return fail(ctx, SVM_ERR_HIP, "graph capture failed");
            }
            int r = SVM_OK;
            for (int i = 0; i < k && !r; ++i) r = enqueue(w);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(w.st, &g);
            if (g) defs.push_back(g);
            hipGraphExec_t ge = nullptr;
            if (!r && (e != hipSuccess || hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess))
                r = fail(ctx, SVM_ERR_HIP, "graph capture failed");
            if (r) {
                discard();
                return r;
            }
            ex.push_back(ge);
        }
        ctx->pcur = p;
        ctx->ncur = q;
        if (lanes) {
            ctx->lane_graphs[key] = ex;
            ctx->lane_graph_defs[key] = defs;
            execs = &ctx->lane_graphs[key];
        } else {
            ctx->graphs[key] = ex[0];
            ctx->graph_defs[key] = defs[0];
            whole_exec = ex[0];
        }
    }
    if (!lanes) {
        for (int i = 0; i < replays; ++i) HIPCHECK(ctx, hipGraphLaunch(whole_exec, ctx->stream));
        return SVM_OK;
    }
    const std::vector<hipGraphExec_t>& ge = *execs;
    return over_lanes(ctx, [&]() -> int {
        for (int i = 0; i < replays; ++i)
            for (int l = 0; l < ctx->nlanes; ++l)
                HIPCHECK(ctx, hipGraphLaunch(ge[l], l == 0 ? ctx->stream : ctx->lane_streams[l - 1]));
        return SVM_OK;
    });
}

int require_topology(svm_context* ctx) {
    for (int s = 0; s < ctx->nsys; ++s)
        if (!ctx->have_topology[s]) return fail(ctx, SVM_ERR_STATE, "sphere " + std::to_string(s) + " has no topology");
    return SVM_OK;
}

}  // namespace

// ======================================================================== API
namespace {

// Device work buffers of one call.  By default each get() is a hipMalloc,
// kept until the object ends; rewind() starts the order again, and later gets
// reuse the buffers of the same position when they are large enough, so a
// loop that repeats the same sequence of gets allocates only on its first
// pass (or when a size grows).  In arena mode
// (constructed over caller-provided device memory) get() hands out 256-byte
// aligned ranges of that memory and never allocates; a request beyond its end
// returns nullptr and sets `overflow`.
struct DevScratch {
    std::vector<void*> ptrs;
    std::vector<size_t> sizes;
    size_t cursor = 0;
    char* arena = nullptr;
    size_t arena_bytes = 0;
    bool overflow = false;
    DevScratch() = default;
    DevScratch(void* base, size_t bytes) : arena((char*)base), arena_bytes(bytes) {}
    DevScratch(const DevScratch&) = delete;
    DevScratch& operator=(const DevScratch&) = delete;
    ~DevScratch() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    void rewind() { cursor = 0; }
    template <class T>
    T* get(size_t n) {
        const size_t bytes = sizeof(T) * std::max<size_t>(n, 1);
        if (arena) {
            const size_t at = (cursor + 255) & ~(size_t)255;
            if (at > arena_bytes || bytes > arena_bytes - at) {
                overflow = true;
                return nullptr;
            }
            cursor = at + bytes;
            return (T*)(arena + at);
        }
        if (cursor < ptrs.size() && sizes[cursor] >= bytes) return (T*)ptrs[cursor++];
        void* p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
        if (cursor < ptrs.size()) {
            (void)hipFree(ptrs[cursor]);   // too small: replaced (hipFree waits for the device)
            ptrs[cursor] = p;
            sizes[cursor] = bytes;
        } else {
            ptrs.push_back(p);
            sizes.push_back(bytes);
        }
        ++cursor;
        return (T*)p;
    }
};

// exclusive scan of n device ints into out (recursive over blocks of 1024)
bool dl_scan(hipStream_t st, DevScratch& m, const int* in, int n, int* out) {
    const int nb = (n + 1023) / 1024;
    int* sums = m.get<int>(nb);
    if (!sums) return false;
    hipLaunchKernelGGL(svm::k_scan_block, dim3(nb), dim3(1024), 0, st, in, n, out, sums);
    if (nb > 1) {
        int* offs = m.get<int>(nb);
        if (!offs || !dl_scan(st, m, sums, nb, offs)) return false;
        hipLaunchKernelGGL(svm::k_scan_add, dim3(nb), dim3(1024), 0, st, out, n, (const int*)offs);
    }
    return hipGetLastError() == hipSuccess;
}

// Cell list of n device points (AoS) in a uniform cube grid over
// [-Rs, Rs]^3 with bins of side about `h`: start[b] .. start[b+1] index the
// points of bin b in `list`, in index order.
struct DlCells {
    svm::DlGrid g;
    int nbins = 0;
    int* bin_of = nullptr;
    int* start = nullptr;
    int* list = nullptr;
};

bool dl_cell_list(hipStream_t st, DevScratch& m, const double* P, int n, double Rs, double h, DlCells& c) {
    c.g.lo = -Rs * (1.0 + 1e-9);
    c.g.G = std::max(1, std::min(256, (int)std::ceil(2.0 * Rs * (1.0 + 1e-9) / h)));
    c.g.h = 2.0 * Rs * (1.0 + 1e-9) / c.g.G;
    c.nbins = c.g.G * c.g.G * c.g.G;
    c.bin_of = m.get<int>(n);
    int* count = m.get<int>(c.nbins + 1);
    c.start = m.get<int>(c.nbins + 1);
    int* fill = m.get<int>(c.nbins);
    c.list = m.get<int>(n);
    if (!c.bin_of || !count || !c.start || !fill || !c.list) return false;
    const int tb = 256;
    hipLaunchKernelGGL(svm::k_fill_int, dim3((c.nbins + 1 + tb - 1) / tb), dim3(tb), 0, st, count, c.nbins + 1, 0);
    hipLaunchKernelGGL(svm::k_fill_int, dim3((c.nbins + tb - 1) / tb), dim3(tb), 0, st, fill, c.nbins, 0);
    hipLaunchKernelGGL(svm::k_dl_bin, dim3((n + tb - 1) / tb), dim3(tb), 0, st, P, n, c.g, c.bin_of, count);
    if (!dl_scan(st, m, count, c.nbins + 1, c.start)) return false;
    hipLaunchKernelGGL(svm::k_dl_fill, dim3((n + tb - 1) / tb), dim3(tb), 0, st, n, (const int*)c.bin_of,
                       (const int*)c.start, fill, c.list);
    hipLaunchKernelGGL(svm::k_dl_sort_bins, dim3((c.nbins + tb - 1) / tb), dim3(tb), 0, st, c.nbins,
                       (const int*)c.start, c.list);
    return hipGetLastError() == hipSuccess;
}

struct DlOut {
    double* P = nullptr;   // points, AoS
    int* T = nullptr;      // canonical triangles
    int* off = nullptr;    // first triangle of each point's slice
    int* nbr = nullptr;    // CCW Delaunay neighbours, cap per point
    int* deg = nullptr;
    int* tid = nullptr;    // triangle of each star wedge
    int* err = nullptr;
    int ntri = 0;
};

// Spherical Delaunay triangulation of n points (host, AoS) on stream `st`,
// exact (see svm_voronoi.h).  Returns SVM_OK or an error code with `msg`.
// Triangulation of n device points P (AoS) whose largest radius is at most
// Rs (used to size the cell list and the search radii; the result does not
// depend on it).
int dl_triangulate_dev(hipStream_t st, DevScratch& m, double* P, int n, double Rs, int cap, DlOut& o,
                       std::string& msg) {
    if (n < 4) { msg = "a spherical triangulation needs at least 4 points"; return SVM_ERR_ARG; }
    const double spacing = std::sqrt(4.0 * SVM_PI * Rs * Rs / n);   // mean distance between neighbours
    const double r0 = 3.5 * spacing;
    o.P = P;
    DlCells cl;
    if (!dl_cell_list(st, m, o.P, n, Rs, r0, cl)) { msg = "cell list failed"; return SVM_ERR_HIP; }
    const svm::DlGrid g = cl.g;
    const int* start = cl.start;
    const int* list = cl.list;
    o.nbr = m.get<int>((size_t)n * cap);
    o.deg = m.get<int>(n);
    int* status = m.get<int>(n);
    int* cnt = m.get<int>(n + 1);
    o.off = m.get<int>(n + 1);
    o.tid = m.get<int>((size_t)n * cap);
    o.err = m.get<int>(1);
    int* outcome = m.get<int>(2);
    double* rad = m.get<double>(n);    // per-star search radius and its bisection bounds
    double* rlo = m.get<double>(n);
    double* rhi = m.get<double>(n);
    if (!o.nbr || !o.deg || !status || !cnt || !o.off || !o.tid || !o.err || !outcome || !rad || !rlo || !rhi) {
        msg = "device allocation failed";
        return SVM_ERR_HIP;
    }
    const int tb = 256;
    auto blocks = [&](long long x) { return dim3((unsigned)((x + tb - 1) / tb)); };
    hipLaunchKernelGGL(svm::k_fill_int, dim3(1), dim3(tb), 0, st, o.err, 1, 0);
    hipLaunchKernelGGL(svm::k_fill_int, blocks(n + 1), dim3(tb), 0, st, cnt, n + 1, 0);
    hipLaunchKernelGGL(svm::k_fill_int, blocks(n), dim3(tb), 0, st, status, n, (int)svm::DL_PENDING);
    hipLaunchKernelGGL(svm::k_fill_double, blocks(n), dim3(tb), 0, st, rad, n, r0);
    hipLaunchKernelGGL(svm::k_fill_double, blocks(n), dim3(tb), 0, st, rlo, n, 0.0);
    hipLaunchKernelGGL(svm::k_fill_double, blocks(n), dim3(tb), 0, st, rhi, n, 0.0);
    // stars, each with its own search radius: r0 (about 40 candidates at
    // uniform density) first, then doubled for the stars that need a wider
    // search and bisected down for the crowded ones (k_dl_radius)
    const double whole = 2.0 * Rs * (1.0 + 1e-9);
    const int max_passes = 40;
    for (int pass = 0; pass < max_passes; ++pass) {
        hipLaunchKernelGGL(svm::k_dl_star, dim3((n + SVM_DL_WAVES - 1) / SVM_DL_WAVES), dim3(64 * SVM_DL_WAVES), 0, st,
                           (const double*)o.P, n, g, (const int*)start, (const int*)list, (const double*)rad, Rs,
                           cap, o.nbr, o.deg, status);
        hipLaunchKernelGGL(svm::k_dl_radius, blocks(n), dim3(tb), 0, st, status, rad, rlo, rhi, n, whole);
        // the pass's outcome, reduced on the device: {stars to redo, first failed point}
        const int init[2] = {0, n};
        int ho[2] = {0, n};
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(outcome, init, sizeof(init), hipMemcpyHostToDevice, st) != hipSuccess) {
            msg = "star kernel failed";
            return SVM_ERR_HIP;
        }
        hipLaunchKernelGGL(svm::k_dl_status, blocks(n), dim3(tb), 0, st, (const int*)status, n, outcome);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(ho, outcome, sizeof(ho), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            msg = "star kernel failed";
            return SVM_ERR_HIP;
        }
        const int retry = ho[0], where = ho[1] < n ? ho[1] : -1;
        int worst = 0;
        if (where >= 0 && (hipMemcpyAsync(&worst, status + where, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
                           hipStreamSynchronize(st) != hipSuccess)) {
            msg = "device copy failed";
            return SVM_ERR_HIP;
        }
        if (worst) {
            msg = worst == svm::DL_DEGENERATE ? "degenerate input: four co-circular points"
                  : worst == svm::DL_OPEN     ? "a cell has more Delaunay neighbours than ring_capacity"
                  : worst == svm::DL_CROWDED  ? "points too crowded for the neighbour search"
                                              : "inconsistent triangulation";
            msg += " (point " + std::to_string(where) + ", status " + std::to_string(worst) + ", pass " +
                   std::to_string(pass) + ")";
            return worst == svm::DL_OPEN ? SVM_ERR_RING_OVERFLOW : SVM_ERR_TOPOLOGY;
        }
        if (!retry) break;
        if (pass == max_passes - 1) { msg = "neighbour search did not converge"; return SVM_ERR_TOPOLOGY; }
    }
    // canonical triangles
    hipLaunchKernelGGL(svm::k_dl_tri_count, blocks(n), dim3(tb), 0, st, n, cap, (const int*)o.nbr,
                       (const int*)o.deg, cnt);
    if (!dl_scan(st, m, cnt, n + 1, o.off)) { msg = "scan failed"; return SVM_ERR_HIP; }
    int total = 0;
    if (hipMemcpyAsync(&total, o.off + n, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        msg = "device copy failed";
        return SVM_ERR_HIP;
    }
    if (total != 2 * n - 4) { msg = "triangle count is not 2N-4"; return SVM_ERR_TOPOLOGY; }
    o.ntri = total;
    o.T = m.get<int>(3 * (size_t)total);
    if (!o.T) { msg = "device allocation failed"; return SVM_ERR_HIP; }
    hipLaunchKernelGGL(svm::k_dl_tri_write, blocks(n), dim3(tb), 0, st, n, cap, (const int*)o.nbr,
                       (const int*)o.deg, (const int*)o.off, o.T);
    hipLaunchKernelGGL(svm::k_dl_tid, blocks(n), dim3(tb), 0, st, n, cap, (const int*)o.nbr, (const int*)o.deg,
                       (const int*)o.T, (const int*)o.off, o.tid, o.err);
    int herr = 0;
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&herr, o.err, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        msg = "triangle kernels failed";
        return SVM_ERR_HIP;
    }
    if (herr) { msg = "stars of neighbouring points disagree"; return SVM_ERR_TOPOLOGY; }
    return SVM_OK;
}

// Triangulation of n host points (AoS), checked to be finite and away from
// the origin.
int dl_triangulate(hipStream_t st, DevScratch& m, const double* pts, int n, int cap, DlOut& o, std::string& msg) {
    if (n < 4) { msg = "a spherical triangulation needs at least 4 points"; return SVM_ERR_ARG; }
    double Rs = 0.0;
    for (int i = 0; i < n; ++i) {
        const double* p = pts + 3 * (size_t)i;
        double r = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        if (!(r > 0.0) || !std::isfinite(r)) { msg = "points must be finite and away from the origin"; return SVM_ERR_ARG; }
        Rs = std::max(Rs, r);
    }
    double* P = m.get<double>(3 * (size_t)n);
    if (!P || hipMemcpyAsync(P, pts, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice, st) != hipSuccess) {
        msg = "device copy failed";
        return SVM_ERR_HIP;
    }
    return dl_triangulate_dev(st, m, P, n, Rs, cap, o, msg);
}

}  // namespace

// rows idx[0..n) of src (width floats each) into dst
__global__ __launch_bounds__(256) void k_gather_rows(const float* src, const int* idx, int n, int width, float* dst) {
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (size_t)n * width) return;
    const int r = (int)(e / width), k = (int)(e % width);
    dst[e] = src[(size_t)idx[r] * width + k];
}

// deterministic pseudo-random descriptors in [-1, 1) for timing runs
__global__ __launch_bounds__(256) void k_fill_hash(float* x, size_t n) {
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    unsigned h = (unsigned)e * 2654435761u;
    h ^= h >> 15;
    h *= 2246822519u;
    h ^= h >> 13;
    x[e] = (float)(h & 0xffffff) * (2.0f / 16777216.0f) - 1.0f;
}

// A plain streaming copy (n doubles read, n written): the yardstick that
// calibrates a profiler's HBM byte counters against known traffic.
__global__ __launch_bounds__(256) void k_copy_probe(const double* __restrict__ a, double* __restrict__ b, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

// ---- phase-boundary probe ----
// One phase of the probe's work: workgroup wg writes its row of `per` doubles
// into row buffer ph & 1, after reading (and checking) what the next
// workgroup wrote in phase ph - 1.  Rows of one buffer are read only after
// the boundary that follows their writes, and overwritten only after the
// boundary after that, so every read must see the previous phase's values.
__device__ inline void probe_phase(double* rows, int nwg, int wg, int ph, int per, unsigned* errors) {
    if (ph > 0) {
        const int src = wg + 1 < nwg ? wg + 1 : 0;
        const double* r = rows + ((size_t)((ph - 1) & 1) * nwg + src) * per;
        const double want = (double)((ph - 1) * nwg + src);
        unsigned bad = 0;
        for (int i = threadIdx.x; i < per; i += blockDim.x) bad += r[i] != want + (double)i;
        if (bad) atomicAdd(errors, bad);
    }
    double* w = rows + ((size_t)(ph & 1) * nwg + wg) * per;
    for (int i = threadIdx.x; i < per; i += blockDim.x) w[i] = (double)(ph * nwg + wg) + (double)i;
}

// The phases as one kernel launch per phase (the library's form).
__global__ __launch_bounds__(256) void k_probe_phase(double* rows, int per, int ph, unsigned* errors) {
    probe_phase(rows, (int)gridDim.x, (int)blockIdx.x, ph, per, errors);
}

// The same phases in one persistent launch, separated by a grid barrier:
// each workgroup publishes its writes (device-scope release), arrives on a
// counter, waits until the phase's nwg arrivals are in, then acquires.  The
// wait is bounded: past spin_limit polls (or once any workgroup gave up) the
// workgroup sets *gave_up and leaves, so every wavefront exits.  Launched
// cooperatively, so all workgroups are resident at once.
__global__ __launch_bounds__(256) void k_probe_persistent(double* rows, int per, int phases, unsigned* arrivals,
                                                          unsigned* errors, int* gave_up, int spin_limit) {
    __shared__ int s_quit;
    const int nwg = (int)gridDim.x, wg = (int)blockIdx.x;
    for (int ph = 0; ph < phases; ++ph) {
        probe_phase(rows, nwg, wg, ph, per, errors);
        __syncthreads();
        if (threadIdx.x == 0) {
            s_quit = 0;
            __threadfence();
            atomicAdd(arrivals, 1u);
            const unsigned target = (unsigned)(ph + 1) * (unsigned)nwg;
            int polls = 0;
            while (__hip_atomic_load(arrivals, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
                if (++polls > spin_limit || __hip_atomic_load(gave_up, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    __hip_atomic_store(gave_up, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    s_quit = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
            __threadfence();
        }
        __syncthreads();
        if (s_quit) return;
    }
    // the last phase's rows, checked as the next launch of k_probe_phase would
    if (phases > 0) {
        const int src = wg + 1 < nwg ? wg + 1 : 0;
        const double* r = rows + ((size_t)((phases - 1) & 1) * nwg + src) * per;
        const double want = (double)((phases - 1) * nwg + src);
        unsigned bad = 0;
        for (int i = threadIdx.x; i < per; i += blockDim.x) bad += r[i] != want + (double)i;
        if (bad) atomicAdd(errors, bad);
    }
}

extern "C" {

const char* svm_last_error(svm_context* ctx) { return ctx ? ctx->err.c_str() : k_ok; }

int svm_create(const svm_system_desc* systems, int32_t n_systems, svm_context** out) {
    if (!systems || n_systems <= 0 || !out) return SVM_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return SVM_ERR_NO_DEVICE;
    svm_context* ctx = new svm_context();
    if (hipGetDevice(&ctx->device) != hipSuccess) { delete ctx; return SVM_ERR_NO_DEVICE; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->device) != hipSuccess || std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        delete ctx;
        return SVM_ERR_NO_DEVICE;
    }
    ctx->nsys = n_systems;
    ctx->desc.assign(systems, systems + n_systems);
    ctx->layout.resize(n_systems);
    ctx->have_topology.assign(n_systems, 0);
    ctx->have_adam.assign(n_systems, 0);
    ctx->voronoi.assign(n_systems, 0);
    ctx->triangles.resize(n_systems);
    int cap = 0;
    long long Cs = 0, Vs = 0;
    for (int s = 0; s < n_systems; ++s) {
        const svm_system_desc& sd = systems[s];
        if (sd.n_cells < 4 || sd.ring_capacity < 0 || sd.ring_capacity > 64) {
            delete ctx;
            return SVM_ERR_ARG;
        }
        cap = std::max(cap, sd.ring_capacity ? (int)sd.ring_capacity : 32);
        SysParams& L = ctx->layout[s];
        std::memset(&L, 0, sizeof(L));
        L.nc = sd.n_cells;
        L.nv = 2 * sd.n_cells - 4;
        L.co = (int)Cs;
        L.vo = (int)Vs;
        L.ctile0 = (int)(Cs / SVM_TILE);
        L.nctiles = round_up(L.nc, SVM_TILE) / SVM_TILE;
        if (L.nctiles > SVM_MAX_TILES || round_up(2 * L.nc - 4, SVM_TILE) / SVM_TILE > SVM_MAX_TILES) {
            delete ctx;
            return SVM_ERR_ARG;   // more than 2^20 rows in one sphere
        }
        L.vtile0 = (int)(Vs / SVM_TILE);
        L.nvtiles = round_up(L.nv, SVM_TILE) / SVM_TILE;
        Cs += round_up(L.nc, SVM_TILE);
        Vs += round_up(L.nv, SVM_TILE);
    }
    if (Cs > (1ll << 30) || Vs > (1ll << 30)) { delete ctx; return SVM_ERR_ARG; }
    ctx->Cs = (int)Cs;
    ctx->Vs = (int)Vs;
    ctx->cap = cap;
    ctx->nct = (int)(Cs / SVM_TILE);
    ctx->nvt = (int)(Vs / SVM_TILE);
    if (hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking) != hipSuccess) { delete ctx; return SVM_ERR_HIP; }
    ctx->stream = ctx->own_stream;

    Dev& d = ctx->d;
    d.nsys = n_systems;
    d.Cs = ctx->Cs;
    d.xcd = 1;
    d.s0 = 0;
    d.ns = n_systems;
    d.vt0 = 0;
    d.vtn = ctx->nvt;
    d.ct0 = 0;
    d.ctn = ctx->nct;
    d.Vs = ctx->Vs;
    d.cap = cap;
    int r = 0;
    int32_t *cts = nullptr, *vts = nullptr;
    r |= dalloc(ctx, &d.sys, n_systems);
    r |= dalloc(ctx, &cts, ctx->nct);
    r |= dalloc(ctx, &vts, ctx->nvt);
    r |= dalloc(ctx, &d.cnn, Cs);
    r |= dalloc(ctx, &d.ring, (size_t)Cs * cap);
    r |= dalloc(ctx, &d.A, Cs);
    r |= dalloc(ctx, &d.P, Cs);
    r |= dalloc(ctx, &d.cx, Cs);
    r |= dalloc(ctx, &d.cy, Cs);
    r |= dalloc(ctx, &d.cz, Cs);
    r |= dalloc(ctx, &d.ctrig, (size_t)4 * Cs);
    r |= dalloc(ctx, &d.cvlist, Vs);
    r |= dalloc(ctx, &d.cvseen, (Vs + 31) / 32);
    r |= dalloc(ctx, &d.a0, Cs);
    r |= dalloc(ctx, &d.p0, Cs);
    for (int b = 0; b < 2; ++b) {
        r |= dalloc(ctx, &d.px[b], Vs);
        r |= dalloc(ctx, &d.py[b], Vs);
        r |= dalloc(ctx, &d.pz[b], Vs);
        r |= dalloc(ctx, &d.nx[b], Vs);
        r |= dalloc(ctx, &d.ny[b], Vs);
        r |= dalloc(ctx, &d.nz[b], Vs);
    }
    r |= dalloc(ctx, &d.vx, Vs);
    r |= dalloc(ctx, &d.vy, Vs);
    r |= dalloc(ctx, &d.vz, Vs);
    r |= dalloc(ctx, &d.fx, Vs);
    r |= dalloc(ctx, &d.fy, Vs);
    r |= dalloc(ctx, &d.fz, Vs);
    r |= dalloc(ctx, &d.vcn, Vs);
    r |= dalloc(ctx, &d.vn, Vs);
    r |= dalloc(ctx, &d.vln, Vs);
    r |= dalloc(ctx, &d.t1mask, Vs / 64);
    r |= dalloc(ctx, &d.t1fmask, Vs / 64);
    r |= dalloc(ctx, &d.t1_vidx, Vs);
    r |= dalloc(ctx, &d.cand, Vs);
    r |= dalloc(ctx, &d.cpart, ctx->nct);
    r |= dalloc(ctx, &d.vpart, 3 * (size_t)ctx->nvt);
    r |= dalloc(ctx, &d.disp, 3 * (size_t)Vs);
    r |= dalloc(ctx, &d.t1_count, 1);
    r |= dalloc(ctx, &d.t1_list, SVM_T1_MAXCL);
    r |= dalloc(ctx, &d.t1_flagged, SVM_T1_MAXCL);
    r |= dalloc(ctx, &d.t1_logged, SVM_T1_MAXCL);
    r |= dalloc(ctx, &d.t1_lead, SVM_T1_MAXCL);
    d.t1_near = SVM_T1_NEAR;
    r |= dalloc(ctx, &d.t1_total, 1);
    r |= dalloc(ctx, &d.t1_btotal, 1);
    r |= dalloc(ctx, &d.t1_bcount, 1);
    r |= dalloc(ctx, &d.t1_blist, SVM_T1_MAXBL);
    r |= dalloc(ctx, &d.t1_bflag, SVM_T1_MAXBL);
    r |= dalloc(ctx, &d.t1_bskip, SVM_T1_MAXBL);
    r |= dalloc(ctx, &d.t1_group, n_systems);
    r |= dalloc(ctx, &d.t1_log, SVM_T1_MAXCL);
    r |= dalloc(ctx, &d.vown, Vs);
    r |= dalloc(ctx, &d.cown, Cs);
    if (r) {
        svm_destroy(ctx);
        return SVM_ERR_HIP;
    }
    d.ctile_sys = cts;
    d.vtile_sys = vts;
    std::vector<int32_t> hct(ctx->nct), hvt(ctx->nvt);
    std::vector<SysParams> hs(n_systems);
    for (int s = 0; s < n_systems; ++s) {
        SysParams& L = ctx->layout[s];
        for (int t = 0; t < L.nctiles; ++t) hct[L.ctile0 + t] = s;
        for (int t = 0; t < L.nvtiles; ++t) hvt[L.vtile0 + t] = s;
        SysParams P = L;
        const svm_system_desc& sd = systems[s];
        P.R = std::sqrt(sd.n_cells / (4.0 * SVM_PI));
        P.kr = sd.kr;
        P.thr = sd.t1_threshold;
        P.thr_cos = svm::t1_cos_cut(P.thr, P.R);
        P.seed = sd.seed;
        P.dt = 0.01; P.mu = 1.0; P.T = 0.0; P.v0 = 0.01; P.tau = 1.0;   // updater defaults
        P.f_dt0 = 0.01; P.f_alpha_start = 0.99; P.f_dt_max = 0.1; P.f_dt_inc = 1.05; P.f_dt_dec = 0.95;
        P.f_alpha_dec = 0.9; P.f_cutoff = 1e-7; P.f_alpha_min = 0.0; P.f_dt_min = 0.01 * 0.01;
        P.f_nmin = 5; P.f_max_iter = 1000;
        P.f_dt = 0.01; P.f_alpha = 0.99; P.f_force_max = 100.;
        P.t1_head = -1;
        P.t1_gen = 1;   // stamps of generation 0 mean "never touched"
        hs[s] = P;
    }
    hipStream_t st = ctx->own_stream;
    if (hipMemcpyAsync(cts, hct.data(), sizeof(int32_t) * ctx->nct, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(vts, hvt.data(), sizeof(int32_t) * ctx->nvt, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(d.sys, hs.data(), sizeof(SysParams) * n_systems, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        svm_destroy(ctx);
        return SVM_ERR_HIP;
    }
    // preferences
    std::vector<double> a0(Cs, 0.0), p0(Cs, 0.0);
    for (int s = 0; s < n_systems; ++s)
        for (int c = 0; c < ctx->layout[s].nc; ++c) {
            a0[ctx->layout[s].co + c] = systems[s].a0;
            p0[ctx->layout[s].co + c] = systems[s].p0;
        }
    if (hipMemcpyAsync(d.a0, a0.data(), sizeof(double) * Cs, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(d.p0, p0.data(), sizeof(double) * Cs, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        svm_destroy(ctx);
        return SVM_ERR_HIP;
    }
    *out = ctx;
    return SVM_OK;
}

int svm_destroy(svm_context* ctx) {
    if (!ctx) return SVM_ERR_ARG;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto st : ctx->lane_streams) (void)hipStreamSynchronize(st);
    drop_graphs(ctx);
    for (auto st : ctx->lane_streams) (void)hipStreamDestroy(st);
    for (auto ev : ctx->lane_joins) (void)hipEventDestroy(ev);
    if (ctx->lane_fork) (void)hipEventDestroy(ctx->lane_fork);
    for (void* p : ctx->allocs) (void)hipFree(p);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return SVM_OK;
}

int svm_set_stream(svm_context* ctx, void* stream) {
    if (!ctx) return SVM_ERR_ARG;
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->stream = stream ? (hipStream_t)stream : ctx->own_stream;
    drop_graphs(ctx);
    return SVM_OK;
}

int svm_synchronize(svm_context* ctx) {
    if (!ctx) return SVM_ERR_ARG;
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    return check_device_errors(ctx);
}

int svm_get_num_vertices(svm_context* ctx, int32_t sys, int32_t* n_vertices) {
    if (!valid_sys(ctx, sys) || !n_vertices) return fail(ctx, SVM_ERR_ARG, "bad argument");
    *n_vertices = ctx->layout[sys].nv;
    return SVM_OK;
}

int svm_generate_cell_points(int32_t n_cells, uint64_t seed, double* cell_points) {
    if (n_cells < 4 || !cell_points) return SVM_ERR_ARG;
    svm::generate_cell_points(n_cells, seed, cell_points);
    return SVM_OK;
}

static uint64_t spread3(uint64_t x) {   // 21 bits -> every third bit of 63
    x &= 0x1fffffull;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

int svm_spatial_order(int32_t n_points, const double* points, int32_t* order) {
    if (n_points < 1 || !points || !order) return SVM_ERR_ARG;
    double lo[3] = {points[0], points[1], points[2]}, hi[3] = {points[0], points[1], points[2]};
    for (int i = 0; i < n_points; ++i)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], points[3 * (size_t)i + k]);
            hi[k] = std::max(hi[k], points[3 * (size_t)i + k]);
        }
    const double span = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-300});
    std::vector<std::pair<uint64_t, int32_t>> key(n_points);
    for (int i = 0; i < n_points; ++i) {
        uint64_t q[3];
        for (int k = 0; k < 3; ++k) {
            double u = (points[3 * (size_t)i + k] - lo[k]) / span;
            q[k] = (uint64_t)std::min(2097151.0, std::max(0.0, u * 2097152.0));
        }
        key[i] = {spread3(q[0]) | spread3(q[1]) << 1 | spread3(q[2]) << 2, i};
    }
    std::sort(key.begin(), key.end());
    for (int i = 0; i < n_points; ++i) order[i] = key[i].second;
    return SVM_OK;
}

int svm_generate_directors(int32_t n_vertices, const double* positions, uint64_t seed, double* directors) {
    if (n_vertices < 1 || !positions || !directors) return SVM_ERR_ARG;
    svm::generate_directors(n_vertices, positions, seed, directors);
    return SVM_OK;
}

int svm_set_topology(svm_context* ctx, int32_t sys, int32_t n_vertices, const double* vertex_positions,
                     const int32_t* vertex_neighbors, const int32_t* vertex_cell_neighbors,
                     const int32_t* cell_vertices, int32_t ring_stride, const int32_t* cell_vertex_counts) {
    if (!valid_sys(ctx, sys) || !vertex_positions || !vertex_neighbors || !vertex_cell_neighbors || !cell_vertices ||
        !cell_vertex_counts)
        return fail(ctx, SVM_ERR_ARG, "null argument");
    const SysParams& L = ctx->layout[sys];
    if (n_vertices != L.nv) return fail(ctx, SVM_ERR_TOPOLOGY, "vertex count must be 2N-4");
    // host-side validation and conversion to global rows
    std::vector<int4> vcn(L.nv), vn(L.nv);
    for (int v = 0; v < L.nv; ++v) {
        int a[3], b[3];
        for (int k = 0; k < 3; ++k) {
            a[k] = vertex_cell_neighbors[3 * (size_t)v + k];
            b[k] = vertex_neighbors[3 * (size_t)v + k];
            if (a[k] < 0 || a[k] >= L.nc || b[k] < 0 || b[k] >= L.nv)
                return fail(ctx, SVM_ERR_TOPOLOGY, "vertex list index out of range");
        }
        vcn[v] = make_int4(a[0] + L.co, a[1] + L.co, a[2] + L.co, sys);
        vn[v] = make_int4(b[0] + L.vo, b[1] + L.vo, b[2] + L.vo, sys);
    }
    std::vector<int32_t> ring((size_t)ctx->cap * L.nc, -1), cnt(L.nc);
    for (int c = 0; c < L.nc; ++c) {
        int n = cell_vertex_counts[c];
        if (n < 3 || n > ctx->cap || n > ring_stride)
            return fail(ctx, SVM_ERR_RING_OVERFLOW, "cell vertex count outside [3, ring_capacity]");
        cnt[c] = n;
        for (int k = 0; k < n; ++k) {
            int v = cell_vertices[(size_t)c * ring_stride + k];
            if (v < 0 || v >= L.nv) return fail(ctx, SVM_ERR_TOPOLOGY, "ring index out of range");
            ring[(size_t)k * L.nc + c] = v + L.vo;
        }
    }
    Dev& d = ctx->d;
    HIPCHECK(ctx, hipMemcpyAsync(d.vcn + L.vo, vcn.data(), sizeof(int4) * L.nv, hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(ctx, hipMemcpyAsync(d.vn + L.vo, vn.data(), sizeof(int4) * L.nv, hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(ctx, hipMemcpyAsync(d.cnn + L.co, cnt.data(), sizeof(int32_t) * L.nc, hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(ctx, hipMemcpy2DAsync(d.ring + L.co, sizeof(int32_t) * ctx->Cs, ring.data(), sizeof(int32_t) * L.nc,
                                   sizeof(int32_t) * L.nc, ctx->cap, hipMemcpyHostToDevice, ctx->stream));
    int r = upload_vec3(ctx, sys, vertex_positions, d.px[ctx->pcur], d.py[ctx->pcur], d.pz[ctx->pcur]);
    if (r) return r;
    hipLaunchKernelGGL(svm::k_build_vln, dim3(ctx->nvt), dim3(SVM_TILE), 0, ctx->stream, ctx->d, sys);
    r = last_launch(ctx);
    if (r) return r;
    r = svm_synchronize(ctx);
    if (r) return r;
    ctx->have_topology[sys] = 1;
    ctx->voronoi[sys] = 0;
    return SVM_OK;
}

// Vertex-model topology from a spherical Delaunay triangulation, in the
// conventions of convexHullCGALInterface::sphericalConvexHullForVertexModel
// (src/model/convexHullCGAL.cpp:82-186): a vertex per triangle at
// R_hull * normalize((a + b + c) / 3) (getFacetCenter, :67-80; R_hull = |p_0|,
// :84), summed in the triangle's circulation order; vertex slot k holds cell
// t[k] and the triangle across the edge (t[k-1], t[k]); cell rings list the
// triangles around each point clockwise seen from outside (the halfedge-
// around-vertex order), starting at the lowest triangle index.  Triangles are
// first put in canonical form: counter-clockwise seen from outside, smallest
// index first, sorted.
static int topology_from_triangles(svm_context* ctx, int sys, const double* pts, int ntri, const int32_t* tris_in) {
    const SysParams& L = ctx->layout[sys];
    const int n = L.nc;
    if (ntri != 2 * n - 4) return fail(ctx, SVM_ERR_TOPOLOGY, "a triangulation of the sphere has 2N-4 triangles");
    std::vector<std::array<int, 3>> T(ntri);
    for (int t = 0; t < ntri; ++t) {
        int a = tris_in[3 * t], b = tris_in[3 * t + 1], c = tris_in[3 * t + 2];
        if (a < 0 || b < 0 || c < 0 || a >= n || b >= n || c >= n || a == b || b == c || a == c)
            return fail(ctx, SVM_ERR_TOPOLOGY, "bad triangle");
        if (svm::orient_origin(pts + 3 * a, pts + 3 * b, pts + 3 * c) < 0) std::swap(b, c);
        int m[3] = {a, b, c};
        int k = (a < b) ? (a < c ? 0 : 2) : (b < c ? 1 : 2);
        T[t] = {m[k], m[(k + 1) % 3], m[(k + 2) % 3]};
    }
    std::sort(T.begin(), T.end());
    std::vector<int32_t> flat(3 * (size_t)ntri);
    for (int t = 0; t < ntri; ++t)
        for (int k = 0; k < 3; ++k) flat[3 * (size_t)t + k] = T[t][k];
    // directed edge a->b -> triangle; each point's out-edges
    std::vector<int> deg(n, 0);
    for (int t = 0; t < ntri; ++t)
        for (int k = 0; k < 3; ++k) deg[T[t][k]]++;
    std::vector<int> off(n + 1, 0);
    for (int i = 0; i < n; ++i) off[i + 1] = off[i] + deg[i];
    std::vector<int> eb(off[n]), et(off[n]), fill(off.begin(), off.end() - 1);
    for (int t = 0; t < ntri; ++t)
        for (int k = 0; k < 3; ++k) {
            int a = T[t][k], b = T[t][(k + 1) % 3];
            eb[fill[a]] = b;
            et[fill[a]] = t;
            fill[a]++;
        }
    auto tri_of = [&](int a, int b) -> int {
        for (int e = off[a]; e < off[a + 1]; ++e)
            if (eb[e] == b) return et[e];
        return -1;
    };
    int maxdeg = 0;
    for (int i = 0; i < n; ++i) maxdeg = std::max(maxdeg, deg[i]);
    if (maxdeg > ctx->cap) return fail(ctx, SVM_ERR_RING_OVERFLOW, "a cell has more vertices than ring_capacity");
    double radius = std::sqrt(pts[0] * pts[0] + pts[1] * pts[1] + pts[2] * pts[2]);
    std::vector<double> vpos(3 * (size_t)ntri);
    std::vector<int32_t> vn(3 * (size_t)ntri), vcn(3 * (size_t)ntri);
    for (int t = 0; t < ntri; ++t) {
        double c[3] = {0.0, 0.0, 0.0};
        for (int k = 0; k < 3; ++k) {
            const double* p = pts + 3 * (size_t)T[t][k];
            c[0] += p[0];
            c[1] += p[1];
            c[2] += p[2];
        }
        for (int k = 0; k < 3; ++k) c[k] = (1.0 / 3) * c[k];
        double nrm = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        for (int k = 0; k < 3; ++k) vpos[3 * (size_t)t + k] = (radius / nrm) * c[k];
        for (int k = 0; k < 3; ++k) {
            int prev = T[t][(k + 2) % 3], cur = T[t][k];
            vcn[3 * (size_t)t + k] = cur;
            int nb = tri_of(cur, prev);
            if (nb < 0) return fail(ctx, SVM_ERR_TOPOLOGY, "triangulation is not closed");
            vn[3 * (size_t)t + k] = nb;
        }
    }
    std::vector<int32_t> ring((size_t)n * ctx->cap, -1), cnt(n);
    for (int c = 0; c < n; ++c) {
        int start = -1, b = -1;
        for (int e = off[c]; e < off[c + 1]; ++e)
            if (start < 0 || et[e] < start) { start = et[e]; b = eb[e]; }
        int t = start;
        for (int k = 0; k < deg[c]; ++k) {
            ring[(size_t)c * ctx->cap + k] = t;
            int tn = tri_of(b, c);
            if (tn < 0) return fail(ctx, SVM_ERR_TOPOLOGY, "triangulation is not closed");
            int y = -1;
            for (int kk = 0; kk < 3; ++kk)
                if (T[tn][kk] == c) y = T[tn][(kk + 1) % 3];
            t = tn;
            b = y;
        }
        if (t != start) return fail(ctx, SVM_ERR_TOPOLOGY, "a point's triangles do not close into one ring");
        cnt[c] = deg[c];
    }
    ctx->triangles[sys] = flat;
    return svm_set_topology(ctx, sys, ntri, vpos.data(), vn.data(), vcn.data(), ring.data(), ctx->cap, cnt.data());
}

int svm_set_topology_from_triangles(svm_context* ctx, int32_t sys, const double* cell_points, int32_t n_triangles,
                                    const int32_t* triangles) {
    if (!valid_sys(ctx, sys) || !cell_points || !triangles) return fail(ctx, SVM_ERR_ARG, "null argument");
    return topology_from_triangles(ctx, sys, cell_points, n_triangles, triangles);
}

int svm_get_topology(svm_context* ctx, int32_t sys, int32_t* vertex_neighbors, int32_t* vertex_cell_neighbors,
                     int32_t* cell_vertices, int32_t ring_stride, int32_t* cell_vertex_counts) {
    if (!valid_sys(ctx, sys) || !vertex_neighbors || !vertex_cell_neighbors || !cell_vertices || !cell_vertex_counts)
        return fail(ctx, SVM_ERR_ARG, "null argument");
    const SysParams& L = ctx->layout[sys];
    Dev& d = ctx->d;
    std::vector<int4> vcn(L.nv), vn(L.nv);
    std::vector<int32_t> ring((size_t)ctx->cap * L.nc), cnt(L.nc);
    HIPCHECK(ctx, hipMemcpyAsync(vcn.data(), d.vcn + L.vo, sizeof(int4) * L.nv, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(ctx, hipMemcpyAsync(vn.data(), d.vn + L.vo, sizeof(int4) * L.nv, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(ctx, hipMemcpyAsync(cnt.data(), d.cnn + L.co, sizeof(int32_t) * L.nc, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(ctx, hipMemcpy2DAsync(ring.data(), sizeof(int32_t) * L.nc, d.ring + L.co, sizeof(int32_t) * ctx->Cs,
                                   sizeof(int32_t) * L.nc, ctx->cap, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    int maxn = 0;
    for (int c = 0; c < L.nc; ++c) maxn = std::max(maxn, cnt[c]);
    if (maxn > ring_stride) return fail(ctx, SVM_ERR_ARG, "ring_stride smaller than the largest cell");
    for (int v = 0; v < L.nv; ++v) {
        const int* a = (const int*)&vcn[v];
        const int* b = (const int*)&vn[v];
        for (int k = 0; k < 3; ++k) {
            vertex_cell_neighbors[3 * (size_t)v + k] = a[k] - L.co;
            vertex_neighbors[3 * (size_t)v + k] = b[k] - L.vo;
        }
    }
    for (int c = 0; c < L.nc; ++c) {
        cell_vertex_counts[c] = cnt[c];
        for (int k = 0; k < ring_stride; ++k)
            cell_vertices[(size_t)c * ring_stride + k] = k < cnt[c] ? ring[(size_t)k * L.nc + c] - L.vo : -1;
    }
    return SVM_OK;
}

// The topology of sphere `sys` from the GPU triangulation of its cell
// points, vertices at the triangles' centroids (circum 0, the vertex model)
// or circumcentres (circum 1, the Voronoi model, whose generators are kept).
static int build_topology(svm_context* ctx, int32_t sys, const double* cell_points, int circum) {
    if (!valid_sys(ctx, sys) || !cell_points) return fail(ctx, SVM_ERR_ARG, "null argument");
    const SysParams& L = ctx->layout[sys];
    const int n = L.nc;
    DevScratch m;
    DlOut o;
    std::string msg;
    int r = dl_triangulate(ctx->stream, m, cell_points, n, ctx->cap, o, msg);
    if (r) return fail(ctx, r, msg);
    const double radius =
        std::sqrt(cell_points[0] * cell_points[0] + cell_points[1] * cell_points[1] + cell_points[2] * cell_points[2]);
    const int tb = 256;
    hipLaunchKernelGGL(svm::k_dl_vertices, dim3((o.ntri + tb - 1) / tb), dim3(tb), 0, ctx->stream, ctx->d, sys,
                       ctx->pcur, (const double*)o.P, o.ntri, ctx->cap, (const int*)o.T, (const int*)o.nbr,
                       (const int*)o.deg, (const int*)o.tid, radius, o.err, circum);
    if (circum) {
        if (!ctx->d.gen) {
            r = dalloc(ctx, &ctx->d.gen, 3 * (size_t)ctx->Cs);
            if (!r) r = dalloc(ctx, &ctx->d.cf, 3 * (size_t)ctx->Cs);
            if (r) return r;
        }
        hipLaunchKernelGGL(svm::k_store_generators, dim3((n + tb - 1) / tb), dim3(tb), 0, ctx->stream, ctx->d, sys,
                           (const double*)o.P, n);
    }
    hipLaunchKernelGGL(svm::k_dl_rings, dim3((n + tb - 1) / tb), dim3(tb), 0, ctx->stream, ctx->d, sys, n, ctx->cap,
                       (const int*)o.deg, (const int*)o.tid);
    hipLaunchKernelGGL(svm::k_build_vln, dim3(ctx->nvt), dim3(SVM_TILE), 0, ctx->stream, ctx->d, sys);
    r = last_launch(ctx);
    if (r) return r;
    int herr = 0;
    std::vector<int32_t> T(3 * (size_t)o.ntri);
    HIPCHECK(ctx, hipMemcpyAsync(&herr, o.err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(ctx, hipMemcpyAsync(T.data(), o.T, sizeof(int32_t) * T.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (herr) return fail(ctx, SVM_ERR_TOPOLOGY, "triangulation is not closed");
    r = check_device_errors(ctx);
    if (r) return r;
    ctx->triangles[sys] = std::move(T);
    ctx->have_topology[sys] = 1;
    ctx->voronoi[sys] = circum;
    return SVM_OK;
}

int svm_build_from_cell_points(svm_context* ctx, int32_t sys, const double* cell_points) {
    return build_topology(ctx, sys, cell_points, 0);
}

int svm_build_voronoi_from_cell_points(svm_context* ctx, int32_t sys, const double* cell_points) {
    return build_topology(ctx, sys, cell_points, 1);
}

int svm_voronoi_forces(svm_context* ctx) {
    if (!ctx) return SVM_ERR_ARG;
    int r = require_topology(ctx);
    if (r) return r;
    for (int s = 0; s < ctx->nsys; ++s)
        if (!ctx->voronoi[s])
            return fail(ctx, SVM_ERR_STATE, "sphere " + std::to_string(s) + " was not built as a Voronoi model");
    r = svm_compute_forces(ctx, ctx->force_fp32);
    if (r) return r;
    hipLaunchKernelGGL(svm::k_voronoi_chain, dim3(ctx->nct), dim3(SVM_TILE), 0, ctx->stream, ctx->d);
    r = last_launch(ctx);
    if (r) return r;
    return check_device_errors(ctx);
}

// Retriangulation of a Voronoi-model sphere from its device-resident
// generators: the topology, circumcentre vertices and vertex lists of
// build_topology(..., circum = 1), without host copies of the points.  The
// circumcentres use the radius of the sphere's first generator, as the
// host-points build does; `keep_triangles` also copies the triangles to the
// host (for svm_get_delaunay_triangles).
static int retopologise_voronoi(svm_context* ctx, int32_t sys, DevScratch& m, bool keep_triangles) {
    const SysParams& L = ctx->layout[sys];
    const int n = L.nc;
    const int tb = 256;
    m.rewind();
    double* P = m.get<double>(3 * (size_t)n);
    if (!P) return fail(ctx, SVM_ERR_HIP, "device allocation failed");
    hipLaunchKernelGGL(svm::k_gen_aos, dim3((n + tb - 1) / tb), dim3(tb), 0, ctx->stream, ctx->d, sys, P);
    double p0[3];
    for (int k = 0; k < 3; ++k)
        HIPCHECK(ctx, hipMemcpyAsync(&p0[k], ctx->d.gen + (size_t)k * ctx->Cs + L.co, sizeof(double),
                                     hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    const double radius = std::sqrt(p0[0] * p0[0] + p0[1] * p0[1] + p0[2] * p0[2]);
    if (!(radius > 0.0) || !std::isfinite(radius)) return fail(ctx, SVM_ERR_STATE, "generators are not finite");
    DlOut o;
    std::string msg;
    // every generator lies on the sphere of the model's radius (to rounding)
    const double R_model = std::sqrt(ctx->desc[sys].n_cells / (4.0 * SVM_PI));
    int r = dl_triangulate_dev(ctx->stream, m, P, n, std::max(radius, R_model) * (1.0 + 1e-12), ctx->cap, o, msg);
    if (r) return fail(ctx, r, msg);
    hipLaunchKernelGGL(svm::k_dl_vertices, dim3((o.ntri + tb - 1) / tb), dim3(tb), 0, ctx->stream, ctx->d, sys,
                       ctx->pcur, (const double*)o.P, o.ntri, ctx->cap, (const int*)o.T, (const int*)o.nbr,
                       (const int*)o.deg, (const int*)o.tid, radius, o.err, 1);
    hipLaunchKernelGGL(svm::k_dl_rings, dim3((n + tb - 1) / tb), dim3(tb), 0, ctx->stream, ctx->d, sys, n, ctx->cap,
                       (const int*)o.deg, (const int*)o.tid);
    hipLaunchKernelGGL(svm::k_build_vln, dim3(ctx->nvt), dim3(SVM_TILE), 0, ctx->stream, ctx->d, sys);
    r = last_launch(ctx);
    if (r) return r;
    int herr = 0;
    HIPCHECK(ctx, hipMemcpyAsync(&herr, o.err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    if (keep_triangles) {
        ctx->triangles[sys].resize(3 * (size_t)o.ntri);
        HIPCHECK(ctx, hipMemcpyAsync(ctx->triangles[sys].data(), o.T, sizeof(int32_t) * 3 * (size_t)o.ntri,
                                     hipMemcpyDeviceToHost, ctx->stream));
    }
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (herr) return fail(ctx, SVM_ERR_TOPOLOGY, "triangulation is not closed");
    return SVM_OK;
}

int svm_voronoi_brownian_steps(svm_context* ctx, int32_t n_steps) {
    if (!ctx || n_steps < 0) return SVM_ERR_ARG;
    int r = require_topology(ctx);
    if (r) return r;
    for (int s = 0; s < ctx->nsys; ++s)
        if (!ctx->voronoi[s])
            return fail(ctx, SVM_ERR_STATE, "sphere " + std::to_string(s) + " was not built as a Voronoi model");
    DevScratch m;   // reused by every retriangulation of the call
    for (int i = 0; i < n_steps; ++i) {
        r = svm_compute_forces(ctx, ctx->force_fp32);
        if (r) return r;
        hipLaunchKernelGGL(svm::k_voronoi_chain, dim3(ctx->nct), dim3(SVM_TILE), 0, ctx->stream, ctx->d);
        hipLaunchKernelGGL(svm::k_voronoi_move, dim3(ctx->nct), dim3(SVM_TILE), 0, ctx->stream, ctx->d);
        hipLaunchKernelGGL(svm::k_advance_step, dim3(ctx->nsys), dim3(64), 0, ctx->stream, ctx->d);
        r = last_launch(ctx);
        if (r) return r;
        for (int s = 0; s < ctx->nsys; ++s) {
            r = retopologise_voronoi(ctx, s, m, i == n_steps - 1);
            if (r) return r;
        }
    }
    return check_device_errors(ctx);
}

int svm_get_generators(svm_context* ctx, int32_t sys, double* points) {
    if (!valid_sys(ctx, sys) || !points) return fail(ctx, SVM_ERR_ARG, "bad argument");
    if (!ctx->voronoi[sys]) return fail(ctx, SVM_ERR_STATE, "sphere was not built as a Voronoi model");
    const SysParams& L = ctx->layout[sys];
    const double* g = ctx->d.gen;
    const size_t Cs = ctx->Cs;
    return download_rows(ctx, L.nc, L.co, g, g + Cs, g + 2 * Cs, points);
}

int svm_get_voronoi_forces(svm_context* ctx, int32_t sys, double* forces) {
    if (!valid_sys(ctx, sys) || !forces) return fail(ctx, SVM_ERR_ARG, "bad argument");
    if (!ctx->voronoi[sys]) return fail(ctx, SVM_ERR_STATE, "sphere was not built as a Voronoi model");
    const SysParams& L = ctx->layout[sys];
    const double* cf = ctx->d.cf;
    const size_t Cs = ctx->Cs;
    return download_rows(ctx, L.nc, L.co, cf, cf + Cs, cf + 2 * Cs, forces);
}

int svm_cell_list(int32_t n_points, const double* points, double cell_size, int32_t* grid_dim, int32_t* bin_of,
                  int32_t* bin_start, int32_t* sorted_points) {
    if (n_points < 1 || !points || !(cell_size > 0.0) || !grid_dim) return SVM_ERR_ARG;
    double Rs = 0.0;
    for (int i = 0; i < 3 * n_points; ++i) {
        if (!std::isfinite(points[i])) return SVM_ERR_ARG;
        Rs = std::max(Rs, std::fabs(points[i]));
    }
    if (!(Rs > 0.0)) return SVM_ERR_ARG;
    const int G = std::max(1, std::min(256, (int)std::ceil(2.0 * Rs * (1.0 + 1e-9) / cell_size)));
    *grid_dim = G;
    if (!bin_of || !bin_start || !sorted_points) return SVM_OK;   // size query
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return SVM_ERR_NO_DEVICE;
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return SVM_ERR_HIP;
    int r = SVM_OK;
    {
        DevScratch m;
        DlCells c;
        double* P = m.get<double>(3 * (size_t)n_points);
        if (!P || hipMemcpyAsync(P, points, sizeof(double) * 3 * (size_t)n_points, hipMemcpyHostToDevice, st) != hipSuccess ||
            !dl_cell_list(st, m, P, n_points, Rs, cell_size, c) ||
            hipMemcpyAsync(bin_of, c.bin_of, sizeof(int) * n_points, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(bin_start, c.start, sizeof(int) * (c.nbins + 1), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(sorted_points, c.list, sizeof(int) * n_points, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            r = SVM_ERR_HIP;
    }
    (void)hipStreamDestroy(st);
    return r;
}

int svm_neighbor_list(int32_t n_points, const double* points, double range, double box_half, int32_t max_neighbors,
                      int32_t* counts, int32_t* indices, double* vectors, double* distances, int32_t* needed) {
    if (n_points < 1 || !points || !(range > 0.0) || !std::isfinite(range) || max_neighbors < 0 || !counts ||
        !needed || (max_neighbors > 0 && (!indices || !vectors || !distances)))
        return SVM_ERR_ARG;
    double Rs = 0.0;
    for (size_t i = 0; i < 3 * (size_t)n_points; ++i) {
        if (!std::isfinite(points[i])) return SVM_ERR_ARG;
        Rs = std::max(Rs, std::fabs(points[i]));
    }
    // no periodic image may change a difference within range or bring a far
    // pair within range: box_half >= range and 2 box_half - 2 Rs >= range
    if (!(box_half >= range) || !(2.0 * box_half - 2.0 * Rs >= range)) return SVM_ERR_ARG;
    if (!(Rs > 0.0)) Rs = range;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return SVM_ERR_NO_DEVICE;
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return SVM_ERR_HIP;
    int r = SVM_OK;
    {
        DevScratch m;
        DlCells c;
        const size_t n = (size_t)n_points;
        double* P = m.get<double>(3 * n);
        int* cnt = m.get<int>(n);
        int* maxc = m.get<int>(1);
        const int tb = 256;
        const dim3 grid((unsigned)((n + tb - 1) / tb));
        int hmax = 0;
        if (!P || !cnt || !maxc ||
            hipMemcpyAsync(P, points, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st) != hipSuccess ||
            !dl_cell_list(st, m, P, n_points, Rs, range, c)) {
            r = SVM_ERR_HIP;
        } else {
            hipLaunchKernelGGL(svm::k_fill_int, dim3(1), dim3(tb), 0, st, maxc, 1, 0);
            hipLaunchKernelGGL(svm::k_neighbors<false>, grid, dim3(tb), 0, st, (const double*)P, n_points, c.g,
                               (const int*)c.start, (const int*)c.list, range, 0, cnt, maxc, (int*)nullptr,
                               (double*)nullptr, (double*)nullptr);
            if (hipGetLastError() != hipSuccess ||
                hipMemcpyAsync(counts, cnt, sizeof(int) * n, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipMemcpyAsync(&hmax, maxc, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                r = SVM_ERR_HIP;
        }
        if (!r) {
            *needed = hmax;
            if (hmax > max_neighbors) {
                r = SVM_ERR_CAPACITY;   // the reference's recompute with a larger Nmax (neighborList.cpp:150-153)
            } else if (max_neighbors > 0) {
                const size_t slots = n * (size_t)max_neighbors;
                int* idx = m.get<int>(slots);
                double* vec = m.get<double>(3 * slots);
                double* dst = m.get<double>(slots);
                if (!idx || !vec || !dst || hipMemsetAsync(idx, 0, sizeof(int) * slots, st) != hipSuccess ||
                    hipMemsetAsync(vec, 0, sizeof(double) * 3 * slots, st) != hipSuccess ||
                    hipMemsetAsync(dst, 0, sizeof(double) * slots, st) != hipSuccess) {
                    r = SVM_ERR_HIP;
                } else {
                    hipLaunchKernelGGL(svm::k_neighbors<true>, grid, dim3(tb), 0, st, (const double*)P, n_points,
                                       c.g, (const int*)c.start, (const int*)c.list, range, max_neighbors, cnt,
                                       maxc, idx, vec, dst);
                    if (hipGetLastError() != hipSuccess ||
                        hipMemcpyAsync(indices, idx, sizeof(int) * slots, hipMemcpyDeviceToHost, st) != hipSuccess ||
                        hipMemcpyAsync(vectors, vec, sizeof(double) * 3 * slots, hipMemcpyDeviceToHost, st) !=
                            hipSuccess ||
                        hipMemcpyAsync(distances, dst, sizeof(double) * slots, hipMemcpyDeviceToHost, st) !=
                            hipSuccess ||
                        hipStreamSynchronize(st) != hipSuccess)
                        r = SVM_ERR_HIP;
                }
            }
        }
    }
    (void)hipStreamDestroy(st);
    return r;
}

int svm_delaunay_triangles(int32_t n_points, const double* points, int32_t* triangles) {
    if (n_points < 4 || !points || !triangles) return SVM_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return SVM_ERR_NO_DEVICE;
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return SVM_ERR_HIP;
    int r;
    {
        DevScratch m;
        DlOut o;
        std::string msg;
        r = dl_triangulate(st, m, points, n_points, 64, o, msg);
        if (!r && (hipMemcpyAsync(triangles, o.T, sizeof(int32_t) * 3 * (size_t)o.ntri, hipMemcpyDeviceToHost, st) !=
                       hipSuccess ||
                   hipStreamSynchronize(st) != hipSuccess))
            r = SVM_ERR_HIP;
    }
    (void)hipStreamDestroy(st);
    return r;
}

}  // extern "C"

// Device memory of svm_delaunay_triangles_noalloc: one array per GPU, held
// until the process ends.  Host code never uses the array's name as an
// address (in host code it is a stand-in, not device memory): the device
// address of each GPU's copy comes from hipGetSymbolAddress, once per device.
static __device__ __align__(256) unsigned char scratch[SVM_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxScratchDevices] = {};

// The device address of the current device's scratch array.
int scratch_on_current_device(void** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return SVM_ERR_HIP;
    if (dev < 0 || dev >= kMaxScratchDevices) return SVM_ERR_NO_DEVICE;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return SVM_ERR_HIP;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return SVM_OK;
}
}  // namespace

extern "C" {

int svm_delaunay_triangles_noalloc(int32_t n_points, const double* points, int32_t* triangles, void* stream) {
    if (n_points < 4 || !points || !triangles) return SVM_ERR_ARG;
    if (n_points > SVM_NOALLOC_MAX_POINTS) return SVM_ERR_CAPACITY;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return SVM_ERR_NO_DEVICE;
    void* base = nullptr;
    int r = scratch_on_current_device(&base);
    if (r) return r;
    hipStream_t st = (hipStream_t)stream;
    DevScratch m(base, SVM_NOALLOC_SCRATCH_BYTES);
    DlOut o;
    std::string msg;
    r = dl_triangulate(st, m, points, n_points, 64, o, msg);
    if (m.overflow) return SVM_ERR_CAPACITY;
    if (!r && (hipMemcpyAsync(triangles, o.T, sizeof(int32_t) * 3 * (size_t)o.ntri, hipMemcpyDeviceToHost, st) !=
                   hipSuccess ||
               hipStreamSynchronize(st) != hipSuccess))
        r = SVM_ERR_HIP;
    return r;
}

int svm_get_delaunay_triangles(svm_context* ctx, int32_t sys, int32_t* triangles) {
    if (!valid_sys(ctx, sys) || !triangles) return fail(ctx, SVM_ERR_ARG, "null argument");
    if (ctx->triangles[sys].empty()) return fail(ctx, SVM_ERR_STATE, "no Delaunay triangulation for this sphere");
    std::memcpy(triangles, ctx->triangles[sys].data(), sizeof(int32_t) * ctx->triangles[sys].size());
    return SVM_OK;
}

#define STATE_IO(NAME, X, Y, Z)                                                                                   \
    int svm_set_##NAME(svm_context* ctx, int32_t sys, const double* src) {                                         \
        if (!valid_sys(ctx, sys) || !src) return fail(ctx, SVM_ERR_ARG, "null argument");                           \
        return upload_vec3(ctx, sys, src, X, Y, Z);                                                                \
    }                                                                                                              \
    int svm_get_##NAME(svm_context* ctx, int32_t sys, double* dst) {                                               \
        if (!valid_sys(ctx, sys) || !dst) return fail(ctx, SVM_ERR_ARG, "null argument");                           \
        return download_rows(ctx, ctx->layout[sys].nv, ctx->layout[sys].vo, X, Y, Z, dst);                         \
    }

STATE_IO(positions, ctx->d.px[ctx->pcur], ctx->d.py[ctx->pcur], ctx->d.pz[ctx->pcur])
STATE_IO(directors, ctx->d.nx[ctx->ncur], ctx->d.ny[ctx->ncur], ctx->d.nz[ctx->ncur])
STATE_IO(velocities, ctx->d.vx, ctx->d.vy, ctx->d.vz)

// Caller-owned device buffers: [n][3] rows copied to or from the state's
// planes by a kernel on the context's stream.  The buffer must be device
// memory of this runtime covering `bytes` from `p`; anything else is refused
// before a kernel could touch it.
int check_device_buffer(svm_context* ctx, const void* p, size_t bytes) {
    if (!p) return fail(ctx, SVM_ERR_ARG, "null device buffer");
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
        (void)hipGetLastError();
        return fail(ctx, SVM_ERR_ARG, "not a device buffer of this process's HIP runtime");
    }
    if (attr.type != hipMemoryTypeDevice) return fail(ctx, SVM_ERR_ARG, "not device memory");
    if (attr.device != ctx->device) return fail(ctx, SVM_ERR_ARG, "device buffer on another device");
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p) != hipSuccess) {
        (void)hipGetLastError();
        return fail(ctx, SVM_ERR_ARG, "device buffer of unknown extent");
    }
    const char* b = (const char*)base;
    const char* q = (const char*)p;
    if (q < b || (size_t)(q - b) + bytes > size) return fail(ctx, SVM_ERR_ARG, "device buffer too small");
    return SVM_OK;
}

__global__ __launch_bounds__(256) void k_rows_to_aos(const double* X, const double* Y, const double* Z, int off, int n,
                                                    double* out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    out[3 * (size_t)i] = X[off + i];
    out[3 * (size_t)i + 1] = Y[off + i];
    out[3 * (size_t)i + 2] = Z[off + i];
}

__global__ __launch_bounds__(256) void k_aos_to_rows(const double* in, int off, int n, double* X, double* Y,
                                                    double* Z) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    X[off + i] = in[3 * (size_t)i];
    Y[off + i] = in[3 * (size_t)i + 1];
    Z[off + i] = in[3 * (size_t)i + 2];
}

#define STATE_IO_DEVICE(NAME, X, Y, Z)                                                                           \
    int svm_set_##NAME##_device(svm_context* ctx, int32_t sys, const double* src) {                                \
        if (!valid_sys(ctx, sys)) return fail(ctx, SVM_ERR_ARG, "bad sphere index");                               \
        const int n = ctx->layout[sys].nv, off = ctx->layout[sys].vo;                                              \
        int r = check_device_buffer(ctx, src, sizeof(double) * 3 * (size_t)n);                                     \
        if (r) return r;                                                                                           \
        hipLaunchKernelGGL(k_aos_to_rows, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, src, off, n, X, Y, Z);  \
        return last_launch(ctx);                                                                                   \
    }                                                                                                              \
    int svm_get_##NAME##_device(svm_context* ctx, int32_t sys, double* dst) {                                      \
        if (!valid_sys(ctx, sys)) return fail(ctx, SVM_ERR_ARG, "bad sphere index");                               \
        const int n = ctx->layout[sys].nv, off = ctx->layout[sys].vo;                                              \
        int r = check_device_buffer(ctx, dst, sizeof(double) * 3 * (size_t)n);                                     \
        if (r) return r;                                                                                           \
        hipLaunchKernelGGL(k_rows_to_aos, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, X, Y, Z, off, n, dst);  \
        return last_launch(ctx);                                                                                   \
    }

STATE_IO_DEVICE(positions, ctx->d.px[ctx->pcur], ctx->d.py[ctx->pcur], ctx->d.pz[ctx->pcur])
STATE_IO_DEVICE(directors, ctx->d.nx[ctx->ncur], ctx->d.ny[ctx->ncur], ctx->d.nz[ctx->ncur])
STATE_IO_DEVICE(velocities, ctx->d.vx, ctx->d.vy, ctx->d.vz)

int svm_get_forces_device(svm_context* ctx, int32_t sys, double* dst) {
    if (!valid_sys(ctx, sys)) return fail(ctx, SVM_ERR_ARG, "bad sphere index");
    const int n = ctx->layout[sys].nv, off = ctx->layout[sys].vo;
    int r = check_device_buffer(ctx, dst, sizeof(double) * 3 * (size_t)n);
    if (r) return r;
    hipLaunchKernelGGL(k_rows_to_aos, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, ctx->d.fx, ctx->d.fy,
                       ctx->d.fz, off, n, dst);
    return last_launch(ctx);
}

int svm_get_forces(svm_context* ctx, int32_t sys, double* forces) {
    if (!valid_sys(ctx, sys) || !forces) return fail(ctx, SVM_ERR_ARG, "null argument");
    return download_rows(ctx, ctx->layout[sys].nv, ctx->layout[sys].vo, ctx->d.fx, ctx->d.fy, ctx->d.fz, forces);
}

int svm_get_cell_positions(svm_context* ctx, int32_t sys, double* out) {
    if (!valid_sys(ctx, sys) || !out) return fail(ctx, SVM_ERR_ARG, "null argument");
    return download_rows(ctx, ctx->layout[sys].nc, ctx->layout[sys].co, ctx->d.cx, ctx->d.cy, ctx->d.cz, out);
}

int svm_get_area_perimeter(svm_context* ctx, int32_t sys, double* out) {
    if (!valid_sys(ctx, sys) || !out) return fail(ctx, SVM_ERR_ARG, "null argument");
    const SysParams& L = ctx->layout[sys];
    std::vector<double> a(L.nc), p(L.nc);
    HIPCHECK(ctx, hipMemcpyAsync(a.data(), ctx->d.A + L.co, sizeof(double) * L.nc, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(ctx, hipMemcpyAsync(p.data(), ctx->d.P + L.co, sizeof(double) * L.nc, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    for (int c = 0; c < L.nc; ++c) { out[2 * (size_t)c] = a[c]; out[2 * (size_t)c + 1] = p[c]; }
    return SVM_OK;
}

int svm_set_preferences(svm_context* ctx, int32_t sys, const double* prefs) {
    if (!valid_sys(ctx, sys) || !prefs) return fail(ctx, SVM_ERR_ARG, "null argument");
    const SysParams& L = ctx->layout[sys];
    std::vector<double> a(L.nc), p(L.nc);
    for (int c = 0; c < L.nc; ++c) { a[c] = prefs[2 * (size_t)c]; p[c] = prefs[2 * (size_t)c + 1]; }
    HIPCHECK(ctx, hipMemcpyAsync(ctx->d.a0 + L.co, a.data(), sizeof(double) * L.nc, hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(ctx, hipMemcpyAsync(ctx->d.p0 + L.co, p.data(), sizeof(double) * L.nc, hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    return SVM_OK;
}

int svm_set_preferred_parameters(svm_context* ctx, int32_t sys, double a0, double p0) {
    if (!valid_sys(ctx, sys)) return fail(ctx, SVM_ERR_ARG, "bad sphere index");
    std::vector<double> pr(2 * (size_t)ctx->layout[sys].nc);
    for (int c = 0; c < ctx->layout[sys].nc; ++c) { pr[2 * (size_t)c] = a0; pr[2 * (size_t)c + 1] = p0; }
    return svm_set_preferences(ctx, sys, pr.data());
}

int svm_set_scalar_model_parameter(svm_context* ctx, int32_t sys, double kr) {
    if (!valid_sys(ctx, sys)) return fail(ctx, SVM_ERR_ARG, "bad sphere index");
    return set_field(ctx, sys, offsetof(SysParams, kr), kr);
}

int svm_set_t1_threshold(svm_context* ctx, int32_t sys, double threshold) {
    if (!valid_sys(ctx, sys)) return fail(ctx, SVM_ERR_ARG, "bad sphere index");
    if (!(threshold >= 0.0)) return fail(ctx, SVM_ERR_ARG, "the T1 threshold must be >= 0");
    const double R = std::sqrt(ctx->layout[sys].nc / (4.0 * SVM_PI));
    int r = set_field(ctx, sys, offsetof(SysParams, thr), threshold);
    return r ? r : set_field(ctx, sys, offsetof(SysParams, thr_cos), svm::t1_cos_cut(threshold, R));
}

int svm_compute_geometry(svm_context* ctx) {
    if (!ctx) return SVM_ERR_ARG;
    int r = require_topology(ctx);
    if (r) return r;
    return launch_geometry(ctx, 0);
}

int svm_compute_forces(svm_context* ctx, int32_t precision) {
    if (!ctx) return SVM_ERR_ARG;
    int r = require_topology(ctx);
    if (r) return r;
    if (precision < 0 || precision > 2)
        return fail(ctx, SVM_ERR_ARG, "precision must be 0 (FP64 reference expressions), 1 (FP32) or 2 (FP64 vector form)");
    r = launch_geometry(ctx, 0);
    if (r) return r;
    const int keep = ctx->force_fp32;
    ctx->force_fp32 = precision;
    r = launch_vertex<svm::VM_FORCE>(ctx, ctx->pcur, 1, "k_vertex_force");
    ctx->force_fp32 = keep;
    return r;
}

int svm_set_force_precision(svm_context* ctx, int32_t precision) {
    if (!ctx || precision < 0 || precision > 2)
        return fail(ctx, SVM_ERR_ARG, "precision must be 0 (FP64 reference expressions), 1 (FP32) or 2 (FP64 vector form)");
    if (ctx->force_fp32 != precision) {
        ctx->force_fp32 = precision;
        drop_graphs(ctx);
    }
    return SVM_OK;
}

int svm_get_force_error(svm_context* ctx, double* stats) {
    if (!ctx || !stats) return SVM_ERR_ARG;
    // forces of every sphere as AoS rows: FP64 reference expressions, FP64
    // vector form, then FP32 (left in the force buffer)
    std::vector<double> f64(3 * (size_t)ctx->Vs), fv(3 * (size_t)ctx->Vs), f32(3 * (size_t)ctx->Vs);
    int r = svm_compute_forces(ctx, 0);
    if (!r) r = download_rows(ctx, ctx->Vs, 0, ctx->d.fx, ctx->d.fy, ctx->d.fz, f64.data());
    if (!r) r = svm_compute_forces(ctx, 2);
    if (!r) r = download_rows(ctx, ctx->Vs, 0, ctx->d.fx, ctx->d.fy, ctx->d.fz, fv.data());
    if (!r) r = svm_compute_forces(ctx, 1);
    if (!r) r = download_rows(ctx, ctx->Vs, 0, ctx->d.fx, ctx->d.fy, ctx->d.fz, f32.data());
    if (r) return r;
    for (int s = 0; s < ctx->nsys; ++s) {
        const SysParams& L = ctx->layout[s];
        double maxf = 0.0, sf = 0.0, se = 0.0, maxe = 0.0, sev = 0.0, maxev = 0.0, ser = 0.0;
        for (int v = L.vo; v < L.vo + L.nv; ++v) {
            double a = 0.0, e = 0.0, ev = 0.0, er = 0.0;
            for (int k = 0; k < 3; ++k) {
                const double x = f64[3 * (size_t)v + k], y = f32[3 * (size_t)v + k], z = fv[3 * (size_t)v + k];
                a += x * x;
                e += (y - x) * (y - x);
                ev += (y - z) * (y - z);
                er += (z - x) * (z - x);
            }
            sf += a;
            se += e;
            sev += ev;
            ser += er;
            maxf = std::max(maxf, std::sqrt(a));
            maxe = std::max(maxe, std::sqrt(e));
            maxev = std::max(maxev, std::sqrt(ev));
        }
        double* o = stats + SVM_FORCE_ERR_COLS * (size_t)s;
        o[0] = maxf > 0 ? maxe / maxf : 0.0;          // FP32 vs FP64 reference: max |dF| / max |F|
        o[1] = sf > 0 ? std::sqrt(se / sf) : 0.0;      //   rms |dF| / rms |F|
        o[2] = maxe;                                   //   max |dF|
        o[3] = std::sqrt(sf / L.nv);                   // rms |F| (FP64)
        o[4] = maxf > 0 ? maxev / maxf : 0.0;         // FP32 vs FP64 vector form: max, rms (relative)
        o[5] = sf > 0 ? std::sqrt(sev / sf) : 0.0;
        o[6] = sf > 0 ? std::sqrt(ser / sf) : 0.0;     // FP64 vector form vs FP64 reference, rms
    }
    return SVM_OK;
}

int svm_move_particles(svm_context* ctx, int32_t sys, const double* displacements, double scale) {
    if (!valid_sys(ctx, sys) || !displacements) return fail(ctx, SVM_ERR_ARG, "null argument");
    int r = require_topology(ctx);
    if (r) return r;
    const SysParams& L = ctx->layout[sys];
    HIPCHECK(ctx, hipMemcpyAsync(ctx->d.disp + 3 * (size_t)L.vo, displacements, sizeof(double) * 3 * L.nv,
                                 hipMemcpyHostToDevice, ctx->stream));
    int pout = 1 - ctx->pcur;
    hipLaunchKernelGGL(svm::k_move_ext, dim3(ctx->nvt), dim3(SVM_TILE), 0, ctx->stream, ctx->d, ctx->pcur, pout, sys,
                       scale);
    r = last_launch(ctx);
    if (r) return r;
    r = launch_t1(ctx, pout, ctx->ncur, ctx->ncur, 0, 0, 0);
    if (r) return r;
    ctx->pcur = pout;
    return svm_synchronize(ctx);
}

int svm_enforce_topology(svm_context* ctx) {
    if (!ctx) return SVM_ERR_ARG;
    int r = require_topology(ctx);
    if (r) return r;
    return launch_t1(ctx, ctx->pcur, ctx->ncur, ctx->ncur, 0, 0, 0);
}

int svm_get_energies(svm_context* ctx, double* potential, double* kinetic) {
    if (!ctx) return SVM_ERR_ARG;
    hipLaunchKernelGGL(svm::k_energy_cells, dim3(ctx->nct), dim3(SVM_TILE), 0, ctx->stream, ctx->d);
    int r = last_launch(ctx);
    if (r) return r;
    hipLaunchKernelGGL(svm::k_energy_vertices, dim3(ctx->nvt), dim3(SVM_TILE), 0, ctx->stream, ctx->d);
    r = last_launch(ctx);
    if (r) return r;
    hipLaunchKernelGGL(svm::k_energy_fold, dim3(ctx->nsys), dim3(SVM_TILE), 0, ctx->stream, ctx->d);
    r = last_launch(ctx);
    if (r) return r;
    std::vector<SysParams> s;
    r = download_sys(ctx, s);
    if (r) return r;
    for (int i = 0; i < ctx->nsys; ++i) {
        if (potential) potential[i] = s[i].energy_pot;
        if (kinetic) kinetic[i] = s[i].energy_kin;
    }
    return SVM_OK;
}

int svm_get_counters(svm_context* ctx, int64_t* counters) {
    if (!ctx || !counters) return SVM_ERR_ARG;
    std::vector<SysParams> s;
    int r = download_sys(ctx, s);
    if (r) return r;
    for (int i = 0; i < ctx->nsys; ++i) {
        int64_t* c = counters + SVM_COUNTER_COLS * (size_t)i;
        c[0] = s[i].n_t1;
        c[1] = s[i].n_fan;
        c[2] = s[i].n_nan;
        c[3] = s[i].n_t1_skipped;
        c[4] = s[i].n_t1_sequential;
        c[5] = s[i].n_t1_group;
    }
    return SVM_OK;
}

int svm_set_brownian(svm_context* ctx, int32_t sys, double dt, double mu, double temperature) {
    if (!valid_sys(ctx, sys) || temperature < 0) return fail(ctx, SVM_ERR_ARG, "bad argument");
    int r = set_field(ctx, sys, offsetof(SysParams, dt), dt);
    r = r ? r : set_field(ctx, sys, offsetof(SysParams, mu), mu);
    return r ? r : set_field(ctx, sys, offsetof(SysParams, T), temperature);
}

int svm_set_active(svm_context* ctx, int32_t sys, double dt, double mu, double v0, double tau) {
    if (!valid_sys(ctx, sys)) return fail(ctx, SVM_ERR_ARG, "bad argument");
    int r = set_field(ctx, sys, offsetof(SysParams, dt), dt);
    r = r ? r : set_field(ctx, sys, offsetof(SysParams, mu), mu);
    r = r ? r : set_field(ctx, sys, offsetof(SysParams, v0), v0);
    return r ? r : set_field(ctx, sys, offsetof(SysParams, tau), tau);
}

int svm_set_vicsek(svm_context* ctx, int32_t sys, double dt, double mu, double v0, double tau, double eta) {
    if (!valid_sys(ctx, sys) || !(tau != 0.0)) return fail(ctx, SVM_ERR_ARG, "bad argument");
    int r = set_field(ctx, sys, offsetof(SysParams, dt), dt);
    r = r ? r : set_field(ctx, sys, offsetof(SysParams, mu), mu);
    r = r ? r : set_field(ctx, sys, offsetof(SysParams, v0), v0);
    r = r ? r : set_field(ctx, sys, offsetof(SysParams, tau), tau);
    return r ? r : set_field(ctx, sys, offsetof(SysParams, eta), eta);
}

static int run_steps(svm_context* ctx, int updater, int n_steps) {
    if (!ctx || n_steps < 0) return SVM_ERR_ARG;
    int r = require_topology(ctx);
    if (r) return r;
    return iterate(ctx, updater, n_steps, 1, false, [&](const Target& w) { return enqueue_step(ctx, w, updater); });
}

int svm_brownian_steps(svm_context* ctx, int32_t n_steps) { return run_steps(ctx, 0, n_steps); }
int svm_active_steps(svm_context* ctx, int32_t n_steps) { return run_steps(ctx, 1, n_steps); }
int svm_vicsek_steps(svm_context* ctx, int32_t n_steps) { return run_steps(ctx, 2, n_steps); }

int svm_run_graph(svm_context* ctx, int32_t updater, int32_t steps_per_graph, int32_t n_replays) {
    if (ctx && steps_per_graph == 0) steps_per_graph = ctx->graph_steps;
    if (!ctx || updater < 0 || updater > 2 || steps_per_graph <= 0 || (steps_per_graph & 1) || n_replays < 0)
        return fail(ctx, SVM_ERR_ARG, "steps_per_graph must be a positive even number");
    int r = require_topology(ctx);
    if (r) return r;
    return iterate(ctx, updater, steps_per_graph, n_replays, true,
                   [&](const Target& w) { return enqueue_step(ctx, w, updater); });
}

int svm_fire_set_parameters(svm_context* ctx, int32_t sys, const svm_fire_params* p) {
    if (!valid_sys(ctx, sys) || !p) return fail(ctx, SVM_ERR_ARG, "bad argument");
    std::vector<SysParams> s;
    int r = download_sys(ctx, s);
    if (r) return r;
    SysParams& S = s[sys];
    S.f_dt0 = p->dt;
    S.f_dt = p->dt;
    S.f_dt_min = p->dt * .01;
    S.f_alpha_start = p->alpha_start;
    S.f_alpha = p->alpha_start;
    S.f_dt_max = p->dt_max;
    S.f_dt_inc = p->dt_inc;
    S.f_dt_dec = p->dt_dec;
    S.f_alpha_dec = p->alpha_dec;
    S.f_nmin = p->n_min;
    S.f_cutoff = p->force_cutoff;
    S.f_alpha_min = p->alpha_min;
    S.f_max_iter = p->max_iterations;
    S.f_stop_rms = p->stop_rms;
    S.f_iter = 0;
    S.f_nsince = 0;
    S.f_power = 0;
    S.f_force_max = 100.;
    S.mix_pending = 0;
    S.f_frozen = 0;   // a minimiser set up afresh has not stopped
    HIPCHECK(ctx, hipMemcpyAsync(ctx->d.sys + sys, &S, sizeof(SysParams), hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    return SVM_OK;
}

int svm_fire_set_max_iterations(svm_context* ctx, int32_t sys, int32_t max_iterations) {
    if (!valid_sys(ctx, sys) || max_iterations < 0) return fail(ctx, SVM_ERR_ARG, "bad argument");
    return set_field(ctx, sys, offsetof(SysParams, f_max_iter), (int32_t)max_iterations);
}

int svm_fire_minimize(svm_context* ctx, int32_t check_every) {
    if (!ctx) return SVM_ERR_ARG;
    int r = require_topology(ctx);
    if (r) return r;
    if (check_every <= 0) check_every = 50;
    // minimize(): forces at the start, forceMax = 110, then iterate while
    // (iterations < max && forceMax > cutoff) (energyMinimizerFIRE.cpp:239-249)
    std::vector<SysParams> s;
    r = download_sys(ctx, s);
    if (r) return r;
    for (int i = 0; i < ctx->nsys; ++i) {
        s[i].f_force_max = 110.0;
        s[i].f_frozen = !((s[i].f_iter < s[i].f_max_iter) && (110.0 > s[i].f_cutoff));
        s[i].mix_pending = 0;
        r = set_field(ctx, i, offsetof(SysParams, f_force_max), s[i].f_force_max);
        r = r ? r : set_field(ctx, i, offsetof(SysParams, f_frozen), s[i].f_frozen);
        r = r ? r : set_field(ctx, i, offsetof(SysParams, mix_pending), 0);
        if (r) return r;
    }
    r = launch_geometry(ctx, 0);
    if (r) return r;
    r = launch_vertex<svm::VM_FORCE>(ctx, ctx->pcur, 1, "k_vertex_force");
    if (r) return r;
    while (true) {
        bool any = false;
        for (int i = 0; i < ctx->nsys; ++i) any |= !s[i].f_frozen;
        if (!any) break;
        r = iterate(ctx, 3, check_every, 1, true, [&](const Target& w) { return enqueue_fire_iteration(ctx, w); });
        if (r) return r;
        r = download_sys(ctx, s);
        if (r) return r;
        r = check_device_errors(ctx);
        if (r) return r;
    }
    hipLaunchKernelGGL(svm::k_fire_finalize, dim3(ctx->nvt), dim3(SVM_TILE), 0, ctx->stream, ctx->d);
    r = last_launch(ctx);
    if (r) return r;
    for (int i = 0; i < ctx->nsys; ++i) {
        r = set_field(ctx, i, offsetof(SysParams, mix_pending), 0);
        if (r) return r;
    }
    return svm_synchronize(ctx);
}

int svm_adam_set_parameters(svm_context* ctx, int32_t sys, const svm_adam_params* p) {
    if (!valid_sys(ctx, sys) || !p || !(p->beta1 < 1.0) || !(p->beta2 < 1.0) || p->max_iterations < 0)
        return fail(ctx, SVM_ERR_ARG, "bad argument");
    Dev& d = ctx->d;
    const size_t Vs = ctx->Vs;
    if (!d.am) {
        int r = dalloc(ctx, &d.am, 3 * Vs);
        if (!r) r = dalloc(ctx, &d.av, 3 * Vs);
        if (r) return r;
    }
    std::vector<SysParams> s;
    int r = download_sys(ctx, s);
    if (r) return r;
    SysParams& S = s[sys];
    S.a_beta1 = p->beta1;
    S.a_beta2 = p->beta2;
    S.a_eps = p->epsilon;
    S.a_dt = p->dt;
    S.a_cutoff = p->force_cutoff;
    S.a_max_iter = p->max_iterations;
    S.a_beta1t = p->beta1;
    S.a_beta2t = p->beta2;
    S.a_iter = 0;
    S.f_frozen = 0;   // a minimiser set up afresh has not stopped
    S.a_force_max = 110.0;
    const SysParams& L = ctx->layout[sys];
    for (int k = 0; k < 3; ++k) {
        HIPCHECK(ctx, hipMemsetAsync(d.am + k * Vs + L.vo, 0, sizeof(double) * L.nv, ctx->stream));
        HIPCHECK(ctx, hipMemsetAsync(d.av + k * Vs + L.vo, 0, sizeof(double) * L.nv, ctx->stream));
    }
    HIPCHECK(ctx, hipMemcpyAsync(d.sys + sys, &S, sizeof(SysParams), hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->have_adam[sys] = 1;
    return SVM_OK;
}

int svm_adam_minimize(svm_context* ctx, int32_t check_every) {
    if (!ctx) return SVM_ERR_ARG;
    int r = require_topology(ctx);
    if (r) return r;
    for (int i = 0; i < ctx->nsys; ++i)
        if (!ctx->have_adam[i]) return fail(ctx, SVM_ERR_STATE, "sphere " + std::to_string(i) + " has no Adam parameters");
    if (check_every <= 0) check_every = 50;
    // minimize(): forceMax = 110, then iterate while (iterations < max &&
    // forceMax > cutoff) (energyMinimizerAdam.cpp:87-98)
    std::vector<SysParams> s;
    r = download_sys(ctx, s);
    if (r) return r;
    for (int i = 0; i < ctx->nsys; ++i) {
        s[i].a_force_max = 110.0;
        s[i].f_frozen = !((s[i].a_iter < s[i].a_max_iter) && (110.0 > s[i].a_cutoff));
        r = set_field(ctx, i, offsetof(SysParams, a_force_max), s[i].a_force_max);
        r = r ? r : set_field(ctx, i, offsetof(SysParams, f_frozen), s[i].f_frozen);
        if (r) return r;
    }
    while (true) {
        bool any = false;
        for (int i = 0; i < ctx->nsys; ++i) any |= !s[i].f_frozen;
        if (!any) break;
        r = iterate(ctx, 4, check_every, 1, true, [&](const Target& w) { return enqueue_adam_iteration(ctx, w); });
        if (r) return r;
        r = download_sys(ctx, s);
        if (r) return r;
        r = check_device_errors(ctx);
        if (r) return r;
    }
    return svm_synchronize(ctx);
}

int svm_adam_set_max_iterations(svm_context* ctx, int32_t sys, int32_t max_iterations) {
    if (!valid_sys(ctx, sys) || max_iterations < 0) return fail(ctx, SVM_ERR_ARG, "bad argument");
    return set_field(ctx, sys, offsetof(SysParams, a_max_iter), (int32_t)max_iterations);
}

int svm_adam_get_state(svm_context* ctx, int32_t sys, svm_adam_state* st) {
    if (!valid_sys(ctx, sys) || !st) return fail(ctx, SVM_ERR_ARG, "bad argument");
    std::vector<SysParams> s;
    int r = download_sys(ctx, s);
    if (r) return r;
    const SysParams& S = s[sys];
    st->iterations = S.a_iter;
    st->converged = S.a_force_max <= S.a_cutoff;
    st->force_max = S.a_force_max;
    st->beta1t = S.a_beta1t;
    st->beta2t = S.a_beta2t;
    return SVM_OK;
}

int svm_fire_get_state(svm_context* ctx, int32_t sys, svm_fire_state* st) {
    if (!valid_sys(ctx, sys) || !st) return fail(ctx, SVM_ERR_ARG, "bad argument");
    std::vector<SysParams> s;
    int r = download_sys(ctx, s);
    if (r) return r;
    const SysParams& S = s[sys];
    st->iterations = S.f_iter;
    st->n_since_negative_power = S.f_nsince;
    st->converged = S.f_force_max <= S.f_cutoff;
    st->reserved = 0;
    st->dt = S.f_dt;
    st->alpha = S.f_alpha;
    st->force_max = S.f_force_max;
    st->power = S.f_power;
    return SVM_OK;
}

int svm_get_t1_diagnostics(svm_context* ctx, int64_t* reasons) {
    if (!ctx || !reasons) return SVM_ERR_ARG;
    std::vector<SysParams> s;
    int r = download_sys(ctx, s);
    if (r) return r;
    for (int i = 0; i < ctx->nsys; ++i) {
        int64_t* r = reasons + SVM_T1_DIAG_COLS * (size_t)i;
        for (int k = 0; k < 4; ++k) r[k] = s[i].n_seq_reason[k];
        r[4] = s[i].n_cand_flip;
        r[5] = s[i].n_cand_block;
        r[6] = s[i].n_grp_members;
        r[7] = s[i].n_grp_attempts;
        r[8] = s[i].n_grp_visits;
        r[9] = s[i].max_grp_visits;
        r[10] = s[i].n_cl_visits;
        r[11] = s[i].max_cl_visits;
    }
    return SVM_OK;
}

int svm_set_t1_mode(svm_context* ctx, int32_t mode) {
    if (!ctx || (mode != 0 && mode != 1)) return fail(ctx, SVM_ERR_ARG, "T1 mode must be 0 or 1");
    ctx->t1_sequential = mode;
    drop_graphs(ctx);
    return SVM_OK;
}

int svm_set_t1_component_distance(svm_context* ctx, double cells) {
    if (!ctx || !(cells >= 0.0 && cells <= 64.0))
        return fail(ctx, SVM_ERR_ARG, "component distance must be in [0, 64] cell sizes");
    ctx->d.t1_near = cells;
    drop_graphs(ctx);
    return SVM_OK;
}
int svm_get_t1_component_distance(svm_context* ctx, double* cells) {
    if (!ctx || !cells) return fail(ctx, SVM_ERR_ARG, "null argument");
    *cells = ctx->d.t1_near;
    return SVM_OK;
}

int svm_set_profiling(svm_context* ctx, int32_t on) {
    if (!ctx) return SVM_ERR_ARG;
    ctx->profiling = on != 0;
    ctx->timing.clear();
    return SVM_OK;
}

int svm_get_timing(svm_context* ctx, const char** names, double* ms, int32_t n, int32_t* n_written) {
    if (!ctx || !n_written) return SVM_ERR_ARG;
    int k = 0;
    ctx->timing_names.clear();
    for (auto& e : ctx->timing) ctx->timing_names.push_back(e.first);
    for (auto& e : ctx->timing) {
        if (k >= n) break;
        if (names) names[k] = ctx->timing_names[k].c_str();
        if (ms) ms[k] = e.second.second ? e.second.first / e.second.second : 0.0;
        ++k;
    }
    *n_written = k;
    return SVM_OK;
}

int svm_time_vertex_pass(svm_context* ctx, int32_t updater, int32_t reps, double* ms_per_launch) {
    if (!ctx || (updater != 0 && updater != 1) || reps < 1 || !ms_per_launch) return SVM_ERR_ARG;
    int r = require_topology(ctx);
    if (r) return r;
    std::vector<SysParams> saved;
    r = download_sys(ctx, saved);
    if (r) return r;
    const int was_profiling = ctx->profiling;
    ctx->profiling = 0;
    r = launch_geometry(ctx, 0);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!r && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) r = SVM_ERR_HIP;
    const int pout = 1 - ctx->pcur;
    if (!r) {
        (void)hipEventRecord(e0, ctx->stream);
        for (int i = 0; i < reps && !r; ++i)
            r = updater == 0 ? launch_vertex<svm::VM_BROWNIAN>(ctx, pout, 0, "k_vertex_brownian")
                             : launch_vertex<svm::VM_ACTIVE>(ctx, pout, 0, "k_vertex_active");
        (void)hipEventRecord(e1, ctx->stream);
        float ms = 0.f;
        if (!r && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            r = SVM_ERR_HIP;
        *ms_per_launch = (double)ms / reps;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    ctx->profiling = was_profiling;
    // the pass is a pure function of the current state; only its counters moved
    if (hipMemcpyAsync(ctx->d.sys, saved.data(), sizeof(SysParams) * ctx->nsys, hipMemcpyHostToDevice,
                       ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        r = r ? r : SVM_ERR_HIP;
    return r;
}

int svm_set_tuning(svm_context* ctx, int32_t cluster_blocks, int32_t graph_steps, int32_t xcd_mapping) {
    if (!ctx || cluster_blocks < 1 || cluster_blocks > 65536 || graph_steps < 2 || (graph_steps & 1) ||
        (xcd_mapping != 0 && xcd_mapping != 1))
        return fail(ctx, SVM_ERR_ARG, "bad tuning values");
    ctx->cluster_blocks = cluster_blocks;
    ctx->graph_steps = graph_steps;
    ctx->d.xcd = xcd_mapping;
    drop_graphs(ctx);
    return SVM_OK;
}

int svm_set_vertex_form(svm_context* ctx, int32_t form) {
    if (!ctx || form < 0 || form > 4) return fail(ctx, SVM_ERR_ARG, "vertex form must be in [0, 4]");
    ctx->vertex_corner = form;
    drop_graphs(ctx);
    return SVM_OK;
}

int svm_get_vertex_form(svm_context* ctx, int32_t* form) {
    if (!ctx || !form) return SVM_ERR_ARG;
    *form = ctx->vertex_corner;
    return SVM_OK;
}

int svm_set_geometry_form(svm_context* ctx, int32_t form) {
    if (!ctx || form < 0 || form > 2) return fail(ctx, SVM_ERR_ARG, "geometry form must be 0, 1 or 2");
    ctx->geometry_form = form;
    drop_graphs(ctx);
    return SVM_OK;
}

int svm_get_geometry_form(svm_context* ctx, int32_t* form) {
    if (!ctx || !form) return SVM_ERR_ARG;
    *form = ctx->geometry_form;
    return SVM_OK;
}

int svm_set_step_form(svm_context* ctx, int32_t form) {
    if (!ctx || form < 0 || form > 3) return fail(ctx, SVM_ERR_ARG, "step form must be in [0, 3]");
    ctx->step_form = form;
    drop_graphs(ctx);
    return SVM_OK;
}

int svm_get_step_form(svm_context* ctx, int32_t* form) {
    if (!ctx || !form) return SVM_ERR_ARG;
    *form = ctx->step_form;
    return SVM_OK;
}

// Lanes over the sphere ranges [s0[k], s0[k+1]): streams, events and T1
// buffers for every lane above the first (kept until svm_destroy).
static int apply_lanes(svm_context* ctx, const std::vector<int>& s0) {
    const int n = (int)s0.size() - 1;
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (!ctx->lane_fork) HIPCHECK(ctx, hipEventCreateWithFlags(&ctx->lane_fork, hipEventDisableTiming));
    while ((int)ctx->lane_streams.size() < n - 1) {
        // default priority: every stream of the library has the same one
        hipStream_t st = nullptr;
        hipEvent_t ev = nullptr;
        HIPCHECK(ctx, hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        ctx->lane_streams.push_back(st);
        HIPCHECK(ctx, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        ctx->lane_joins.push_back(ev);
        Dev L{};
        int r = 0;
        r |= dalloc(ctx, &L.t1_count, 1);
        r |= dalloc(ctx, &L.t1_list, SVM_T1_MAXCL);
        r |= dalloc(ctx, &L.t1_flagged, SVM_T1_MAXCL);
        r |= dalloc(ctx, &L.t1_logged, SVM_T1_MAXCL);
        r |= dalloc(ctx, &L.t1_lead, SVM_T1_MAXCL);
        r |= dalloc(ctx, &L.t1_total, 1);
        r |= dalloc(ctx, &L.t1_btotal, 1);
        r |= dalloc(ctx, &L.t1_bcount, 1);
        r |= dalloc(ctx, &L.t1_blist, SVM_T1_MAXBL);
        r |= dalloc(ctx, &L.t1_bflag, SVM_T1_MAXBL);
        r |= dalloc(ctx, &L.t1_bskip, SVM_T1_MAXBL);
        r |= dalloc(ctx, &L.t1_log, SVM_T1_MAXCL);
        if (r) return fail(ctx, SVM_ERR_HIP, "lane buffers: out of device memory");
        ctx->lane_t1.push_back(L);
    }
    ctx->lane_s0 = s0;
    ctx->nlanes = n;
    drop_graphs(ctx);
    return SVM_OK;
}

int svm_set_lanes(svm_context* ctx, int32_t n_lanes) {
    if (!ctx || n_lanes < 1 || n_lanes > SVM_MAX_LANES)
        return fail(ctx, SVM_ERR_ARG, "n_lanes must be in [1, " + std::to_string(SVM_MAX_LANES) + "]");
    const int n = std::min((int)n_lanes, ctx->nsys);
    // contiguous sphere ranges with about equal numbers of vertex tiles
    std::vector<int> s0(1, 0);
    long long acc = 0;
    for (int s = 0; s < ctx->nsys && (int)s0.size() < n; ++s) {
        acc += ctx->layout[s].nvtiles;
        const int k = (int)s0.size();   // lane whose start is next
        const bool room = ctx->nsys - (s + 1) >= n - k;
        if (room && acc * n >= (long long)ctx->nvt * k) s0.push_back(s + 1);
    }
    while ((int)s0.size() < n) s0.push_back(ctx->nsys - (n - (int)s0.size()));
    s0.push_back(ctx->nsys);
    return apply_lanes(ctx, s0);
}

int svm_set_lane_spheres(svm_context* ctx, int32_t n_lanes, const int32_t* first_sphere) {
    if (!ctx || n_lanes < 1 || n_lanes > SVM_MAX_LANES || n_lanes > ctx->nsys || !first_sphere)
        return fail(ctx, SVM_ERR_ARG, "n_lanes must be in [1, " + std::to_string(SVM_MAX_LANES) +
                                          "] and at most the number of spheres, first_sphere non-null");
    std::vector<int> s0(first_sphere, first_sphere + n_lanes);
    if (s0[0] != 0) return fail(ctx, SVM_ERR_ARG, "the first lane must start at sphere 0");
    for (int k = 1; k < n_lanes; ++k)
        if (s0[k] <= s0[k - 1] || s0[k] >= ctx->nsys)
            return fail(ctx, SVM_ERR_ARG, "lane starts must increase strictly and stay below the sphere count");
    s0.push_back(ctx->nsys);
    return apply_lanes(ctx, s0);
}

int svm_get_lanes(svm_context* ctx, int32_t* n_lanes, int32_t* first_sphere) {
    if (!ctx || !n_lanes) return SVM_ERR_ARG;
    *n_lanes = ctx->nlanes;
    if (first_sphere)
        for (int k = 0; k < ctx->nlanes; ++k) first_sphere[k] = ctx->nlanes == 1 ? 0 : ctx->lane_s0[k];
    return SVM_OK;
}

int svm_get_tuning(svm_context* ctx, int32_t* cluster_blocks, int32_t* graph_steps, int32_t* xcd_mapping) {
    if (!ctx) return SVM_ERR_ARG;
    if (cluster_blocks) *cluster_blocks = ctx->cluster_blocks;
    if (graph_steps) *graph_steps = ctx->graph_steps;
    if (xcd_mapping) *xcd_mapping = ctx->d.xcd;
    return SVM_OK;
}

int svm_autotune(svm_context* ctx, int32_t updater, int32_t steps_per_trial, double* table, int32_t table_rows,
                 int32_t* rows_written) {
    if (!ctx || (updater != 0 && updater != 1) || steps_per_trial < 6) return fail(ctx, SVM_ERR_ARG, "bad argument");
    int r = require_topology(ctx);
    if (r) return r;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return fail(ctx, SVM_ERR_HIP, "event creation failed");
    int nrows = 0;
    // ms per step of `steps` steps under the current settings: median of three
    // samples (kernelTuner's median-of-samples, src/utilities/kernelTuner.cpp:35-126)
    auto trial = [&](int G, double* ms) -> int {
        const int reps = std::max(1, steps_per_trial / (3 * G));
        double samples[3];
        int rr = svm_run_graph(ctx, updater, G, 1);   // capture + one warm replay
        if (rr) return rr;
        for (int k = 0; k < 3; ++k) {
            (void)hipEventRecord(e0, ctx->stream);
            rr = svm_run_graph(ctx, updater, G, reps);
            if (rr) return rr;
            (void)hipEventRecord(e1, ctx->stream);
            float t = 0.f;
            if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess)
                return SVM_ERR_HIP;
            samples[k] = (double)t / (reps * G);
        }
        std::sort(samples, samples + 3);
        *ms = samples[1];
        return SVM_OK;
    };
    int best_cb = ctx->cluster_blocks, best_G = ctx->graph_steps, best_x = ctx->d.xcd, best_l = ctx->nlanes;
    int best_f = 0, best_s = 0, best_g = 0;
    auto record = [&](int lanes, int cb, int G, int xcd, int form, int sform, int gform, double ms) {
        if (table && nrows < table_rows) {
            double* row = table + (size_t)SVM_TUNE_COLS * nrows;
            row[0] = lanes;
            row[1] = cb;
            row[2] = G;
            row[3] = xcd;
            row[4] = form;
            row[5] = sform;
            row[6] = gform;
            row[7] = ms;
        }
        ++nrows;
    };
    double best = 1e300;
    r = svm_set_vertex_form(ctx, 0);
    if (!r) r = svm_set_step_form(ctx, 0);
    if (!r) r = svm_set_geometry_form(ctx, 0);
    // lanes x two dry-run grids, at 10 steps per graph with the XCD mapping
    const int cb_pair[] = {256, 1024};
    for (int l = 1; l <= std::min(ctx->nsys, (int)SVM_MAX_LANES) && !r; ++l)
        for (int cb : cb_pair) {
            r = svm_set_lanes(ctx, l);
            if (!r) r = svm_set_tuning(ctx, cb, 10, 1);
            double ms = 0.0;
            if (!r) r = trial(10, &ms);
            if (r) break;
            record(l, cb, 10, 1, 0, 0, 0, ms);
            if (ms < best) { best = ms; best_l = l; best_cb = cb; best_x = 1; best_G = 10; }
        }
    if (!r) r = svm_set_lanes(ctx, best_l);
    // the launch structures of the step (svm_set_step_form)
    for (int sform = 1; sform <= 3 && !r; ++sform) {
        r = svm_set_tuning(ctx, best_cb, 10, 1);
        if (!r) r = svm_set_step_form(ctx, sform);
        double ms = 0.0;
        if (!r) r = trial(10, &ms);
        if (r) break;
        record(best_l, best_cb, 10, 1, 0, sform, 0, ms);
        if (ms < best) { best = ms; best_s = sform; best_x = 1; best_G = 10; }
    }
    if (!r) r = svm_set_step_form(ctx, best_s);
    // the geometry pass with 2 and 4 cells per thread
    for (int gform = 1; gform <= 2 && !r; ++gform) {
        r = svm_set_tuning(ctx, best_cb, 10, 1);
        if (!r) r = svm_set_geometry_form(ctx, gform);
        double ms = 0.0;
        if (!r) r = trial(10, &ms);
        if (r) break;
        record(best_l, best_cb, 10, 1, 0, best_s, gform, ms);
        if (ms < best) { best = ms; best_g = gform; best_x = 1; best_G = 10; }
    }
    if (!r) r = svm_set_geometry_form(ctx, best_g);
    // the corner forms and the three-wavefront form of the vertex pass (the
    // FP64 parity arithmetic only)
    for (int form = 1; form <= 4 && !r && ctx->force_fp32 == 0; ++form) {
        r = svm_set_tuning(ctx, best_cb, 10, 1);
        if (!r) r = svm_set_vertex_form(ctx, form);
        double ms = 0.0;
        if (!r) r = trial(10, &ms);
        if (r) break;
        record(best_l, best_cb, 10, 1, form, best_s, best_g, ms);
        if (ms < best) { best = ms; best_f = form; best_x = 1; best_G = 10; }
    }
    if (!r) r = svm_set_vertex_form(ctx, best_f);
    // the other kernel schedules at the default graph length
    const int cbs[] = {128, 256, 512, 1024};
    for (int xcd = 0; xcd <= 1 && !r; ++xcd)
        for (int cb : cbs) {
            if (xcd == 1 && (cb == 256 || cb == 1024)) continue;   // measured above
            r = svm_set_tuning(ctx, cb, 10, xcd);
            double ms = 0.0;
            if (!r) r = trial(10, &ms);
            if (r) break;
            record(best_l, cb, 10, xcd, best_f, best_s, best_g, ms);
            if (ms < best) { best = ms; best_cb = cb; best_x = xcd; best_G = 10; }
        }
    // then the graph length
    const int Gs[] = {4, 20, 40};
    for (int G : Gs) {
        if (r) break;
        r = svm_set_tuning(ctx, best_cb, G, best_x);
        double ms = 0.0;
        if (!r) r = trial(G, &ms);
        if (r) break;
        record(best_l, best_cb, G, best_x, best_f, best_s, best_g, ms);
        if (ms < best) { best = ms; best_G = G; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (r) return r;
    if (rows_written) *rows_written = nrows;
    r = svm_set_lanes(ctx, best_l);
    r = r ? r : svm_set_vertex_form(ctx, best_f);
    r = r ? r : svm_set_step_form(ctx, best_s);
    r = r ? r : svm_set_geometry_form(ctx, best_g);
    r = r ? r : svm_set_tuning(ctx, best_cb, best_G, best_x);
    if (!r) ctx->tuned_ms = best;
    return r;
}

}  // extern "C"

// ---- persisted schedules ----
namespace {
int floor_log2(long long x) {
    int k = -1;
    while (x > 0) { x >>= 1; ++k; }
    return k;
}
// The key of a batch's schedule: updater, force precision, and the binary
// orders of magnitude of its sphere count and of its mean sphere size.
std::array<int, 4> schedule_key(svm_context* ctx, int updater) {
    long long cells = 0;
    for (int s = 0; s < ctx->nsys; ++s) cells += ctx->layout[s].nc;
    return {updater, ctx->force_fp32, floor_log2(ctx->nsys), floor_log2(cells / std::max(1, ctx->nsys))};
}
// v2 lines carry the step and geometry forms; v1 lines (without them) are
// read as forms 0
const char* kScheduleTag = "svm-schedule-v2";
const char* kScheduleTagV1 = "svm-schedule-v1";
}  // namespace

extern "C" {

int svm_tuning_save(svm_context* ctx, int32_t updater, const char* path) {
    if (!ctx || !path || updater < 0 || updater > 2) return fail(ctx, SVM_ERR_ARG, "bad argument");
    const auto key = schedule_key(ctx, updater);
    std::vector<std::string> keep;
    if (FILE* f = std::fopen(path, "r")) {
        char line[512];
        while (std::fgets(line, sizeof line, f)) {
            char tag[64];
            int k[4];
            if (std::sscanf(line, "%63s %d %d %d %d", tag, &k[0], &k[1], &k[2], &k[3]) == 5 &&
                (std::strcmp(tag, kScheduleTag) == 0 || std::strcmp(tag, kScheduleTagV1) == 0) && k[0] == key[0] &&
                k[1] == key[1] && k[2] == key[2] && k[3] == key[3])
                continue;   // replaced below
            keep.push_back(line);
        }
        std::fclose(f);
    }
    char row[512];
    std::snprintf(row, sizeof row, "%s %d %d %d %d %d %d %d %d %d %d %d %.6f\n", kScheduleTag, key[0], key[1],
                  key[2], key[3], ctx->nlanes, ctx->cluster_blocks, ctx->graph_steps, ctx->d.xcd, ctx->vertex_corner,
                  ctx->step_form, ctx->geometry_form, ctx->tuned_ms);
    keep.push_back(row);
    const std::string tmp = std::string(path) + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "w");
    if (!f) return fail(ctx, SVM_ERR_IO, "cannot write " + tmp);
    bool ok = true;
    for (const auto& l : keep) ok &= std::fputs(l.c_str(), f) >= 0;
    ok &= std::fclose(f) == 0;
    if (!ok || std::rename(tmp.c_str(), path) != 0) return fail(ctx, SVM_ERR_IO, std::string("cannot write ") + path);
    return SVM_OK;
}

int svm_tuning_load(svm_context* ctx, int32_t updater, const char* path, int32_t* found) {
    if (!ctx || !path || !found || updater < 0 || updater > 2) return fail(ctx, SVM_ERR_ARG, "bad argument");
    *found = 0;
    const auto key = schedule_key(ctx, updater);
    FILE* f = std::fopen(path, "r");
    if (!f) return SVM_OK;
    char line[512];
    int sched[7];
    double ms = 0.0;
    while (std::fgets(line, sizeof line, f)) {
        char tag[64] = {0};
        int k[4], v[7] = {0, 0, 0, 0, 0, 0, 0};
        double t = 0.0;
        bool ok = false;
        if (std::sscanf(line, "%63s", tag) == 1 && std::strcmp(tag, kScheduleTag) == 0)
            ok = std::sscanf(line, "%63s %d %d %d %d %d %d %d %d %d %d %d %lf", tag, &k[0], &k[1], &k[2], &k[3],
                             &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6], &t) == 13;
        else if (std::strcmp(tag, kScheduleTagV1) == 0)
            ok = std::sscanf(line, "%63s %d %d %d %d %d %d %d %d %d %lf", tag, &k[0], &k[1], &k[2], &k[3], &v[0],
                             &v[1], &v[2], &v[3], &v[4], &t) == 11;
        if (ok && k[0] == key[0] && k[1] == key[1] && k[2] == key[2] && k[3] == key[3]) {
            std::memcpy(sched, v, sizeof sched);
            ms = t;
            *found = 1;
        }
    }
    std::fclose(f);
    if (!*found) return SVM_OK;
    int r = svm_set_lanes(ctx, sched[0]);
    r = r ? r : svm_set_tuning(ctx, sched[1], sched[2], sched[3]);
    r = r ? r : svm_set_vertex_form(ctx, sched[4]);
    r = r ? r : svm_set_step_form(ctx, sched[5]);
    r = r ? r : svm_set_geometry_form(ctx, sched[6]);
    if (r) {
        *found = 0;
        return r;
    }
    ctx->tuned_ms = ms;
    return SVM_OK;
}

int svm_phase_boundary_probe(int32_t workgroups, int32_t phases, int64_t bytes_per_phase, double* us_per_launch,
                             double* us_per_barrier, int64_t* errors) {
    if (workgroups < 1 || phases < 2 || bytes_per_phase < 8 * 256 || !us_per_launch || !us_per_barrier || !errors)
        return SVM_ERR_ARG;
    int ndev = 0, dev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1 || hipGetDevice(&dev) != hipSuccess)
        return SVM_ERR_NO_DEVICE;
    // co-residency of the persistent grid, checked before it is launched
    int per_cu = 0, cus = 0, coop = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_probe_persistent, 256, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess)
        return SVM_ERR_HIP;
    if (!coop || workgroups > per_cu * cus) return SVM_ERR_ARG;
    const int per = (int)std::max<int64_t>(256, bytes_per_phase / 8 / workgroups);
    double* rows = nullptr;
    unsigned* counters = nullptr;   // arrivals, launch-form errors, persistent-form errors
    int* gave_up = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int rc = SVM_OK;
    auto check = [&](hipError_t e) { if (e != hipSuccess && rc == SVM_OK) rc = SVM_ERR_HIP; return rc == SVM_OK; };
    if (check(hipMalloc(&rows, sizeof(double) * 2 * (size_t)workgroups * per)) &&
        check(hipMalloc(&counters, 3 * sizeof(unsigned))) && check(hipMalloc(&gave_up, sizeof(int))) &&
        check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) && check(hipEventCreate(&e0)) &&
        check(hipEventCreate(&e1)) && check(hipMemsetAsync(rows, 0, sizeof(double) * 2 * (size_t)workgroups * per, st)) &&
        check(hipMemsetAsync(counters, 0, 3 * sizeof(unsigned), st)) && check(hipMemsetAsync(gave_up, 0, sizeof(int), st))) {
        // kernel boundaries: `phases` launches captured in one graph
        check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        for (int ph = 0; ph < phases; ++ph)
            hipLaunchKernelGGL(k_probe_phase, dim3(workgroups), dim3(256), 0, st, rows, per, ph, counters + 1);
        check(hipStreamEndCapture(st, &graph));
        if (rc == SVM_OK && check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0)) &&
            check(hipGraphLaunch(exec, st)) && check(hipEventRecord(e0, st)) && check(hipGraphLaunch(exec, st)) &&
            check(hipEventRecord(e1, st)) && check(hipEventSynchronize(e1))) {
            float ms = 0.f;
            check(hipEventElapsedTime(&ms, e0, e1));
            *us_per_launch = 1000.0 * ms / phases;
        }
        // grid barriers: one cooperative launch (a warm one, then the timed one)
        int nph = phases, lim = 1 << 22;
        unsigned* arrivals = counters;
        unsigned* perr = counters + 2;
        void* args[] = {&rows, (void*)&per, &nph, &arrivals, &perr, &gave_up, &lim};
        for (int pass = 0; pass < 2 && rc == SVM_OK; ++pass) {
            check(hipMemsetAsync(arrivals, 0, sizeof(unsigned), st));
            if (pass) check(hipEventRecord(e0, st));
            check(hipLaunchCooperativeKernel((const void*)k_probe_persistent, dim3(workgroups), dim3(256), args, 0, st));
            if (pass) check(hipEventRecord(e1, st));
            check(hipStreamSynchronize(st));
        }
        if (rc == SVM_OK) {
            float ms = 0.f;
            check(hipEventElapsedTime(&ms, e0, e1));
            *us_per_barrier = 1000.0 * ms / phases;
            unsigned h[3] = {0, 0, 0};
            int g = 0;
            check(hipMemcpy(h, counters, sizeof h, hipMemcpyDeviceToHost));
            check(hipMemcpy(&g, gave_up, sizeof g, hipMemcpyDeviceToHost));
            *errors = (int64_t)h[1] + h[2];
            if (g) rc = SVM_ERR_HIP;   // a barrier wait ran out: no timing
        }
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (rows) (void)hipFree(rows);
    if (counters) (void)hipFree(counters);
    if (gave_up) (void)hipFree(gave_up);
    return rc;
}

// The short root and quotient sequences of the vertex pass (svm::FastOps)
// against the operators: sample i draws a divisor b with a random mantissa
// and an exponent in [-420, 520), a numerator a = (random mantissa, random
// sign) * 2^(e_b + d) with d in [-430, 430), a root argument x with an
// exponent in [-800, 1024) (random sign one time in 8), and, one sample in 16,
// the mantissas all ones or all zeros (all-ones divisors over powers of two
// among them: quotients just above a rounding midpoint), zero numerators,
// zero, infinite, NaN and subnormal root arguments and all-ones root
// mantissas.  Counts per kind: samples in the sequence's precondition
// (divisor in [2^-384, 2^500)), accepted results (not replaced by a NaN),
// accepted results whose bits differ from the operator's, and accepted
// results that are not the double nearest to the exact quotient or root.
__device__ inline double bits_to_d(uint32_t hi, uint32_t lo) {
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
// ulp of a normal double of biased exponent e > 52 (a power of two)
__device__ inline double ulp_of(uint64_t u) { return bits_to_d((uint32_t)(((u >> 52) & 0x7ff) - 52) << 20, 0u); }
// Is q (normal, |q| in [2^-400, 2^400), |b| in [2^-384, 2^500)) the double
// nearest to a / b, ties to even?  r = a - b q is exact for a q within an
// ulp of the quotient, and a / b = q + r / b lies on the side of q that
// r / b's sign gives, where the spacing is an ulp (half an ulp below a power
// of two); |b| times half the spacing is exact.
__device__ inline bool nearest_quotient(double a, double b, double q) {
    const double r = __builtin_fma(-b, q, a);
    if (r == 0.0) return true;
    const uint64_t u = __builtin_bit_cast(uint64_t, q);
    double sp = ulp_of(u);
    const bool away = ((r > 0) == (b > 0)) ? (q > 0) : (q < 0);
    if (!away && (u & 0xfffffffffffffull) == 0) sp *= 0.5;
    const double half = __builtin_fabs(b) * (0.5 * sp);
    const double ar = __builtin_fabs(r);
    return ar < half || (ar == half && (u & 1) == 0);
}
// Is g (positive normal, g^2 normal) the double nearest to sqrt(x)?  With
// r = x - g^2 (exact for a faithful g), sqrt(x) <= g + s/2 iff
// r <= g s + s^2 / 4, and >= g - s'/2 iff r >= -(g s' - s'^2 / 4), s and s'
// the spacings above and below g; r - g s and r + g s' are exact where the
// answer is close (Sterbenz).  Roots are never exactly at a midpoint.
__device__ inline bool nearest_root(double x, double g) {
    const double r = __builtin_fma(-g, g, x);
    const uint64_t u = __builtin_bit_cast(uint64_t, g);
    const double s = ulp_of(u);
    const double sl = (u & 0xfffffffffffffull) == 0 ? 0.5 * s : s;
    if (r > 0) return r - g * s <= 0.25 * s * s;
    return r + g * sl >= 0.25 * sl * sl;
}
__global__ __launch_bounds__(256) void k_fast_ops_check(uint64_t seed, int64_t n, unsigned long long* counts) {
    unsigned long long c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const svm::FastOps fo{};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const svm::u32x4 r = svm::philox4x32_10((uint32_t)i, (uint32_t)(i >> 32), 0u, 0x46534f31u, (uint32_t)seed,
                                                (uint32_t)(seed >> 32));
        const svm::u32x4 t = svm::philox4x32_10((uint32_t)i, (uint32_t)(i >> 32), 1u, 0x46534f31u, (uint32_t)seed,
                                                (uint32_t)(seed >> 32));
        const int special = (t.v[3] & 15u) == 0u;
        const int eb = (int)(t.v[0] % 940u) - 420;
        const int d = (int)(t.v[1] % 861u) - 430;
        uint32_t mb_hi = r.v[0] & 0xfffffu, mb_lo = r.v[1];
        uint32_t ma_hi = r.v[2] & 0xfffffu, ma_lo = r.v[3];
        if (special) {
            const uint32_t k = (t.v[3] >> 4) & 3u;
            if (k == 0) {
                mb_hi = 0xfffffu;
                mb_lo = 0xffffffffu;
                if (t.v[2] & 2u) { ma_hi = 0u; ma_lo = 0u; }   // a power of two
            }
            if (k == 1) { mb_hi = 0u; mb_lo = 0u; }
            if (k == 2) { ma_hi = 0xfffffu; ma_lo = 0xffffffffu; }
        }
        const double b = bits_to_d((uint32_t)(eb + 1023) << 20 | mb_hi, mb_lo);
        int ea = eb + d;
        ea = ea < -1022 ? -1022 : (ea > 1023 ? 1023 : ea);
        double a = bits_to_d((uint32_t)(ea + 1023) << 20 | ma_hi | (r.v[0] & 0x80000000u), ma_lo);
        if (special && ((t.v[3] >> 4) & 3u) == 3u) a = (t.v[2] & 1u) ? -0.0 : 0.0;
        if (b >= 0x1p-384 && b < 0x1p500) {
            ++c[0];
            const double q = fo.div(a, b, fo.rcp(b));
            if (!(q != q)) {
                ++c[1];
                c[2] += __builtin_bit_cast(uint64_t, q) != __builtin_bit_cast(uint64_t, a / b);
                c[6] += !nearest_quotient(a, b, q);
            }
        }
        const int ex = (int)(t.v[2] % 1824u) - 800;
        double x = bits_to_d((uint32_t)(ex + 1023) << 20 | (r.v[1] & 0xfffffu), r.v[2] ^ r.v[3]);
        if ((t.v[1] & 7u) == 0u) x = -x;
        if (special) {
            const uint32_t k = (t.v[2] >> 8) & 7u;
            if (k == 0) x = 0.0;
            if (k == 1) x = __builtin_inf();
            if (k == 2) x = __builtin_nan("");
            if (k == 3) x = bits_to_d(r.v[0] & 0xfffffu, r.v[1]);   // subnormal
            if (k == 4) x = bits_to_d((uint32_t)(ex + 1023) << 20 | 0xfffffu, 0xffffffffu);
        }
        ++c[3];
        const double y = fo.sqrt(x);
        if (!(y != y)) {
            ++c[4];
            c[5] += __builtin_bit_cast(uint64_t, y) != __builtin_bit_cast(uint64_t, __builtin_sqrt(x));
            c[7] += !nearest_root(x, y);
        }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        unsigned long long v = c[k];
        for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
        if ((threadIdx.x & 63) == 0 && v) atomicAdd(counts + k, v);
    }
}

int svm_fast_ops_check(int64_t n, uint64_t seed, int64_t* counts) {
    if (n < 1 || !counts) return SVM_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return SVM_ERR_NO_DEVICE;
    unsigned long long* dc = nullptr;
    int rc = SVM_OK;
    if (hipMalloc(&dc, 8 * sizeof(unsigned long long)) != hipSuccess) return SVM_ERR_HIP;
    unsigned long long h[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (hipMemcpy(dc, h, sizeof h, hipMemcpyHostToDevice) != hipSuccess) rc = SVM_ERR_HIP;
    if (!rc) {
        const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
        hipLaunchKernelGGL(k_fast_ops_check, dim3((unsigned)blocks), dim3(256), 0, nullptr, seed, n, dc);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(h, dc, sizeof h, hipMemcpyDeviceToHost) != hipSuccess)
            rc = SVM_ERR_HIP;
    }
    (void)hipFree(dc);
    for (int k = 0; k < 8; ++k) counts[k] = (int64_t)h[k];
    return rc;
}

int svm_copy_probe(int64_t n_bytes, int32_t reps, double* ms_per_copy) {
    if (n_bytes < 8 || reps < 1 || !ms_per_copy) return SVM_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return SVM_ERR_NO_DEVICE;
    const size_t n = (size_t)(n_bytes / 8);
    double *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = SVM_OK;
    if (hipMalloc(&a, n * 8) != hipSuccess || hipMalloc(&b, n * 8) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        rc = SVM_ERR_HIP;
    } else {
        const int threads = 256;
        const int blocks = (int)std::min<size_t>((n + threads - 1) / threads, 16384);
        hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(threads), 0, nullptr, (const double*)a, a, n);  // touch
        if (hipEventRecord(e0, nullptr) != hipSuccess) rc = SVM_ERR_HIP;
        for (int r = 0; r < reps && !rc; ++r)
            hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(threads), 0, nullptr, (const double*)a, b, n);
        float ms = 0.f;
        if (rc || hipEventRecord(e1, nullptr) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            rc = SVM_ERR_HIP;
        *ms_per_copy = ms / reps;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

}  // extern "C"

// ============================================================ learned warm start
struct svm_mlp {
    float* theta = nullptr;           // FP32 master weights [SVM_MLP_NP]
    unsigned short* wb = nullptr;     // bf16 copies: w1 [1024], w2 [4096], w3 [4096], w4 [1024]
    float* musig = nullptr;           // mu [16], sigma [16]
    float* mom = nullptr;
    float* vel = nullptr;
    float* grad = nullptr;
    float out_scale = 1.0f;
    long long adam_t = 0;
    hipStream_t stream = nullptr;
    // workspace, grown on demand
    size_t cap = 0;
    float *X = nullptr, *T = nullptr, *Y = nullptr, *Z = nullptr, *H = nullptr, *D = nullptr, *partial = nullptr,
          *lossp = nullptr;
    int* idx = nullptr;
};

namespace {

svm::MlpParams mlp_params(const svm_mlp* m) {
    svm::MlpParams p;
    p.theta = m->theta;
    p.w1 = m->wb;
    p.w2 = m->wb + 1024;
    p.w3 = m->wb + 5120;
    p.w4 = m->wb + 9216;
    p.mu = m->musig;
    p.sig = m->musig + 16;
    p.out_scale = m->out_scale;
    return p;
}

int mlp_refresh_bf16(svm_mlp* m) {
    hipLaunchKernelGGL(svm::k_mlp_to_bf16, dim3(16), dim3(256), 0, m->stream, (const float*)m->theta, m->wb,
                       m->wb + 1024, m->wb + 5120, m->wb + 9216);
    return hipGetLastError() == hipSuccess ? SVM_OK : SVM_ERR_HIP;
}

int mlp_reserve(svm_mlp* m, size_t n) {
    if (n <= m->cap) return SVM_OK;
    void* old[9] = {m->X, m->T, m->Y, m->Z, m->H, m->D, m->partial, m->lossp, m->idx};
    (void)hipStreamSynchronize(m->stream);
    for (void* p : old)
        if (p) (void)hipFree(p);
    const size_t chunks = (n + 63) / 64, blocks = (n + 255) / 256;
    if (hipMalloc(&m->X, sizeof(float) * 16 * n) != hipSuccess || hipMalloc(&m->T, sizeof(float) * 2 * n) != hipSuccess ||
        hipMalloc(&m->Y, sizeof(float) * 2 * n) != hipSuccess ||
        hipMalloc(&m->Z, sizeof(float) * SVM_MLP_ZS * n) != hipSuccess ||
        hipMalloc(&m->H, sizeof(float) * SVM_MLP_ZS * n) != hipSuccess ||
        hipMalloc(&m->D, sizeof(float) * SVM_MLP_DS * n) != hipSuccess ||
        hipMalloc(&m->partial, sizeof(float) * SVM_MLP_NP * chunks) != hipSuccess ||
        hipMalloc(&m->lossp, sizeof(float) * blocks) != hipSuccess || hipMalloc(&m->idx, sizeof(int) * n) != hipSuccess) {
        m->cap = 0;
        return SVM_ERR_HIP;
    }
    m->cap = n;
    return SVM_OK;
}

// forward on device rows X[0..n) -> Y; with training buffers when `train`
int mlp_forward_dev(svm_mlp* m, const float* X, int n, float* Y, bool train) {
    const int blocks = std::max(1, std::min(1024, (n + 63) / 64));   // 512-2048 measure alike; weights load once per wave
    hipLaunchKernelGGL(svm::k_mlp_forward, dim3(blocks), dim3(256), 0, m->stream, mlp_params(m), X, n, Y,
                       train ? m->Z : nullptr, train ? m->H : nullptr);
    return hipGetLastError() == hipSuccess ? SVM_OK : SVM_ERR_HIP;
}

// gradient of (1/n) sum |Y - T|^2 over device rows X, T [0..n) into m->grad;
// the per-block loss partials stay in m->lossp
int mlp_gradient_dev(svm_mlp* m, const float* X, const float* T, int n) {
    int r = mlp_forward_dev(m, X, n, m->Y, true);
    if (r) return r;
    const int blocks = (n + 255) / 256, chunks = (n + 63) / 64;
    hipLaunchKernelGGL(svm::k_mlp_backward, dim3(blocks), dim3(256), 0, m->stream, mlp_params(m), (const float*)m->Z,
                       (const float*)m->Y, T, n, 1.0f / n, m->D, m->lossp);
    hipLaunchKernelGGL(svm::k_mlp_wgrad, dim3(chunks), dim3(256), 0, m->stream, mlp_params(m), X, (const float*)m->H,
                       (const float*)m->D, n, m->partial);
    hipLaunchKernelGGL(svm::k_mlp_reduce, dim3((SVM_MLP_NP + 255) / 256), dim3(256), 0, m->stream,
                       (const float*)m->partial, chunks, m->grad);
    return hipGetLastError() == hipSuccess ? SVM_OK : SVM_ERR_HIP;
}

double mlp_loss_host(svm_mlp* m, int n) {
    const int blocks = (n + 255) / 256;
    std::vector<float> lp(blocks);
    if (hipMemcpyAsync(lp.data(), m->lossp, sizeof(float) * blocks, hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
        hipStreamSynchronize(m->stream) != hipSuccess)
        return -1.0;
    double s = 0.0;
    for (float x : lp) s += x;
    return s / n;
}

}  // namespace

extern "C" {

int svm_mlp_create(svm_mlp** out) {
    if (!out) return SVM_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return SVM_ERR_NO_DEVICE;
    svm_mlp* m = new svm_mlp();
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&m->theta, sizeof(float) * SVM_MLP_NP) != hipSuccess ||
        hipMalloc(&m->wb, sizeof(unsigned short) * 10240) != hipSuccess ||
        hipMalloc(&m->musig, sizeof(float) * 32) != hipSuccess ||
        hipMalloc(&m->mom, sizeof(float) * SVM_MLP_NP) != hipSuccess ||
        hipMalloc(&m->vel, sizeof(float) * SVM_MLP_NP) != hipSuccess ||
        hipMalloc(&m->grad, sizeof(float) * SVM_MLP_NP) != hipSuccess) {
        svm_mlp_destroy(m);
        return SVM_ERR_HIP;
    }
    std::vector<float> ms(32, 0.0f);
    for (int k = 16; k < 32; ++k) ms[k] = 1.0f;
    if (hipMemsetAsync(m->theta, 0, sizeof(float) * SVM_MLP_NP, m->stream) != hipSuccess ||
        hipMemsetAsync(m->mom, 0, sizeof(float) * SVM_MLP_NP, m->stream) != hipSuccess ||
        hipMemsetAsync(m->vel, 0, sizeof(float) * SVM_MLP_NP, m->stream) != hipSuccess ||
        hipMemcpyAsync(m->musig, ms.data(), sizeof(float) * 32, hipMemcpyHostToDevice, m->stream) != hipSuccess ||
        mlp_refresh_bf16(m) != SVM_OK || hipStreamSynchronize(m->stream) != hipSuccess) {
        svm_mlp_destroy(m);
        return SVM_ERR_HIP;
    }
    *out = m;
    return SVM_OK;
}

int svm_mlp_destroy(svm_mlp* m) {
    if (!m) return SVM_ERR_ARG;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    void* ptrs[] = {m->theta, m->wb, m->musig, m->mom, m->vel, m->grad, m->X, m->T, m->Y, m->Z, m->H, m->D,
                    m->partial, m->lossp, m->idx};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return SVM_OK;
}

int svm_mlp_set_parameters(svm_mlp* m, const float* theta, const float* mu, const float* sigma, float out_scale) {
    if (!m || !theta || !mu || !sigma) return SVM_ERR_ARG;
    std::vector<float> th(theta, theta + SVM_MLP_NP), ms(32);
    // structural zeros: inputs 12..15 of layer 1, outputs 2..15 of layer 4
    for (int o = 0; o < 64; ++o)
        for (int k = SVM_MLP_IN; k < 16; ++k) th[SVM_MLP_W1 + o * 16 + k] = 0.0f;
    for (int o = 2; o < 16; ++o) {
        for (int k = 0; k < 64; ++k) th[SVM_MLP_W4 + o * 64 + k] = 0.0f;
        th[SVM_MLP_B4 + o] = 0.0f;
    }
    for (int k = 0; k < 16; ++k) {
        ms[k] = k < SVM_MLP_IN ? mu[k] : 0.0f;
        ms[16 + k] = k < SVM_MLP_IN ? sigma[k] : 1.0f;
        if (!(ms[16 + k] > 0.0f)) return SVM_ERR_ARG;
    }
    m->out_scale = out_scale;
    m->adam_t = 0;
    if (hipMemcpyAsync(m->theta, th.data(), sizeof(float) * SVM_MLP_NP, hipMemcpyHostToDevice, m->stream) != hipSuccess ||
        hipMemcpyAsync(m->musig, ms.data(), sizeof(float) * 32, hipMemcpyHostToDevice, m->stream) != hipSuccess ||
        hipMemsetAsync(m->mom, 0, sizeof(float) * SVM_MLP_NP, m->stream) != hipSuccess ||
        hipMemsetAsync(m->vel, 0, sizeof(float) * SVM_MLP_NP, m->stream) != hipSuccess || mlp_refresh_bf16(m) ||
        hipStreamSynchronize(m->stream) != hipSuccess)
        return SVM_ERR_HIP;
    return SVM_OK;
}

int svm_mlp_get_parameters(svm_mlp* m, float* theta, float* mu, float* sigma, float* out_scale) {
    if (!m || !theta) return SVM_ERR_ARG;
    std::vector<float> ms(32);
    if (hipMemcpyAsync(theta, m->theta, sizeof(float) * SVM_MLP_NP, hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
        hipMemcpyAsync(ms.data(), m->musig, sizeof(float) * 32, hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
        hipStreamSynchronize(m->stream) != hipSuccess)
        return SVM_ERR_HIP;
    for (int k = 0; k < SVM_MLP_IN; ++k) {
        if (mu) mu[k] = ms[k];
        if (sigma) sigma[k] = ms[16 + k];
    }
    if (out_scale) *out_scale = m->out_scale;
    return SVM_OK;
}

int svm_mlp_forward(svm_mlp* m, const float* X, int32_t n, float* Y) {
    if (!m || !X || !Y || n < 1) return SVM_ERR_ARG;
    if (mlp_reserve(m, n)) return SVM_ERR_HIP;
    if (hipMemcpyAsync(m->X, X, sizeof(float) * 16 * n, hipMemcpyHostToDevice, m->stream) != hipSuccess ||
        mlp_forward_dev(m, m->X, n, m->Y, false) ||
        hipMemcpyAsync(Y, m->Y, sizeof(float) * 2 * n, hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
        hipStreamSynchronize(m->stream) != hipSuccess)
        return SVM_ERR_HIP;
    return SVM_OK;
}

int svm_mlp_gradient(svm_mlp* m, const float* X, const float* T, int32_t n, float* grad, double* loss) {
    if (!m || !X || !T || !grad || n < 1) return SVM_ERR_ARG;
    if (mlp_reserve(m, n)) return SVM_ERR_HIP;
    if (hipMemcpyAsync(m->X, X, sizeof(float) * 16 * n, hipMemcpyHostToDevice, m->stream) != hipSuccess ||
        hipMemcpyAsync(m->T, T, sizeof(float) * 2 * n, hipMemcpyHostToDevice, m->stream) != hipSuccess ||
        mlp_gradient_dev(m, m->X, m->T, n) ||
        hipMemcpyAsync(grad, m->grad, sizeof(float) * SVM_MLP_NP, hipMemcpyDeviceToHost, m->stream) != hipSuccess)
        return SVM_ERR_HIP;
    const double l = mlp_loss_host(m, n);
    if (l < 0) return SVM_ERR_HIP;
    if (loss) *loss = l;
    return SVM_OK;
}

int svm_mlp_train(svm_mlp* m, const float* X, const float* T, int32_t n, int32_t epochs, int32_t batch, float lr,
                  uint64_t seed, double* losses) {
    if (!m || !X || !T || n < 1 || epochs < 1 || batch < 1 || !(lr > 0.0f)) return SVM_ERR_ARG;
    batch = std::min(batch, n);
    // the whole set plus one gathered batch
    float *Xall = nullptr, *Tall = nullptr;
    if (mlp_reserve(m, batch)) return SVM_ERR_HIP;
    if (hipMalloc(&Xall, sizeof(float) * 16 * (size_t)n) != hipSuccess ||
        hipMalloc(&Tall, sizeof(float) * 2 * (size_t)n) != hipSuccess) {
        if (Xall) (void)hipFree(Xall);
        return SVM_ERR_HIP;
    }
    int r = SVM_OK;
    if (hipMemcpyAsync(Xall, X, sizeof(float) * 16 * (size_t)n, hipMemcpyHostToDevice, m->stream) != hipSuccess ||
        hipMemcpyAsync(Tall, T, sizeof(float) * 2 * (size_t)n, hipMemcpyHostToDevice, m->stream) != hipSuccess)
        r = SVM_ERR_HIP;
    std::vector<int> perm(n);
    for (int i = 0; i < n; ++i) perm[i] = i;
    uint64_t state = seed * 0x9e3779b97f4a7c15ull + 1;
    auto next = [&]() {   // splitmix64
        uint64_t z = (state += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    };
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    for (int ep = 0; ep < epochs && !r; ++ep) {
        for (int i = n - 1; i > 0; --i) std::swap(perm[i], perm[(int)(next() % (uint64_t)(i + 1))]);
        double lsum = 0.0;
        for (int b0 = 0; b0 < n && !r; b0 += batch) {
            const int nb = std::min(batch, n - b0);
            const int tb = 256;
            if (hipMemcpyAsync(m->idx, perm.data() + b0, sizeof(int) * nb, hipMemcpyHostToDevice, m->stream) !=
                hipSuccess) {
                r = SVM_ERR_HIP;
                break;
            }
            hipLaunchKernelGGL(k_gather_rows, dim3((nb * 16 + tb - 1) / tb), dim3(tb), 0, m->stream,
                               (const float*)Xall, (const int*)m->idx, nb, 16, m->X);
            hipLaunchKernelGGL(k_gather_rows, dim3((nb * 2 + tb - 1) / tb), dim3(tb), 0, m->stream,
                               (const float*)Tall, (const int*)m->idx, nb, 2, m->T);
            r = mlp_gradient_dev(m, m->X, m->T, nb);
            if (r) break;
            m->adam_t += 1;
            const float c1 = 1.0f - std::pow(beta1, (float)m->adam_t), c2 = 1.0f - std::pow(beta2, (float)m->adam_t);
            hipLaunchKernelGGL(svm::k_mlp_adam, dim3((SVM_MLP_NP + 255) / 256), dim3(256), 0, m->stream, m->theta,
                               (const float*)m->grad, m->mom, m->vel, lr, beta1, beta2, eps, c1, c2);
            r = mlp_refresh_bf16(m);
            if (r) break;
            const double l = mlp_loss_host(m, nb);
            if (l < 0) { r = SVM_ERR_HIP; break; }
            lsum += l * nb;
        }
        if (losses) losses[ep] = lsum / n;
    }
    (void)hipStreamSynchronize(m->stream);
    (void)hipFree(Xall);
    (void)hipFree(Tall);
    return r;
}

int svm_mlp_time_forward(svm_mlp* m, int32_t n, int32_t reps, double* ms_per_launch) {
    if (!m || n < 16 || reps < 1 || !ms_per_launch) return SVM_ERR_ARG;
    if (mlp_reserve(m, n)) return SVM_ERR_HIP;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int r = SVM_OK;
    hipLaunchKernelGGL(k_fill_hash, dim3((unsigned)((16 * (size_t)n + 255) / 256)), dim3(256), 0, m->stream, m->X,
                       16 * (size_t)n);
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) r = SVM_ERR_HIP;
    if (!r) {
        r = mlp_forward_dev(m, m->X, n, m->Y, false);   // warm
        (void)hipEventRecord(e0, m->stream);
        for (int i = 0; i < reps && !r; ++i) r = mlp_forward_dev(m, m->X, n, m->Y, false);
        (void)hipEventRecord(e1, m->stream);
        float t = 0.f;
        if (!r && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess)) r = SVM_ERR_HIP;
        *ms_per_launch = (double)t / reps;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return r;
}

int svm_cell_descriptors(svm_context* ctx, int32_t sys, float* X, float* frames) {
    if (!valid_sys(ctx, sys) || !X) return fail(ctx, SVM_ERR_ARG, "null argument");
    int r = require_topology(ctx);
    if (r) return r;
    r = svm_compute_forces(ctx, 0);
    if (r) return r;
    float *dX = nullptr, *dF = nullptr;
    if (hipMalloc(&dX, sizeof(float) * 16 * (size_t)ctx->Cs) != hipSuccess ||
        hipMalloc(&dF, sizeof(float) * 6 * (size_t)ctx->Cs) != hipSuccess) {
        if (dX) (void)hipFree(dX);
        return fail(ctx, SVM_ERR_HIP, "device allocation failed");
    }
    const SysParams& L = ctx->layout[sys];
    hipLaunchKernelGGL(svm::k_cell_descriptors, dim3(ctx->nct), dim3(SVM_TILE), 0, ctx->stream, ctx->d, ctx->pcur, dX,
                       dF);
    r = last_launch(ctx);
    if (!r && (hipMemcpyAsync(X, dX + 16 * (size_t)L.co, sizeof(float) * 16 * L.nc, hipMemcpyDeviceToHost,
                              ctx->stream) != hipSuccess ||
               (frames && hipMemcpyAsync(frames, dF + 6 * (size_t)L.co, sizeof(float) * 6 * L.nc,
                                         hipMemcpyDeviceToHost, ctx->stream) != hipSuccess) ||
               hipStreamSynchronize(ctx->stream) != hipSuccess))
        r = fail(ctx, SVM_ERR_HIP, "descriptor copy failed");
    (void)hipFree(dX);
    (void)hipFree(dF);
    return r;
}

int svm_warm_start(svm_context* ctx, svm_mlp* m, double scale, double max_step, int32_t* reverted) {
    if (!ctx || !m || !(max_step == max_step)) return SVM_ERR_ARG;
    int r = require_topology(ctx);
    if (r) return r;
    r = svm_compute_forces(ctx, 0);
    if (r) return r;
    Dev& d = ctx->d;
    const size_t Vs = ctx->Vs, Cs = ctx->Cs, cap = ctx->cap;
    DevScratch w;
    float* dX = w.get<float>(16 * Cs);
    float* dF = w.get<float>(6 * Cs);
    float* dY = w.get<float>(2 * Cs);
    // the state the move may change, to put back per sphere
    double* sp = w.get<double>(3 * Vs);
    int4* svn = w.get<int4>(Vs);
    int4* svcn = w.get<int4>(Vs);
    int32_t* svln = w.get<int32_t>(Vs);
    int32_t* scnn = w.get<int32_t>(Cs);
    int32_t* sring = w.get<int32_t>(cap * Cs);
    if (!dX || !dF || !dY || !sp || !svn || !svcn || !svln || !scnn || !sring)
        return fail(ctx, SVM_ERR_HIP, "device allocation failed");
    const int p0 = ctx->pcur;
    // how far each sphere's areas are from tiling it before the move (a raw
    // random start can already hold folded cells)
    std::vector<double> A(Cs), err0(ctx->nsys);
    HIPCHECK(ctx, hipMemcpyAsync(A.data(), d.A, sizeof(double) * Cs, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    auto tiling_error = [&](int s, bool* ok) {
        const SysParams& L = ctx->layout[s];
        double sum = 0.0;
        *ok = true;
        for (int c = L.co; c < L.co + L.nc; ++c) {
            *ok = *ok && A[c] > 0.0 && std::isfinite(A[c]);
            sum += A[c];
        }
        return std::fabs(sum / (double)L.nc - 1.0);   // R = sqrt(N / 4 pi): the sphere's area is N
    };
    for (int s = 0; s < ctx->nsys; ++s) {
        bool ok;
        err0[s] = tiling_error(s, &ok);
    }
    HIPCHECK(ctx, hipMemcpyAsync(sp, d.px[p0], sizeof(double) * Vs, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHECK(ctx, hipMemcpyAsync(sp + Vs, d.py[p0], sizeof(double) * Vs, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHECK(ctx, hipMemcpyAsync(sp + 2 * Vs, d.pz[p0], sizeof(double) * Vs, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHECK(ctx, hipMemcpyAsync(svn, d.vn, sizeof(int4) * Vs, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHECK(ctx, hipMemcpyAsync(svcn, d.vcn, sizeof(int4) * Vs, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHECK(ctx, hipMemcpyAsync(svln, d.vln, sizeof(int32_t) * Vs, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHECK(ctx, hipMemcpyAsync(scnn, d.cnn, sizeof(int32_t) * Cs, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHECK(ctx, hipMemcpyAsync(sring, d.ring, sizeof(int32_t) * cap * Cs, hipMemcpyDeviceToDevice, ctx->stream));
    // the network runs on the context's stream
    hipStream_t keep = m->stream;
    m->stream = ctx->stream;
    hipLaunchKernelGGL(svm::k_cell_descriptors, dim3(ctx->nct), dim3(SVM_TILE), 0, ctx->stream, d, p0, dX, dF);
    r = last_launch(ctx);
    if (!r) r = mlp_forward_dev(m, dX, ctx->Cs, dY, false);
    m->stream = keep;
    if (r) return r;
    hipLaunchKernelGGL(svm::k_warm_disp, dim3(ctx->nvt), dim3(SVM_TILE), 0, ctx->stream, d, p0, (const float*)dY,
                       (const float*)dF, scale, max_step > 0.0 ? max_step : 0.0);
    const int pout = 1 - p0;
    // converged spheres keep their positions and topology
    hipLaunchKernelGGL(svm::k_move_ext, dim3(ctx->nvt), dim3(SVM_TILE), 0, ctx->stream, d, p0, pout, -2, 1.0);
    r = last_launch(ctx);
    if (!r) r = launch_t1(ctx, pout, ctx->ncur, ctx->ncur, 0, 0, 0, 2);
    if (r) return r;
    ctx->pcur = pout;
    // a sphere whose angle-sum areas tile it worse than before (a cell
    // folded) gets its state back: the warm start never leaves a sphere worse
    // than it was
    r = launch_geometry(ctx, 0);
    if (r) return r;
    HIPCHECK(ctx, hipMemcpyAsync(A.data(), d.A, sizeof(double) * Cs, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    int nrev = 0;
    for (int s = 0; s < ctx->nsys; ++s) {
        const SysParams& L = ctx->layout[s];
        bool ok;
        const double err1 = tiling_error(s, &ok);
        if (ok && err1 <= err0[s] + 1e-8) continue;
        ++nrev;
        const size_t vo = L.vo, nv = L.nv, co = L.co, nc = L.nc;
        HIPCHECK(ctx, hipMemcpyAsync(d.px[pout] + vo, sp + vo, sizeof(double) * nv, hipMemcpyDeviceToDevice, ctx->stream));
        HIPCHECK(ctx, hipMemcpyAsync(d.py[pout] + vo, sp + Vs + vo, sizeof(double) * nv, hipMemcpyDeviceToDevice,
                                     ctx->stream));
        HIPCHECK(ctx, hipMemcpyAsync(d.pz[pout] + vo, sp + 2 * Vs + vo, sizeof(double) * nv, hipMemcpyDeviceToDevice,
                                     ctx->stream));
        HIPCHECK(ctx, hipMemcpyAsync(d.vn + vo, svn + vo, sizeof(int4) * nv, hipMemcpyDeviceToDevice, ctx->stream));
        HIPCHECK(ctx, hipMemcpyAsync(d.vcn + vo, svcn + vo, sizeof(int4) * nv, hipMemcpyDeviceToDevice, ctx->stream));
        HIPCHECK(ctx, hipMemcpyAsync(d.vln + vo, svln + vo, sizeof(int32_t) * nv, hipMemcpyDeviceToDevice, ctx->stream));
        HIPCHECK(ctx, hipMemcpyAsync(d.cnn + co, scnn + co, sizeof(int32_t) * nc, hipMemcpyDeviceToDevice, ctx->stream));
        HIPCHECK(ctx, hipMemcpy2DAsync(d.ring + co, sizeof(int32_t) * Cs, sring + co, sizeof(int32_t) * Cs,
                                       sizeof(int32_t) * nc, cap, hipMemcpyDeviceToDevice, ctx->stream));
    }
    if (reverted) *reverted = nrev;
    return svm_synchronize(ctx);
}

int svm_dyn_set_reference(svm_context* ctx) {
    if (!ctx) return SVM_ERR_ARG;
    int r = require_topology(ctx);
    if (r) return r;
    Dev& d = ctx->d;
    const size_t Cs = ctx->Cs;
    if (!d.dref) {
        r = dalloc(ctx, &d.dref, 3 * Cs);
        if (!r) r = dalloc(ctx, &d.dpart, 2 * (size_t)ctx->nct);
        if (!r) r = dalloc(ctx, &d.dout, 2 * (size_t)ctx->nsys);
        if (r) return r;
    }
    HIPCHECK(ctx, hipMemcpyAsync(d.dref, d.cx, sizeof(double) * Cs, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHECK(ctx, hipMemcpyAsync(d.dref + Cs, d.cy, sizeof(double) * Cs, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHECK(ctx, hipMemcpyAsync(d.dref + 2 * Cs, d.cz, sizeof(double) * Cs, hipMemcpyDeviceToDevice, ctx->stream));
    return svm_synchronize(ctx);
}

int svm_dyn_measure(svm_context* ctx, double fraction, double cutoff, double* msd, double* overlap) {
    if (!ctx || !msd || !overlap || !(fraction > 0.0)) return fail(ctx, SVM_ERR_ARG, "bad argument");
    if (!ctx->d.dref) return fail(ctx, SVM_ERR_STATE, "no reference positions (svm_dyn_set_reference)");
    hipLaunchKernelGGL(svm::k_dyn_tiles, dim3(ctx->nct), dim3(SVM_TILE), 0, ctx->stream, ctx->d, fraction, cutoff);
    int r = last_launch(ctx);
    if (r) return r;
    hipLaunchKernelGGL(svm::k_dyn_fold, dim3(ctx->nsys), dim3(SVM_TILE), 0, ctx->stream, ctx->d, fraction);
    r = last_launch(ctx);
    if (r) return r;
    std::vector<double> out(2 * (size_t)ctx->nsys);
    HIPCHECK(ctx, hipMemcpyAsync(out.data(), ctx->d.dout, sizeof(double) * out.size(), hipMemcpyDeviceToHost,
                                 ctx->stream));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    for (int s = 0; s < ctx->nsys; ++s) {
        msd[s] = out[2 * s];
        overlap[s] = out[2 * s + 1];
    }
    return SVM_OK;
}

int svm_nc_open(const char* path, int32_t n, svm_nc_file** out) {
    if (!path || !out || n <= 0 || (int64_t)n * 8 > 0x7fffffffLL) return SVM_ERR_ARG;
    *out = nullptr;
    FILE* f = std::fopen(path, "wb+");
    if (!f) return SVM_ERR_IO;
    const std::string probe = svm_nc::header(n, 0, 0);
    const uint32_t hsize = (uint32_t)probe.size();
    const std::string h = svm_nc::header(n, hsize, hsize + 8);
    if (std::fwrite(h.data(), 1, h.size(), f) != h.size() || std::fflush(f) != 0) {
        std::fclose(f);
        return SVM_ERR_IO;
    }
    svm_nc_file* nc = new svm_nc_file;
    nc->f = f;
    nc->n = n;
    nc->header = (long)hsize;
    nc->recsize = 8 + 8 * (long)n;
    *out = nc;
    return SVM_OK;
}

int svm_nc_write(svm_nc_file* nc, const double* vec, double value) {
    if (!nc || !nc->f || !vec) return SVM_ERR_ARG;
    std::vector<unsigned char> rec((size_t)nc->recsize);
    svm_nc::be64d(rec.data(), value);
    for (int32_t i = 0; i < nc->n; ++i) svm_nc::be64d(rec.data() + 8 + 8 * (size_t)i, vec[i]);
    if (std::fseek(nc->f, nc->header + nc->recsize * (long)nc->numrecs, SEEK_SET) != 0 ||
        std::fwrite(rec.data(), 1, rec.size(), nc->f) != rec.size())
        return SVM_ERR_IO;
    nc->numrecs += 1;
    unsigned char cnt[4];
    for (int k = 0; k < 4; ++k) cnt[k] = (unsigned char)(((uint32_t)nc->numrecs >> (8 * (3 - k))) & 0xff);
    if (std::fseek(nc->f, 4, SEEK_SET) != 0 || std::fwrite(cnt, 1, 4, nc->f) != 4 || std::fflush(nc->f) != 0)
        return SVM_ERR_IO;
    return SVM_OK;
}

int svm_nc_num_records(svm_nc_file* nc, int32_t* n_records) {
    if (!nc || !n_records) return SVM_ERR_ARG;
    *n_records = nc->numrecs;
    return SVM_OK;
}

int svm_nc_close(svm_nc_file* nc) {
    if (!nc) return SVM_ERR_ARG;
    int r = (nc->f && std::fclose(nc->f) != 0) ? SVM_ERR_IO : SVM_OK;
    delete nc;
    return r;
}

}  // extern "C"
