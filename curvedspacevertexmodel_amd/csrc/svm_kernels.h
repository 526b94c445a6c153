// svm_kernels.h -- HIP kernels of the per-timestep hot path (gfx950, wave64).
//
// One dynamics step of a batch of spheres (or of one lane of it) is a few
// launches, each over every sphere's tiles in one grid:
//   k_geometry  (thread per cell)   cell centre, area, perimeter; per-sphere
//                                   total-area check with the fan fallback
//   k_vertex    (thread per vertex) chain-rule force over the 3 cells of the
//                                   vertex, fused with the updater epilogue
//                                   (Brownian / active displacement, the move
//                                   on the sphere and the active-director
//                                   update, or the FIRE half-kick and its
//                                   power/norm reductions)
//   k_t1_detect (thread per vertex) short-edge test -> candidate lists
//   k_t1_cluster .. k_t1_resolve    the T1 dry runs and their commit, in
//                                   exactly the order of the reference's
//                                   sequential CPU loop (svm_t1.h)
// Per-sphere reductions are deterministic: 256-row tiles folded by a fixed
// halving tree, then the tile partials folded the same way by the last block
// of that sphere to arrive.
#pragma once
#include <hip/hip_runtime.h>
#include "svm_device.h"
#include "svm_geom.h"

namespace svm {

// Workgroups are dealt round-robin to the 8 XCDs (workgroup b runs on XCD
// b % 8), each with its own L2.  Giving XCD x the contiguous tile range
// [x*per, (x+1)*per) keeps a sphere's rows, and the rows its tiles gather, in
// one L2.  Grids are launched with 8*per workgroups (xcd_grid); the few
// spare workgroups get a tile >= ntiles and return at once.
__device__ inline int xcd_tile(const Dev& d, int ntiles) {
    if (!d.xcd) return (int)blockIdx.x;
    const int per = (ntiles + 7) >> 3;
    return (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
}

// The vertex / cell tile of this workgroup within the launch's lane, or -1
// for a spare workgroup of the XCD-padded grid.
__device__ inline int lane_vtile(const Dev& d) {
    const int t = xcd_tile(d, d.vtn);
    return t < d.vtn ? d.vt0 + t : -1;
}
__device__ inline int lane_ctile(const Dev& d) {
    const int t = xcd_tile(d, d.ctn);
    return t < d.ctn ? d.ct0 + t : -1;
}


enum VertexMode { VM_FORCE = 0, VM_BROWNIAN = 1, VM_ACTIVE = 2, VM_FIRE = 3, VM_FORCE_F32 = 4, VM_ADAM = 5 };

// A sphere whose minimiser (FIRE or Adam) has converged: its force statistic
// at or below the cutoff after at least one iteration.
__device__ inline bool minimiser_converged(const SysParams* S) {
    return (S->f_iter > 0 && !(S->f_force_max > S->f_cutoff)) || (S->a_iter > 0 && !(S->a_force_max > S->a_cutoff));
}

// ------------------------------------------------------------ reductions --
// Halving tree over the 256 values of a block (s[i] += s[i+h], h = 128..1).
// The result is valid in every thread.
__device__ inline double block_tree256(double v, double* red) {
    int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    if (t < 128) red[t] = red[t] + red[t + 128];
    __syncthreads();
    if (t < 64) {
        double x = red[t] + red[t + 64];
        // h = 32..1 inside wave 0: lane i adds lane i+h
        x = x + __shfl_down(x, 32, 64);
        x = x + __shfl_down(x, 16, 64);
        x = x + __shfl_down(x, 8, 64);
        x = x + __shfl_down(x, 4, 64);
        x = x + __shfl_down(x, 2, 64);
        x = x + __shfl_down(x, 1, 64);
        if (t == 0) red[256] = x;
    }
    __syncthreads();
    double r = red[256];
    __syncthreads();
    return r;
}

// Halving tree over part[i*stride], i < n, padded with zeros to the next
// power of two m (s[i] += s[i+h], h = m/2..1), folded in the LDS buffer buf
// (m doubles).  Whole block; the partials were written by an earlier kernel,
// so plain loads see them.  The result is valid in every thread.
#define SVM_MAX_TILES 4096
__device__ inline double fold_lds(const double* part, int stride, int n, double* buf) {
    int m = 1;
    while (m < n) m <<= 1;
    for (int i = threadIdx.x; i < m; i += blockDim.x) buf[i] = i < n ? part[(size_t)i * stride] : 0.0;
    __syncthreads();
    for (int h = m >> 1; h >= 1; h >>= 1) {
        for (int i = threadIdx.x; i < h; i += blockDim.x) buf[i] = buf[i] + buf[i + h];
        __syncthreads();
    }
    double r = buf[0];
    __syncthreads();
    return r;
}

// The same fixed tree by one wavefront in registers, for m = n padded to a
// power of two at most 64 * SVM_FOLD_K: lane l holds entries l + 64 k.  The
// tree's levels with h >= 64 add a lane's own entries k and k + h / 64, the
// last six take entry i + h from lane i + h by a shuffle; every addition is
// buf[i] + buf[i + h], as in fold_lds.  Each wavefront folds for itself (no
// LDS, no barrier); all lanes return the sum.
#define SVM_FOLD_K 4
__device__ inline double fold_wave(const double* part, int stride, int n, int m) {
    const int lane = threadIdx.x & 63;
    double v[SVM_FOLD_K];
#pragma unroll
    for (int k = 0; k < SVM_FOLD_K; ++k) {
        const int i = lane + 64 * k;
        v[k] = i < n ? part[(size_t)i * stride] : 0.0;
    }
#pragma unroll
    for (int q = SVM_FOLD_K / 2; q >= 1; q >>= 1)      // h = 64 q
        if (64 * q < m) {
#pragma unroll
            for (int k = 0; k < q; ++k) v[k] = v[k] + v[k + q];
        }
    double x = v[0];
#pragma unroll
    for (int h = 32; h >= 1; h >>= 1) {
        const double y = __shfl_down(x, h, 64);
        if (h < m && lane < h) x = x + y;
    }
    return __shfl(x, 0, 64);
}

// The same fixed tree by a whole 256-thread block, for m = n padded to a
// power of two in (64 * SVM_FOLD_K, 256 * SVM_FOLD_B]: thread t holds entries
// t + 256 k in registers and adds them in pairs for the levels h >= 256;
// block_tree256 takes the last eight levels.  The same additions of the same
// pairs as fold_lds, with no LDS round trip per level: a tile's workgroup
// folds a large sphere's tile partials in a few loads and three barriers.
// Every thread returns the sum.
#define SVM_FOLD_B 16
static_assert(SVM_MAX_TILES <= 256 * SVM_FOLD_B, "fold_block covers every sphere's tiles");
__device__ inline double fold_block(const double* part, int stride, int n, int m, double* red) {
    const int t = threadIdx.x;
    double v[SVM_FOLD_B];
#pragma unroll
    for (int k = 0; k < SVM_FOLD_B; ++k) {
        const int i = t + 256 * k;
        v[k] = i < n ? part[(size_t)i * stride] : 0.0;
    }
#pragma unroll
    for (int q = SVM_FOLD_B / 2; q >= 1; q >>= 1)      // h = 256 q
        if (256 * q < m) {
#pragma unroll
            for (int k = 0; k < q; ++k) v[k] = v[k] + v[k + q];
        }
    return block_tree256(v[0], red);
}

__device__ inline double dot0(const double* a, const double* b) {
    double s = 0.0;
    s += a[0] * b[0];
    s += a[1] * b[1];
    s += a[2] * b[2];
    return s;
}

// --------------------------------------------------------------- geometry --
// sphericalVertexModel::computeGeometryCPU (src/model/sphericalVertexModel.cpp:209-311)
// for cell c of sphere S: centre (and its spherical terms for the force
// pass), perimeter and angle-sum area, stored; returns the area.
//
// PF > 0 (prefetched ring): the first PF ring slots are read in one round
// of independent loads and their positions in a second, so the centre sum
// waits for two memory round trips instead of two per slot; the angle loop's
// first three vertices come from those indices, and the loop takes the
// slots' vertex indices from LDS (`sidx`, this thread's column of a
// [PF][SVM_TILE] array).  The arithmetic is the same in the same order.
// With 8 slots (97% of rings) the pass takes 0.091 ms on the 256-sphere
// batch against 0.100 ms without (98 against 76 VGPRs, 4 against 6 waves per
// SIMD); 6 slots (80 VGPRs) gave 0.092 ms.
#define SVM_GEO_PF 8
template <int PF = 0>
__device__ __attribute__((always_inline)) inline double geometry_cell(const Dev& d, const SysParams* S,
                                                                     const double* X, const double* Y,
                                                                     const double* Z, int c,
                                                                     int* sidx = nullptr) {
    const double R = S->R;
    const int Cs = d.Cs;
    double area = 0.0;
    const int n = d.cnn[c];
    const int32_t* rc = d.ring + c;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    int idx[PF ? PF : 1];
    if (PF) {
#pragma unroll
        for (int k = 0; k < PF; ++k) {
            idx[k] = rc[(size_t)(k < n ? k : n - 1) * Cs];
            sidx[k * SVM_TILE] = idx[k];
        }
        double gx[PF ? PF : 1], gy[PF ? PF : 1], gz[PF ? PF : 1];
#pragma unroll
        for (int k = 0; k < PF; ++k) {
            gx[k] = X[idx[k]];
            gy[k] = Y[idx[k]];
            gz[k] = Z[idx[k]];
        }
#pragma unroll
        for (int k = 0; k < PF; ++k)
            if (k < n) {
                sx = sx + gx[k];
                sy = sy + gy[k];
                sz = sz + gz[k];
            }
    }
    for (int k = PF; k < n; ++k) {
        int v = rc[(size_t)k * Cs];
        sx = sx + X[v];
        sy = sy + Y[v];
        sz = sz + Z[v];
    }
    // the angle loop's first three vertices, read before the centre's stores
    int il, ic, in;
    if (PF && n <= PF) {   // the last three slots, picked from the prefetched indices
        il = idx[0], ic = idx[0], in = idx[0];
#pragma unroll
        for (int k = 1; k < PF; ++k) {
            il = k == n - 3 ? idx[k] : il;
            ic = k == n - 2 ? idx[k] : ic;
            in = k == n - 1 ? idx[k] : in;
        }
    } else {
        il = rc[(size_t)(n - 3) * Cs], ic = rc[(size_t)(n - 2) * Cs], in = rc[(size_t)(n - 1) * Cs];
    }
    double pl[3] = {X[il], Y[il], Z[il]};
    double pc[3] = {X[ic], Y[ic], Z[ic]};
    double pn[3] = {X[in], Y[in], Z[in]};
    double cs[3] = {sx, sy, sz};
    double inv = 1.0 / svm::sqrt_rn(dot0(cs, cs));
    double cxv = R * (inv * sx), cyv = R * (inv * sy), czv = R * (inv * sz);
    d.cx[c] = cxv;
    d.cy[c] = cyv;
    d.cz[c] = czv;
    {   // the centre's spherical-coordinate terms, for every vertex of the cell
        const PQ q = make_pq<double>(cxv, cyv, czv);
        d.ctrig[c] = q.cosT;
        d.ctrig[(size_t)Cs + c] = q.sinT;
        d.ctrig[2 * (size_t)Cs + c] = q.cosP;
        d.ctrig[3 * (size_t)Cs + c] = q.sinP;
    }

    double nl = pnorm(pl), nc = pnorm(pc), nnx = pnorm(pn);
    double cm[3] = {pl[1] * pc[2] - pl[2] * pc[1], pl[2] * pc[0] - pl[0] * pc[2], pl[0] * pc[1] - pl[1] * pc[0]};
    double ncm = svm::sqrt_rn(dot0(cm, cm));
    double perimeter = 0.0, angles = 0.0;
    for (int k = 0; k < n; ++k) {
        int ip = PF && k < PF ? sidx[k * SVM_TILE] : rc[(size_t)k * Cs];
        double pp[3] = {X[ip], Y[ip], Z[ip]};
        perimeter += geodesic(pl, pc, nl, nc, R);
        double ci[3] = {pc[1] * pn[2] - pc[2] * pn[1], pc[2] * pn[0] - pc[0] * pn[2], pc[0] * pn[1] - pc[1] * pn[0]};
        double nci = svm::sqrt_rn(dot0(ci, ci));
        double determinant = pl[0] * ci[0] + pl[1] * ci[1] + pl[2] * ci[2];
        double angle = svm::acos_lo(-dot0(ci, cm) / (nci * ncm));
        if (determinant > 0) angle *= -1;
        if (angle < 0) angle += 2. * SVM_PI;
        angles += angle;
        pl[0] = pc[0]; pl[1] = pc[1]; pl[2] = pc[2]; nl = nc;
        pc[0] = pn[0]; pc[1] = pn[1]; pc[2] = pn[2]; nc = nnx;
        pn[0] = pp[0]; pn[1] = pp[1]; pn[2] = pp[2]; nnx = pnorm(pp);
        cm[0] = ci[0]; cm[1] = ci[1]; cm[2] = ci[2]; ncm = nci;
    }
    area = (angles - (n - 2) * SVM_PI);
    int extraAngularArea = (int)__builtin_floor(area / (1.0 * SVM_PI));
    if (extraAngularArea > 0) area -= extraAngularArea * SVM_PI;
    area = area * R * R;
    d.A[c] = area;
    d.P[c] = perimeter;
    return area;
}

// The 256-row tile's halving tree (block_tree256) over the areas of a
// workgroup whose threads take CPT cells each: thread t holds rows
// t + k * (256 / CPT), k < CPT.  The tree's first log2(CPT) levels pair rows
// that one thread holds, so they are added in registers, in the tree's order;
// the rest is block_tree256's.  Result in thread 0.
template <int CPT>
__device__ inline double tile_tree(const double* a, double* red) {
    const int t = threadIdx.x;
    if (CPT == 1) return block_tree256(a[0], red);
    double x;
    if (CPT == 2) {                       // rows t, t + 128
        red[t] = a[0] + a[1];
        __syncthreads();
        x = t < 64 ? red[t] + red[t + 64] : 0.0;
    } else {                              // rows t, t + 64, t + 128, t + 192
        x = (a[0] + a[2]) + (a[1] + a[3]);
    }
    if (t < 64) {
        x = x + __shfl_down(x, 32, 64);
        x = x + __shfl_down(x, 16, 64);
        x = x + __shfl_down(x, 8, 64);
        x = x + __shfl_down(x, 4, 64);
        x = x + __shfl_down(x, 2, 64);
        x = x + __shfl_down(x, 1, 64);
    }
    return x;
}

// One workgroup per 256-row cell tile, CPT cells per thread (256 / CPT
// threads); the tile's area partial for the fan check.  Every CPT gives the
// same values, bit for bit: each cell is the same function, and the tile's
// sum the same tree.  PF: the prefetched ring of geometry_cell (one cell per
// thread only).
template <int CPT, int PF = 0>
__global__ __launch_bounds__(SVM_TILE / CPT) void k_geometry(Dev d, int pin, int skip_frozen) {
    __shared__ double red[SVM_TILE + 1];
    __shared__ int s_idx[PF ? PF * SVM_TILE : 1];
    const int tile = lane_ctile(d);
    if (tile < 0) return;
    const int s = d.ctile_sys[tile];
    SysParams* S = d.sys + s;
    if (skip_frozen && S->f_frozen) return;
    const double* X = d.px[pin];
    const double* Y = d.py[pin];
    const double* Z = d.pz[pin];
    static_assert(!PF || CPT == 1, "the prefetched ring is for one cell per thread");
    double area[CPT];
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int c = tile * SVM_TILE + threadIdx.x + k * (SVM_TILE / CPT);
        area[k] = c - S->co < S->nc ? geometry_cell<PF>(d, S, X, Y, Z, c, s_idx + threadIdx.x) : 0.0;
    }
    const double part = tile_tree<CPT>(area, red);
    if (threadIdx.x == 0) d.cpart[tile] = part;
}

// The total-area check of computeGeometryCPU (src/model/sphericalVertexModel.cpp:304-310):
// true when the angle-sum areas of sphere S, summed to `total` (the fixed-tree
// fold of its tile partials from k_geometry), miss 4 pi R^2 by more than
// 1e-6, so that its cells take their triangle-fan areas.
__device__ inline bool fan_needed(const SysParams* S, double total) {
    const double R = S->R;
    double excessArea = total - 4.0 * SVM_PI * R * R;
    return __builtin_fabs(excessArea) > 1e-6;
}
// the power of two the tile partials of n tiles are padded to
__device__ inline int fold_width(int n) {
    int m = 1;
    while (m < n) m <<= 1;
    return m;
}

// Triangle-fan area of cell c (recomputeAreasCPU, :162-207): the fan's
// triangles (centre, ring[k-2], ring[k-1]) in ring order, each
// triangle_area(cp, pl, pc, R) term for term; the centre's unit vector, and
// the previous triangle's unit pc, its dot with the centre and
// sqrt(1 - dot^2), are the same expressions on the same inputs, carried
// instead of recomputed.  A pure function of the ring, the positions and the
// cell centre from k_geometry.
__device__ inline double fan_area(const Dev& d, const double* X, const double* Y, const double* Z, int c,
                                  double R) {
    const int Cs = d.Cs;
    const int n = d.cnn[c];
    const int32_t* rc = d.ring + c;
    const double cp[3] = {d.cx[c], d.cy[c], d.cz[c]};
    auto unit = [](const double* p, double* u) {
        double dn = 0.0; dn += p[0] * p[0]; dn += p[1] * p[1]; dn += p[2] * p[2];
        const double inv = 1.0 / svm::sqrt_rn(dn);
        for (int k = 0; k < 3; ++k) u[k] = inv * p[k];
    };
    double ua[3], ub[3], uc[3];
    unit(cp, ua);
    const int il = rc[(size_t)(n - 2) * Cs], ic = rc[(size_t)(n - 1) * Cs];
    {
        const double pl[3] = {X[il], Y[il], Z[il]};
        unit(pl, ub);
    }
    double d12 = 0.0; d12 += ua[0] * ub[0]; d12 += ua[1] * ub[1]; d12 += ua[2] * ub[2];
    double s12 = svm::sqrt_rn(1 - d12 * d12);
    double pc[3] = {X[ic], Y[ic], Z[ic]};
    double a = 0.0;
    for (int k = 0; k < n; ++k) {
        const int ip = rc[(size_t)k * Cs];
        const double pn[3] = {X[ip], Y[ip], Z[ip]};
        unit(pc, uc);
        double d13 = 0.0; d13 += ua[0] * uc[0]; d13 += ua[1] * uc[1]; d13 += ua[2] * uc[2];
        double d23 = 0.0; d23 += ub[0] * uc[0]; d23 += ub[1] * uc[1]; d23 += ub[2] * uc[2];
        const double s13 = svm::sqrt_rn(1 - d13 * d13);
        const double s23 = svm::sqrt_rn(1 - d23 * d23);
        double area = -SVM_PI;
        area += svm::acos((d23 - d12 * d13) / (s12 * s13));
        area += svm::acos((d12 - d13 * d23) / (s13 * s23));
        area += svm::acos((d13 - d12 * d23) / (s12 * s23));
        area *= (R * R);
        a += area;
        ub[0] = uc[0]; ub[1] = uc[1]; ub[2] = uc[2];
        d12 = d13;
        s12 = s13;
        pc[0] = pn[0]; pc[1] = pn[1]; pc[2] = pn[2];
    }
    return a;
}

// The total-area check of computeGeometryCPU (src/model/sphericalVertexModel.cpp:304-310)
// and, where it fails, triangle-fan areas (recomputeAreasCPU, :162-207) for
// every cell of that sphere.  Each block folds its sphere's tile partials
// from k_geometry with the same fixed tree, so every block reaches the same
// decision without any cross-block synchronisation.
// Spheres of up to 64 * SVM_FOLD_K tiles fold in one wavefront's registers,
// larger ones in the block's (fold_block).
__global__ __launch_bounds__(SVM_TILE) void k_fan_fix(Dev d, int pin, int skip_frozen) {
    __shared__ double red[SVM_TILE + 1];
    const int tile = lane_ctile(d);
    if (tile < 0) return;
    SysParams* S = d.sys + d.ctile_sys[tile];
    if (skip_frozen && S->f_frozen) return;
    const int m = fold_width(S->nctiles);
    const double total = m <= 64 * SVM_FOLD_K ? fold_wave(d.cpart + S->ctile0, 1, S->nctiles, m)
                                              : fold_block(d.cpart + S->ctile0, 1, S->nctiles, m, red);
    if (!fan_needed(S, total)) return;
    if (tile == S->ctile0 && threadIdx.x == 0) S->n_fan += 1;
    const int c = tile * SVM_TILE + threadIdx.x;
    if (c - S->co >= S->nc) return;
    d.A[c] = fan_area(d, d.px[pin], d.py[pin], d.pz[pin], c, S->R);
}

// Active-director update of sphericalSelfPropelledParticle::integrateEOMCPU
// (src/updaters/sphericalSelfPropelledParticle.cpp:37-45) at position p.
__device__ inline void director_update(const SysParams* S, int local, const double* p, const double* nin,
                                       double* nout) {
    double u, unused;
    uniform2(S->seed, STREAM_ACTIVE, S->step, (uint32_t)local, &u, &unused);
    double randomNumber = u * (.5 - (-.5)) + (-.5);
    double theta = randomNumber * __builtin_sqrt(2.0 * S->dt * S->tau);
    double n[3] = {nin[0], nin[1], nin[2]};
    project_normalize(n, p);
    rodrigues(n, p, theta);
    nout[0] = n[0]; nout[1] = n[1]; nout[2] = n[2];
}

// --------------------------------------------------------------- vertices --
template <class T>
__device__ inline PQT<T> pq_at(const double* X, const double* Y, const double* Z, int v) {
    return make_pq<T>((T)X[v], (T)Y[v], (T)Z[v]);
}

template <class T>
__device__ inline const PQT<T>& pick3(int k, const PQT<T>& a, const PQT<T>& b, const PQT<T>& c) {
    return k == 0 ? a : (k == 1 ? b : c);
}

// Force on vertex v (sphericalVertexModel::computeForceCPU,
// src/model/sphericalVertexModel.cpp:380-416).  Returns the number of skipped
// NaN contributions.  T = double is the parity path; T = float evaluates the
// same expression trees in FP32 from the FP64 positions, cell centres and
// (A - A0, P - P0) differences.
template <class T>
__device__ inline int vertex_force(const Dev& d, const SysParams* S, const double* X, const double* Y,
                                   const double* Z, int v, double* f) {
    const T R = (T)S->R, Kr = (T)S->kr;
    PQT<T> qs = pq_at<T>(X, Y, Z, v);
    int4 nb = d.vn[v];
    int4 cl = d.vcn[v];
    int ln = d.vln[v];
    PQT<T> q0 = pq_at<T>(X, Y, Z, nb.x);
    PQT<T> q1 = pq_at<T>(X, Y, Z, nb.y);
    PQT<T> q2 = pq_at<T>(X, Y, Z, nb.z);
    T f0 = T(0.0), f1 = T(0.0), f2 = T(0.0);
    int nan = 0;
#pragma unroll
    for (int sl = 0; sl < 3; ++sl) {
        int c = sl == 0 ? cl.x : (sl == 1 ? cl.y : cl.z);
        int ls = (ln >> (4 * sl)) & 3;
        int ns = (ln >> (4 * sl + 2)) & 3;
        T areaDifference = (T)(d.A[c] - d.a0[c]);
        T perimeterDifference = (T)(d.P[c] - d.p0[c]);
        PQT<T> qc = make_pq<T>((T)d.cx[c], (T)d.cy[c], (T)d.cz[c]);
        T fs[3];
        if (pair_force(qs, pick3(ls, q0, q1, q2), pick3(ns, q0, q1, q2), qc, areaDifference, perimeterDifference,
                       Kr, R, fs)) {
            f0 = f0 + fs[0];
            f1 = f1 + fs[1];
            f2 = f2 + fs[2];
        } else {
            ++nan;
        }
    }
    f[0] = (double)f0; f[1] = (double)f1; f[2] = (double)f2;
    return nan;
}

// The parity path (T = double) of vertex_force, arranged for fewer
// operations with every value computed by the same expression as above: the
// vertex's basis and its pair terms with each neighbour are evaluated once
// (each neighbour's arc gradient serves two cells), and the cell centres'
// spherical-coordinate terms come from the geometry pass.
//
// Which two neighbours bound a cell is data (vln), but around a vertex the
// map from a cell's previous ring neighbour to its next one is a cyclic
// permutation of the three neighbour slots: 0->1->2 or 0->2->1.  Naming the
// neighbours so that it is always 0->1->2 (swapping slots 1 and 2 for the
// other orientation, an integer select of rows), cell j of the loop below is
// the one whose previous neighbour is j and next j+1 (mod 3): its operands
// are fixed registers, not per-cell selections of every neighbour term.  The
// contributions are then added in the reference's cell-slot order.
// Where the parity force reads its inputs: positions of the vertex and its
// neighbours, and each cell's (A - A0, P - P0, centre, centre's spherical
// terms).  GlobalSrc reads the arrays; the staged vertex pass reads most of
// them from its workgroup's LDS windows (StagedSrc).  Either way the same
// values enter the same expressions.
struct GlobalSrc {
    const Dev& d;
    const double *X, *Y, *Z;
    template <class Ops>
    __device__ PQ pq(int w, const Ops& ops) const { return make_pq<double>(X[w], Y[w], Z[w], ops); }
    __device__ void cell(int c, double& aD, double& pD, PQ& qc) const {
        const size_t Cs = (size_t)d.Cs;
        aD = d.A[c] - d.a0[c];
        pD = d.P[c] - d.p0[c];
        qc.x = d.cx[c]; qc.y = d.cy[c]; qc.z = d.cz[c];
        qc.cosT = d.ctrig[c]; qc.sinT = d.ctrig[Cs + c]; qc.cosP = d.ctrig[2 * Cs + c]; qc.sinP = d.ctrig[3 * Cs + c];
    }
};

template <class Src, class Ops = ExactOps>
__device__ inline int vertex_force_f64(const Src& src, const Dev& d, const SysParams* S, int v, double* f,
                                       const Ops& ops = Ops()) {
    const double R = S->R, Kr = S->kr;
    const PQ qs = src.pq(v, ops);
    const int4 nb = d.vn[v];
    const int4 cl = d.vcn[v];
    const int ln = d.vln[v];
    const int ls0 = ln & 3, ns0 = (ln >> 2) & 3, ls1 = (ln >> 4) & 3, ls2 = (ln >> 8) & 3;
    const int o = ns0 == (ls0 == 2 ? 0 : ls0 + 1) ? 0 : 1;   // 1: the cycle runs 0->2->1
    auto rn = [o](int k) { return o && k ? 3 - k : k; };
    // the renamed previous-neighbour slot of cell slots 0, 1, 2
    const int L0 = rn(ls0), L1 = rn(ls1), L2 = rn(ls2);
    const int r1 = o ? nb.z : nb.y, r2 = o ? nb.y : nb.z;
    const PQ q0 = src.pq(nb.x, ops);
    const PQ q1 = src.pq(r1, ops);
    const PQ q2 = src.pq(r2, ops);
    double tH[3], pH[3];
    basis(qs, tH, pH);
    const PairT<double> w0 = pair_terms(qs, q0), w1 = pair_terms(qs, q1), w2 = pair_terms(qs, q2);
    double g0[3], g1[3], g2[3];
    grad_geodesic_p(qs, q0, w0, tH, pH, g0, ops);
    grad_geodesic_p(qs, q1, w1, tH, pH, g1, ops);
    grad_geodesic_p(qs, q2, w2, tH, pH, g2, ops);
    // pair_force's sequence (sphericalVertexModel.cpp:386-413) for the cell
    // whose previous and next ring neighbours are (ql, qn)
    auto cell_force = [&](int c, const PQ& ql, const PQ& qn, const PairT<double>& wl, const PairT<double>& wn,
                          const double* gl, const double* gn, double* fSet) {
        double areaDifference, perimeterDifference;
        PQ qc;
        src.cell(c, areaDifference, perimeterDifference, qc);
        const PairT<double> wc = pair_terms(qs, qc);
        double t[3];
        fSet[0] = 0.0; fSet[1] = 0.0; fSet[2] = 0.0;
        const double cp = 2.0 * perimeterDifference;
        const double ca = 2.0 * Kr * areaDifference;
        fSet[0] -= cp * gl[0]; fSet[1] -= cp * gl[1]; fSet[2] -= cp * gl[2];
        fSet[0] -= cp * gn[0]; fSet[1] -= cp * gn[1]; fSet[2] -= cp * gn[2];
        grad_triangle_p(qs, ql, qc, wl, wc, tH, pH, R, t, ops);
        fSet[0] -= ca * t[0]; fSet[1] -= ca * t[1]; fSet[2] -= ca * t[2];
        grad_triangle_p(qs, qc, qn, wc, wn, tH, pH, R, t, ops);
        fSet[0] -= ca * t[0]; fSet[1] -= ca * t[1]; fSet[2] -= ca * t[2];
    };
    // cell rows of loop cells 0, 1, 2 (previous neighbour 0, 1, 2)
    const int c0 = L0 == 0 ? cl.x : (L1 == 0 ? cl.y : cl.z);
    const int c1 = L0 == 1 ? cl.x : (L1 == 1 ? cl.y : cl.z);
    const int c2 = L0 == 2 ? cl.x : (L1 == 2 ? cl.y : cl.z);
    double fa[3], fb[3], fc[3];
    cell_force(c0, q0, q1, w0, w1, g0, g1, fa);
    cell_force(c1, q1, q2, w1, w2, g1, g2, fb);
    cell_force(c2, q2, q0, w2, w0, g2, g0, fc);
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    int nan = 0;
#pragma unroll
    for (int sl = 0; sl < 3; ++sl) {
        const int j = sl == 0 ? L0 : (sl == 1 ? L1 : L2);
        const double* fSet = j == 0 ? fa : (j == 1 ? fb : fc);
        if (!(fSet[0] != fSet[0])) {
            f0 = f0 + fSet[0];
            f1 = f1 + fSet[1];
            f2 = f2 + fSet[2];
        } else {
            ++nan;
        }
    }
    f[0] = f0; f[1] = f1; f[2] = f2;
    return nan;
}

template <>
__device__ inline int vertex_force<double>(const Dev& d, const SysParams* S, const double* X, const double* Y,
                                           const double* Z, int v, double* f) {
    // the short root and quotient sequences first; a vertex with any NaN
    // contribution (an argument outside their range, or a NaN the reference
    // would skip too) is evaluated again with the operators themselves
    const int nan = vertex_force_f64(GlobalSrc{d, X, Y, Z}, d, S, v, f, FastOps{});
    if (__builtin_expect(nan != 0, 0)) {
        // an opaque copy of the row, so that the compiler does not keep the
        // first evaluation's loads live for this one (it reads its inputs again)
        int w = v;
        asm volatile("" : "+v"(w));
        return vertex_force_f64(GlobalSrc{d, X, Y, Z}, d, S, w, f);
    }
    return nan;
}

// The staged vertex pass's windows: SVM_STAGE_C cell rows from the least cell
// of the tile's vertices, and the tile's vertex rows with SVM_STAGE_M rows on
// either side.  In the canonical numbering (triangles sorted by their least
// cell, cells in spatial order) about 96% of a tile's cell references and 95%
// of its neighbour references fall inside.
#define SVM_STAGE_C 512
#define SVM_STAGE_M 64
#define SVM_STAGE_P (SVM_TILE + 2 * SVM_STAGE_M)
#define SVM_STAGE_DOUBLES (11 * SVM_STAGE_C + 3 * SVM_STAGE_P)

// Inputs from the LDS windows where they hold them, else from the arrays:
// each value's address is one or the other (a flat load either way, no
// branch), and the values are the same.
struct StagedSrc {
    const Dev& d;
    const double *X, *Y, *Z;
    const double* st;   // LDS: 11 planes of SVM_STAGE_C cell rows, then 3 planes of SVM_STAGE_P vertex rows
    int c_lo, p_lo;
    template <class Ops>
    __device__ PQ pq(int w, const Ops& ops) const {
        const unsigned k = (unsigned)(w - p_lo);
        const bool in = k < (unsigned)SVM_STAGE_P;
        const double* ps = st + 11 * SVM_STAGE_C;
        const double* qx = in ? ps + k : X + w;
        const double* qy = in ? ps + SVM_STAGE_P + k : Y + w;
        const double* qz = in ? ps + 2 * SVM_STAGE_P + k : Z + w;
        return make_pq<double>(*qx, *qy, *qz, ops);
    }
    __device__ void cell(int c, double& aD, double& pD, PQ& qc) const {
        const unsigned k = (unsigned)(c - c_lo);
        const bool in = k < (unsigned)SVM_STAGE_C;
        const size_t Cs = (size_t)d.Cs;
        auto at = [&](int plane, const double* g) { return *(in ? st + plane * SVM_STAGE_C + k : g + c); };
        aD = at(0, d.A) - at(1, d.a0);
        pD = at(2, d.P) - at(3, d.p0);
        qc.x = at(4, d.cx); qc.y = at(5, d.cy); qc.z = at(6, d.cz);
        qc.cosT = at(7, d.ctrig); qc.sinT = at(8, d.ctrig + Cs); qc.cosP = at(9, d.ctrig + 2 * Cs);
        qc.sinP = at(10, d.ctrig + 3 * Cs);
    }
};

// Fills the staged windows of vertex tile `tile` (whole block; every thread
// reaches the barriers).  Returns the window origins in c_lo, p_lo.
__device__ inline void stage_windows(const Dev& d, const SysParams* S, int tile, bool valid, int v, const double* X,
                                     const double* Y, const double* Z, double* st, int* s_min, int& c_lo,
                                     int& p_lo) {
    int cmin = 0x7fffffff;
    if (valid) {
        const int4 cl = d.vcn[v];
        cmin = min(cl.x, min(cl.y, cl.z));
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) cmin = min(cmin, __shfl_xor(cmin, o, 64));
    if ((threadIdx.x & 63) == 0) s_min[threadIdx.x >> 6] = cmin;
    __syncthreads();
    c_lo = min(min(s_min[0], s_min[1]), min(s_min[2], s_min[3]));
    p_lo = max(S->vo, tile * SVM_TILE - SVM_STAGE_M);
    const int c_end = S->co + S->nc, p_end = S->vo + S->nv;
    const size_t Cs = (size_t)d.Cs;
    const int c_rows = c_lo < c_end ? min(SVM_STAGE_C, c_end - c_lo) : 0;   // c_lo is INT_MAX if no vertex is valid
    for (int r = threadIdx.x; r < c_rows; r += SVM_TILE) {
        const int c = c_lo + r;
        {
            st[r] = d.A[c];
            st[SVM_STAGE_C + r] = d.a0[c];
            st[2 * SVM_STAGE_C + r] = d.P[c];
            st[3 * SVM_STAGE_C + r] = d.p0[c];
            st[4 * SVM_STAGE_C + r] = d.cx[c];
            st[5 * SVM_STAGE_C + r] = d.cy[c];
            st[6 * SVM_STAGE_C + r] = d.cz[c];
            st[7 * SVM_STAGE_C + r] = d.ctrig[c];
            st[8 * SVM_STAGE_C + r] = d.ctrig[Cs + c];
            st[9 * SVM_STAGE_C + r] = d.ctrig[2 * Cs + c];
            st[10 * SVM_STAGE_C + r] = d.ctrig[3 * Cs + c];
        }
    }
    double* ps = st + 11 * SVM_STAGE_C;
    for (int r = threadIdx.x; r < SVM_STAGE_P; r += SVM_TILE) {
        const int p = p_lo + r;
        if (p < p_end) {
            ps[r] = X[p];
            ps[SVM_STAGE_P + r] = Y[p];
            ps[2 * SVM_STAGE_P + r] = Z[p];
        }
    }
    __syncthreads();
}

// Force of the same energy from the vector-form gradients (grad_arc_v /
// grad_tri_v; no basis singularity at the poles), in T arithmetic from FP64
// positions and geometry: T = float is the FP32 force variant, T = double a
// well-conditioned FP64 variant.
template <class T>
__device__ inline int vertex_force_vec(const Dev& d, const SysParams* S, const double* X, const double* Y,
                                       const double* Z, int v, double* f) {
    const double R = S->R, iR = 1.0 / R;
    const T Kr = (T)S->kr, Rt = (T)R;
    const double px = X[v], py = Y[v], pz = Z[v];
    const V3<T> a{(T)(px * iR), (T)(py * iR), (T)(pz * iR)};
    auto rel = [&](double x, double y, double z) {
        return V3<T>{(T)((x - px) * iR), (T)((y - py) * iR), (T)((z - pz) * iR)};
    };
    int4 nb = d.vn[v];
    int4 cl = d.vcn[v];
    int ln = d.vln[v];
    // neighbours named so that each cell's previous -> next neighbour is
    // j -> j+1 (mod 3), as in vertex_force<double>
    const int ls0 = ln & 3, ns0 = (ln >> 2) & 3, ls1 = (ln >> 4) & 3, ls2 = (ln >> 8) & 3;
    const int o = ns0 == (ls0 == 2 ? 0 : ls0 + 1) ? 0 : 1;
    auto rn = [o](int k) { return o && k ? 3 - k : k; };
    const int L0 = rn(ls0), L1 = rn(ls1), L2 = rn(ls2);
    const int r1 = o ? nb.z : nb.y, r2 = o ? nb.y : nb.z;
    const V3<T> e0 = rel(X[nb.x], Y[nb.x], Z[nb.x]);
    const V3<T> e1 = rel(X[r1], Y[r1], Z[r1]);
    const V3<T> e2 = rel(X[r2], Y[r2], Z[r2]);
    const V3<T> g0 = grad_arc_v(a, e0), g1 = grad_arc_v(a, e1), g2 = grad_arc_v(a, e2);
    auto cell_force = [&](int c, const V3<T>& el, const V3<T>& en, const V3<T>& gl, const V3<T>& gn) {
        const T cp = T(2) * (T)(d.P[c] - d.p0[c]);
        const T ca = T(2) * Kr * (T)(d.A[c] - d.a0[c]);
        const V3<T> ec = rel(d.cx[c], d.cy[c], d.cz[c]);
        return (-cp) * (gl + gn) - ca * (grad_tri_v(a, el, ec, Rt) + grad_tri_v(a, ec, en, Rt));
    };
    const int c0 = L0 == 0 ? cl.x : (L1 == 0 ? cl.y : cl.z);
    const int c1 = L0 == 1 ? cl.x : (L1 == 1 ? cl.y : cl.z);
    const int c2 = L0 == 2 ? cl.x : (L1 == 2 ? cl.y : cl.z);
    const V3<T> fa = cell_force(c0, e0, e1, g0, g1);
    const V3<T> fb = cell_force(c1, e1, e2, g1, g2);
    const V3<T> fc = cell_force(c2, e2, e0, g2, g0);
    V3<T> F{T(0), T(0), T(0)};
    int nan = 0;
#pragma unroll
    for (int sl = 0; sl < 3; ++sl) {
        const int j = sl == 0 ? L0 : (sl == 1 ? L1 : L2);
        const V3<T> fs = j == 0 ? fa : (j == 1 ? fb : fc);
        // skipped like the reference's NaN contributions; here a degenerate
        // edge (coincident vertices) gives +-inf, which x - x turns into NaN
        if (!(fs.x - fs.x == T(0) && fs.y - fs.y == T(0) && fs.z - fs.z == T(0))) ++nan;
        else F = F + fs;
    }
    f[0] = (double)F.x; f[1] = (double)F.y; f[2] = (double)F.z;
    return nan;
}

template <>
__device__ inline int vertex_force<float>(const Dev& d, const SysParams* S, const double* X, const double* Y,
                                          const double* Z, int v, double* f) {
    return vertex_force_vec<float>(d, S, X, Y, Z, v, f);
}

// Force arithmetic selector of k_vertex: double (reference expressions,
// parity), float (vector form, FP32), VecF64 (vector form, FP64).
struct VecF64 {};
template <>
__device__ inline int vertex_force<VecF64>(const Dev& d, const SysParams* S, const double* X, const double* Y,
                                           const double* Z, int v, double* f) {
    return vertex_force_vec<double>(d, S, X, Y, Z, v, f);
}

// FAN = true: the total-area check and the fan fallback of the geometry
// (k_fan_fix) are done here, for spheres of at most 64 * SVM_FOLD_K cell
// tiles.  Every wavefront folds its sphere's tile partials itself; where the
// check fails, each thread gives the three cells of its vertex their fan
// areas (fan_area) before the force reads them.  Every thread whose vertex
// touches a cell stores the same value there before it loads it, so each
// load returns that value; every cell has vertices, so every cell's area is
// replaced, as k_fan_fix replaces it.
template <int MODE, class FT, bool FAN, bool STAGED = false>
__device__ __attribute__((always_inline)) inline void vertex_pass(const Dev& d, int pin, int pout, int nin, int nout,
                                                                 int store_forces, double* red, double* stage = nullptr,
                                                                 int* s_min = nullptr) {
    const int tile = lane_vtile(d);
    if (tile < 0) return;
    const int s = d.vtile_sys[tile];
    SysParams* S = d.sys + s;
    if (MODE == VM_FIRE && S->f_frozen) return;
    const int v = tile * SVM_TILE + threadIdx.x;
    const int local = v - S->vo;
    const bool valid = local < S->nv;
    const double* X = d.px[pin];
    const double* Y = d.py[pin];
    const double* Z = d.pz[pin];
    if (MODE == VM_ADAM && S->f_frozen) {   // a converged sphere keeps its positions
        if (valid) { d.px[pout][v] = X[v]; d.py[pout][v] = Y[v]; d.pz[pout][v] = Z[v]; }
        return;
    }
    const bool fan = FAN && fan_needed(S, fold_wave(d.cpart + S->ctile0, 1, S->nctiles, fold_width(S->nctiles)));
    int c_lo = 0, p_lo = 0;
    if (STAGED) stage_windows(d, S, tile, valid, v, X, Y, Z, stage, s_min, c_lo, p_lo);
    double f[3] = {0.0, 0.0, 0.0};
    double tp = 0.0, tf = 0.0, tv = 0.0;
    int nan = 0;
    if (valid) {
        if (FAN && fan) {   // rare: one loop body, not three inlined copies
            if (local == 0) S->n_fan += 1;
            const int4 cl = d.vcn[v];
#pragma unroll 1
            for (int k = 0; k < 3; ++k) {
                const int c = k == 0 ? cl.x : (k == 1 ? cl.y : cl.z);
                d.A[c] = fan_area(d, X, Y, Z, c, S->R);
            }
        }
        if constexpr (STAGED)
            nan = vertex_force_f64(StagedSrc{d, X, Y, Z, stage, c_lo, p_lo}, d, S, v, f);
        else
            nan = vertex_force<FT>(d, S, X, Y, Z, v, f);
        if (MODE == VM_FORCE || store_forces) {
            d.fx[v] = f[0]; d.fy[v] = f[1]; d.fz[v] = f[2];
        }
        if (MODE == VM_BROWNIAN || MODE == VM_ACTIVE) {
            double disp[3], nI[3];
            if (MODE == VM_BROWNIAN) {
                double forcePrefactor = S->dt * S->mu;
                double noisePrefactor = __builtin_sqrt(2.0 * forcePrefactor * S->T);
                double xi[3] = {0.0, 0.0, 0.0};
                if (S->T != 0.0) normal3(S->seed, S->step, (uint32_t)local, xi);
                for (int k = 0; k < 3; ++k) disp[k] = xi[k] * noisePrefactor + f[k] * forcePrefactor;
            } else {
                nI[0] = d.nx[nin][v]; nI[1] = d.ny[nin][v]; nI[2] = d.nz[nin][v];
                double dt = S->dt, v0 = S->v0, mu = S->mu;
                disp[0] = dt * (v0 * nI[0] + mu * f[0]);
                disp[1] = dt * (v0 * nI[1] + mu * f[1]);
                disp[2] = dt * (v0 * nI[2] + mu * f[2]);
            }
            double p[3] = {X[v], Y[v], Z[v]};
            move_on_sphere(p, disp, S->R);
            d.px[pout][v] = p[0];
            d.py[pout][v] = p[1];
            d.pz[pout][v] = p[2];
            if (MODE == VM_ACTIVE) {
                // the director's rotational-diffusion step at the moved
                // position; a vertex that a T1 flip then moves gets it redone
                // at its final position by the T1 commit
                double nO[3];
                director_update(S, local, p, nI, nO);
                d.nx[nout][v] = nO[0];
                d.ny[nout][v] = nO[1];
                d.nz[nout][v] = nO[2];
            }
        }
        if (MODE == VM_ADAM) {
            // adamStepCPU (src/updaters/energyMinimizerAdam.cpp:59-80), f the negative gradient
            const size_t Vs = (size_t)d.Vs;
            const double beta1 = S->a_beta1, beta2 = S->a_beta2;
            const double ic1 = 1.0 / (1.0 - S->a_beta1t), ic2 = 1.0 / (1.0 - S->a_beta2t);
            double disp[3];
            for (int k = 0; k < 3; ++k) {
                const double mm = beta1 * d.am[k * Vs + v] + (beta1 - 1) * f[k];
                const double vv = beta2 * d.av[k * Vs + v] + (1 - beta2) * (f[k] * f[k]);
                d.am[k * Vs + v] = mm;
                d.av[k * Vs + v] = vv;
                const double mc = ic1 * mm;
                const double vc = ic2 * vv;
                double rootvc = __builtin_sqrt(vc);
                if (rootvc == 0) rootvc = S->a_eps;
                disp[k] = -S->a_dt * mc / (rootvc);
            }
            tf = dot0(f, f);
            double p[3] = {X[v], Y[v], Z[v]};
            move_on_sphere(p, disp, S->R);
            d.px[pout][v] = p[0];
            d.py[pout][v] = p[1];
            d.pz[pout][v] = p[2];
        }
        if (MODE == VM_FIRE) {
            double dt = S->f_dt;
            double h = (0.5 / 1.0) * dt;
            double vel[3] = {d.vx[v], d.vy[v], d.vz[v]};
            vel[0] += h * f[0]; vel[1] += h * f[1]; vel[2] += h * f[2];
            d.vx[v] = vel[0]; d.vy[v] = vel[1]; d.vz[v] = vel[2];
            d.fx[v] = f[0]; d.fy[v] = f[1]; d.fz[v] = f[2];
            tp = dot0(f, vel);
            tf = dot0(f, f);
            tv = dot0(vel, vel);
        }
    }
    if (nan) atomicAdd((unsigned long long*)&S->n_nan, (unsigned long long)nan);
    if (MODE == VM_ADAM) {
        double sf = block_tree256(tf, red);
        if (threadIdx.x == 0) d.vpart[3 * tile + 1] = sf;
    }
    if (MODE == VM_FIRE) {
        double sp = block_tree256(tp, red);
        double sf = block_tree256(tf, red);
        double sv = block_tree256(tv, red);
        if (threadIdx.x == 0) {
            d.vpart[3 * tile + 0] = sp;
            d.vpart[3 * tile + 1] = sf;
            d.vpart[3 * tile + 2] = sv;
        }
    }
}

template <int MODE, class FT, bool FAN = false>
__global__ __launch_bounds__(SVM_TILE) void k_vertex(Dev d, int pin, int pout, int nin, int nout, int store_forces) {
    __shared__ double red[SVM_TILE + 1];
    vertex_pass<MODE, FT, FAN>(d, pin, pout, nin, nout, store_forces, red);
}

// The same pass with its inputs staged in LDS (the vertex form "staged";
// Brownian and active steps in the FP64 parity arithmetic): the tile's cell
// window and vertex window loaded once, with coalesced loads, by the whole
// workgroup; the force then reads them from LDS (StagedSrc).
template <int MODE>
__global__ __launch_bounds__(SVM_TILE) void k_vertex_staged(Dev d, int pin, int pout, int nin, int nout) {
    __shared__ double red[SVM_TILE + 1];
    __shared__ double stage[SVM_STAGE_DOUBLES];
    __shared__ int s_min[SVM_TILE / 64];
    vertex_pass<MODE, double, false, true>(d, pin, pout, nin, nout, 0, red, stage, s_min);
}

// The same pass compiled for three wavefronts per SIMD (at most 168 VGPRs,
// against the 212 the pass takes unconstrained: the compiler spills the
// rest), an autotuner form of the Brownian and active steps.
template <int MODE, bool FAN = false>
__global__ __launch_bounds__(SVM_TILE) __attribute__((amdgpu_waves_per_eu(3))) void k_vertex_w3(
    Dev d, int pin, int pout, int nin, int nout) {
    __shared__ double red[SVM_TILE + 1];
    vertex_pass<MODE, double, FAN>(d, pin, pout, nin, nout, 0, red);
}

// Corner form of k_vertex for the Brownian and active updaters in the parity
// arithmetic.  A block takes 64 vertices (a quarter of a 256-row tile) with
// three wavefronts and works in three phases:
//   1. wavefront k evaluates, for every vertex, the terms of its neighbour
//      slot k that vertex_force<double> shares between two cells (the
//      neighbour's spherical coordinates, the pair terms with the vertex and
//      the arc-length gradient) and leaves them in LDS;
//   2. wavefront k evaluates the contribution of cell slot k (the body of the
//      reference's per-cell loop, sphericalVertexModel.cpp:386-413) from them;
//   3. wavefront 0 adds the three contributions in slot order, skipping NaN
//      ones as the reference does, and applies the updater.
// Every value is computed by the same operations on the same inputs as in
// k_vertex, so the results are identical bit for bit.  A thread carries the
// terms of one neighbour and one cell instead of three of each, which leaves
// room for four or five wavefronts per SIMD (WAVES: at most 128 or 102
// VGPRs) to hide the gathers' latency.
#define SVM_CORNER_V 64
__device__ inline int corner_block(const Dev& d) {
    const int nq = d.vtn * (SVM_TILE / SVM_CORNER_V);
    int q = (int)blockIdx.x;
    if (d.xcd) {
        const int per = (nq + 7) >> 3;
        q = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    }
    return q < nq ? q : -1;
}

template <int MODE, int WAVES>
__global__ __launch_bounds__(3 * SVM_CORNER_V) __attribute__((amdgpu_waves_per_eu(WAVES))) void k_vertex_corner(
    Dev d, int pin, int pout, int nin, int nout) {
    // per neighbour slot: cosT, sinT, cosP, sinP, the pair terms sinPd,
    // cosPd, s, dd, and the arc gradient (3); the neighbour's coordinates are
    // read again from memory (cache-resident after phase 1)
    __shared__ double nsh[3][11][SVM_CORNER_V];
    __shared__ double fsh[3][3][SVM_CORNER_V];
    __shared__ int oksh[3][SVM_CORNER_V];
    const int q = corner_block(d);
    if (q < 0) return;
    const int tile = d.vt0 + q / (SVM_TILE / SVM_CORNER_V);
    const int lane = threadIdx.x & (SVM_CORNER_V - 1);
    const int slot = threadIdx.x / SVM_CORNER_V;   // the same for a whole wavefront
    SysParams* S = d.sys + d.vtile_sys[tile];
    const int v = tile * SVM_TILE + (q % (SVM_TILE / SVM_CORNER_V)) * SVM_CORNER_V + lane;
    const int local = v - S->vo;
    const bool valid = local < S->nv;
    const double* X = d.px[pin];
    const double* Y = d.py[pin];
    const double* Z = d.pz[pin];
    PQ qs{};
    double tH[3] = {0.0, 0.0, 0.0}, pH[3] = {0.0, 0.0, 0.0};
    if (valid) {
        qs = pq_at<double>(X, Y, Z, v);
        basis(qs, tH, pH);
        const int4 nb = d.vn[v];
        const int j = slot == 0 ? nb.x : (slot == 1 ? nb.y : nb.z);
        const PQ qj = pq_at<double>(X, Y, Z, j);
        const PairT<double> w = pair_terms(qs, qj);
        double g[3];
        grad_geodesic_p(qs, qj, w, tH, pH, g);
        double* o = &nsh[slot][0][lane];
        o[0 * SVM_CORNER_V] = qj.cosT; o[1 * SVM_CORNER_V] = qj.sinT;
        o[2 * SVM_CORNER_V] = qj.cosP; o[3 * SVM_CORNER_V] = qj.sinP;
        o[4 * SVM_CORNER_V] = w.sinPd; o[5 * SVM_CORNER_V] = w.cosPd; o[6 * SVM_CORNER_V] = w.s;
        o[7 * SVM_CORNER_V] = w.dd;
        o[8 * SVM_CORNER_V] = g[0]; o[9 * SVM_CORNER_V] = g[1]; o[10 * SVM_CORNER_V] = g[2];
    }
    __syncthreads();
    if (valid) {
        const double R = S->R, Kr = S->kr;
        const int4 cl = d.vcn[v];
        const int4 nb = d.vn[v];
        const int ln = d.vln[v];
        const int c = slot == 0 ? cl.x : (slot == 1 ? cl.y : cl.z);
        // neighbour slots of the cell's previous and next ring vertices; a
        // field value of 3 selects slot 2, as pick3 does in k_vertex
        const int ls = min((ln >> (4 * slot)) & 3, 2);
        const int ns = min((ln >> (4 * slot + 2)) & 3, 2);
        const int jl = ls == 0 ? nb.x : (ls == 1 ? nb.y : nb.z);
        const int jn = ns == 0 ? nb.x : (ns == 1 ? nb.y : nb.z);
        const size_t Cs = (size_t)d.Cs;
        const double areaDifference = d.A[c] - d.a0[c];
        const double perimeterDifference = d.P[c] - d.p0[c];
        PQ qc;
        qc.x = d.cx[c]; qc.y = d.cy[c]; qc.z = d.cz[c];
        qc.cosT = d.ctrig[c]; qc.sinT = d.ctrig[Cs + c]; qc.cosP = d.ctrig[2 * Cs + c]; qc.sinP = d.ctrig[3 * Cs + c];
        auto load_pq = [&](int k, int j) {
            const double* o = &nsh[k][0][lane];
            PQ r;
            r.x = X[j]; r.y = Y[j]; r.z = Z[j];
            r.cosT = o[0 * SVM_CORNER_V]; r.sinT = o[1 * SVM_CORNER_V];
            r.cosP = o[2 * SVM_CORNER_V]; r.sinP = o[3 * SVM_CORNER_V];
            return r;
        };
        auto load_pair = [&](int k) {
            const double* o = &nsh[k][0][lane];
            PairT<double> r;
            r.sinPd = o[4 * SVM_CORNER_V]; r.cosPd = o[5 * SVM_CORNER_V]; r.s = o[6 * SVM_CORNER_V];
            r.dd = o[7 * SVM_CORNER_V];
            return r;
        };
        const PairT<double> wc = pair_terms(qs, qc);
        double fSet[3] = {0.0, 0.0, 0.0}, t[3];
        const double cp = 2.0 * perimeterDifference;
        const double ca = 2.0 * Kr * areaDifference;
        {
            const double* gl = &nsh[ls][8][lane];
            const double* gn = &nsh[ns][8][lane];
            fSet[0] -= cp * gl[0]; fSet[1] -= cp * gl[SVM_CORNER_V]; fSet[2] -= cp * gl[2 * SVM_CORNER_V];
            fSet[0] -= cp * gn[0]; fSet[1] -= cp * gn[SVM_CORNER_V]; fSet[2] -= cp * gn[2 * SVM_CORNER_V];
        }
        {
            const PQ ql = load_pq(ls, jl);
            grad_triangle_p(qs, ql, qc, load_pair(ls), wc, tH, pH, R, t);
        }
        fSet[0] -= ca * t[0]; fSet[1] -= ca * t[1]; fSet[2] -= ca * t[2];
        {
            const PQ qn = load_pq(ns, jn);
            grad_triangle_p(qs, qc, qn, wc, load_pair(ns), tH, pH, R, t);
        }
        fSet[0] -= ca * t[0]; fSet[1] -= ca * t[1]; fSet[2] -= ca * t[2];
        fsh[slot][0][lane] = fSet[0];
        fsh[slot][1][lane] = fSet[1];
        fsh[slot][2][lane] = fSet[2];
        oksh[slot][lane] = !(fSet[0] != fSet[0]);
    }
    __syncthreads();
    if (slot != 0 || !valid) return;
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    int nan = 0;
#pragma unroll
    for (int sl = 0; sl < 3; ++sl) {
        if (oksh[sl][lane]) {
            f0 = f0 + fsh[sl][0][lane];
            f1 = f1 + fsh[sl][1][lane];
            f2 = f2 + fsh[sl][2][lane];
        } else {
            ++nan;
        }
    }
    const double f[3] = {f0, f1, f2};
    double disp[3], nI[3];
    if (MODE == VM_BROWNIAN) {
        double forcePrefactor = S->dt * S->mu;
        double noisePrefactor = svm::sqrt_rn(2.0 * forcePrefactor * S->T);
        double xi[3] = {0.0, 0.0, 0.0};
        if (S->T != 0.0) normal3(S->seed, S->step, (uint32_t)local, xi);
        for (int k = 0; k < 3; ++k) disp[k] = xi[k] * noisePrefactor + f[k] * forcePrefactor;
    } else {
        nI[0] = d.nx[nin][v]; nI[1] = d.ny[nin][v]; nI[2] = d.nz[nin][v];
        double dt = S->dt, v0 = S->v0, mu = S->mu;
        disp[0] = dt * (v0 * nI[0] + mu * f[0]);
        disp[1] = dt * (v0 * nI[1] + mu * f[1]);
        disp[2] = dt * (v0 * nI[2] + mu * f[2]);
    }
    double p[3] = {X[v], Y[v], Z[v]};
    move_on_sphere(p, disp, S->R);
    d.px[pout][v] = p[0];
    d.py[pout][v] = p[1];
    d.pz[pout][v] = p[2];
    if (MODE == VM_ACTIVE) {
        double nO[3];
        director_update(S, local, p, nI, nO);
        d.nx[nout][v] = nO[0];
        d.ny[nout][v] = nO[1];
        d.nz[nout][v] = nO[2];
    }
    if (nan) atomicAdd((unsigned long long*)&S->n_nan, (unsigned long long)nan);
}

// The FIRE step's reductions and controller (energyMinimizerFIRE::fireStepCPU,
// src/updaters/energyMinimizerFIRE.cpp:181-233), one block per sphere, after
// k_vertex<VM_FIRE> wrote the per-tile partial sums of F.v, F.F and v.v.
__global__ __launch_bounds__(SVM_TILE) void k_fire_control(Dev d) {
    __shared__ double buf[SVM_MAX_TILES];
    SysParams* S = d.sys + d.s0 + blockIdx.x;
    if (S->f_frozen) return;
    double Power = fold_lds(d.vpart + 3 * S->vtile0 + 0, 3, S->nvtiles, buf);
    double forceNorm = fold_lds(d.vpart + 3 * S->vtile0 + 1, 3, S->nvtiles, buf);
    double velocityNorm = fold_lds(d.vpart + 3 * S->vtile0 + 2, 3, S->nvtiles, buf);
    if (threadIdx.x == 0) {
        double Ndof = (double)S->nv;
        double forceMax = S->f_stop_rms ? svm::sqrt_rn(forceNorm / Ndof) : svm::sqrt_rn(forceNorm) / Ndof;
        double scaling = 0.0;
        if (forceNorm > 0.) scaling = svm::sqrt_rn(velocityNorm / forceNorm);
        double alpha = S->f_alpha, dtf = S->f_dt;
        S->mix_alpha = alpha;
        S->mix_scaling = scaling;
        S->mix_pending = 1;
        S->mix_zero = 0;
        if (Power > 0) {
            if (S->f_nsince > S->f_nmin) {
                double a = dtf * S->f_dt_inc;
                dtf = (S->f_dt_max < a) ? S->f_dt_max : a;
                alpha = alpha * S->f_alpha_dec;
                alpha = (alpha < S->f_alpha_min) ? S->f_alpha_min : alpha;
            }
            S->f_nsince += 1;
        } else {
            S->f_nsince = 0;
            dtf = dtf * S->f_dt_dec;
            dtf = (dtf < S->f_dt_min) ? S->f_dt_min : dtf;
            alpha = S->f_alpha_start;
            S->mix_zero = 1;
        }
        S->f_dt = dtf;
        S->f_alpha = alpha;
        S->f_iter += 1;
        S->f_force_max = forceMax;
        S->f_power = Power;
        S->f_frozen = !((S->f_iter < S->f_max_iter) && (forceMax > S->f_cutoff));
    }
}

// The end of an Adam iteration (energyMinimizerAdam.cpp:79-82, :90-95), one
// block per sphere, after k_vertex<VM_ADAM> wrote the per-tile partial sums
// of F.F: forceMax = sqrt(sum F^2) / N, the bias-correction powers, and the
// loop condition (iterations < max and forceMax > cutoff).
__global__ __launch_bounds__(SVM_TILE) void k_adam_control(Dev d) {
    __shared__ double buf[SVM_MAX_TILES];
    SysParams* S = d.sys + d.s0 + blockIdx.x;
    if (S->f_frozen) return;
    double forceNorm = fold_lds(d.vpart + 3 * S->vtile0 + 1, 3, S->nvtiles, buf);
    if (threadIdx.x == 0) {
        double forceMax = svm::sqrt_rn(forceNorm) / (double)S->nv;
        S->a_iter += 1;
        S->a_force_max = forceMax;
        S->a_beta1t *= S->a_beta1;
        S->a_beta2t *= S->a_beta2;
        S->f_frozen = !((S->a_iter < S->a_max_iter) && (forceMax > S->a_cutoff));
    }
}

// dynamicalFeatures::computeMSD / computeOverlapFunction
// (src/analysis/dynamicalFeatures.cpp:17-53) on the cell positions against
// the reference snapshot d.dref, for the first floor(N * fraction) cells of
// each sphere (all of them when fraction >= 1): per-tile tree partials of the
// squared geodesic displacement (NaN adds nothing) and of the count below the
// cutoff, then one block per sphere folds them.
__device__ inline int dyn_count(const SysParams* S, double fraction) {
    return fraction < 1 ? (int)__builtin_floor(S->nc * fraction) : S->nc;
}
__global__ __launch_bounds__(SVM_TILE) void k_dyn_tiles(Dev d, double fraction, double cutoff) {
    __shared__ double red[SVM_TILE + 1];
    const int tile = blockIdx.x;
    const SysParams* S = d.sys + d.ctile_sys[tile];
    const int c = tile * SVM_TILE + threadIdx.x;
    const size_t Cs = (size_t)d.Cs;
    double sq = 0.0, ov = 0.0;
    if (c - S->co < dyn_count(S, fraction)) {
        const double init[3] = {d.dref[c], d.dref[Cs + c], d.dref[2 * Cs + c]};
        const double cur[3] = {d.cx[c], d.cy[c], d.cz[c]};
        const double disp = geodesic(init, cur, pnorm(init), pnorm(cur), S->R);
        sq = disp != disp ? 0.0 : disp * disp;
        ov = disp < cutoff ? 1.0 : 0.0;
    }
    const double a = block_tree256(sq, red);
    const double b = block_tree256(ov, red);
    if (threadIdx.x == 0) {
        d.dpart[2 * tile] = a;
        d.dpart[2 * tile + 1] = b;
    }
}
__global__ __launch_bounds__(SVM_TILE) void k_dyn_fold(Dev d, double fraction) {
    __shared__ double buf[SVM_MAX_TILES];
    const SysParams* S = d.sys + blockIdx.x;
    const double msd = fold_lds(d.dpart + 2 * S->ctile0, 2, S->nctiles, buf);
    const double overlap = fold_lds(d.dpart + 2 * S->ctile0 + 1, 2, S->nctiles, buf);
    if (threadIdx.x == 0) {
        const int n = dyn_count(S, fraction);
        d.dout[2 * blockIdx.x] = msd / n;
        d.dout[2 * blockIdx.x + 1] = overlap / n;
    }
}

// Pending FIRE velocity mixing (fireStepCPU :205-208, :227-231) of vertex v.
__device__ inline void fire_mix(const Dev& d, const SysParams* S, int v, double* vel, const double* f) {
    if (S->mix_pending) {
        if (S->mix_zero) {
            vel[0] = 0.0; vel[1] = 0.0; vel[2] = 0.0;
        } else {
            double a = 1.0 - S->mix_alpha;
            double b = S->mix_alpha * S->mix_scaling;
            vel[0] = a * vel[0] + b * f[0];
            vel[1] = a * vel[1] + b * f[1];
            vel[2] = a * vel[2] + b * f[2];
        }
    }
}

// First half of velocityVerlet::integrateEOMCPU (src/updaters/velocityVerlet.cpp:14-23)
// preceded by the previous iteration's FIRE mixing, then the move.
__global__ __launch_bounds__(SVM_TILE) void k_fire_vv1(Dev d, int pin, int pout) {
    const int tile = lane_vtile(d);
    if (tile < 0) return;
    const int s = d.vtile_sys[tile];
    SysParams* S = d.sys + s;
    const int v = tile * SVM_TILE + threadIdx.x;
    const int local = v - S->vo;
    if (local >= S->nv) return;
    double p[3] = {d.px[pin][v], d.py[pin][v], d.pz[pin][v]};
    if (!S->f_frozen) {
        double f[3] = {d.fx[v], d.fy[v], d.fz[v]};
        double vel[3] = {d.vx[v], d.vy[v], d.vz[v]};
        fire_mix(d, S, v, vel, f);
        double dt = S->f_dt;
        double disp[3];
        double hh = 0.5 * dt * dt;
        double h = (0.5 / 1.0) * dt;
        for (int k = 0; k < 3; ++k) {
            disp[k] = dt * vel[k] + hh * f[k];
            vel[k] += h * f[k];
        }
        d.vx[v] = vel[0]; d.vy[v] = vel[1]; d.vz[v] = vel[2];
        move_on_sphere(p, disp, S->R);
    }
    d.px[pout][v] = p[0];
    d.py[pout][v] = p[1];
    d.pz[pout][v] = p[2];
}

// Apply a pending FIRE mixing (end of a minimisation).
__global__ __launch_bounds__(SVM_TILE) void k_fire_finalize(Dev d) {
    const int tile = blockIdx.x;
    const SysParams* S = d.sys + d.vtile_sys[tile];
    const int v = tile * SVM_TILE + threadIdx.x;
    if (v - S->vo >= S->nv || !S->mix_pending) return;
    double f[3] = {d.fx[v], d.fy[v], d.fz[v]};
    double vel[3] = {d.vx[v], d.vy[v], d.vz[v]};
    fire_mix(d, S, v, vel, f);
    d.vx[v] = vel[0]; d.vy[v] = vel[1]; d.vz[v] = vel[2];
}

// External displacement (moveParticles(displacements, scale),
// src/model/sphericalVertexModel.cpp:111-131) for sphere `only` (-1: all;
// -2: all whose minimiser has not converged).  Other spheres keep their
// positions bit for bit.
__global__ __launch_bounds__(SVM_TILE) void k_move_ext(Dev d, int pin, int pout, int only, double scale) {
    const int tile = blockIdx.x;
    const int s = d.vtile_sys[tile];
    const SysParams* S = d.sys + s;
    const int v = tile * SVM_TILE + threadIdx.x;
    if (v - S->vo >= S->nv) return;
    double p[3] = {d.px[pin][v], d.py[pin][v], d.pz[pin][v]};
    if (only == -1 || only == s || (only == -2 && !minimiser_converged(S))) {
        double disp[3] = {d.disp[3 * (size_t)v], d.disp[3 * (size_t)v + 1], d.disp[3 * (size_t)v + 2]};
        if (scale != 1.) { disp[0] = scale * disp[0]; disp[1] = scale * disp[1]; disp[2] = scale * disp[2]; }
        move_on_sphere(p, disp, S->R);
    }
    d.px[pout][v] = p[0];
    d.py[pout][v] = p[1];
    d.pz[pout][v] = p[2];
}


// Director update of sphericalVectorialVicsek::integrateEOMCPU
// (src/updaters/sphericalVectorialVicsek.cpp:24-64) on the vertex model, after
// the move and the T1 pass: a random unit tangent direction (uniform on the
// sphere, projected), the mean of the vertex's and its three neighbours'
// directors plus eta times that direction, projected and normalised; then the
// director relaxes towards it at rate dt / tau and is put back in the tangent
// plane with unit length.  Reads directors nin, writes nout; the draws are
// keyed by the step count after this step.
__global__ __launch_bounds__(SVM_TILE) void k_vicsek_dir(Dev d, int pin, int nin, int nout) {
    const int tile = lane_vtile(d);
    if (tile < 0) return;
    const SysParams* S = d.sys + d.vtile_sys[tile];
    const int v = tile * SVM_TILE + threadIdx.x;
    const int local = v - S->vo;
    if (local >= S->nv) return;
    const double p[3] = {d.px[pin][v], d.py[pin][v], d.pz[pin][v]};
    double u, w;
    uniform2(S->seed, STREAM_VICSEK, S->step, (uint32_t)local, &u, &w);
    const double phi = 2.0 * SVM_PI * u;
    const double theta = svm::acos(2.0 * w - 1);
    double sp[3] = {1.0 * svm::sin(theta) * svm::cos(phi), 1.0 * svm::sin(theta) * svm::sin(phi),
                    1.0 * svm::cos(theta)};
    {   // projectToTangentPlane, then * (1 / norm)
        double dvn = 0.0; dvn += sp[0] * p[0]; dvn += sp[1] * p[1]; dvn += sp[2] * p[2];
        double dnn = 0.0; dnn += p[0] * p[0]; dnn += p[1] * p[1]; dnn += p[2] * p[2];
        const double q = dvn / dnn;
        sp[0] = sp[0] - q * p[0]; sp[1] = sp[1] - q * p[1]; sp[2] = sp[2] - q * p[2];
        double ss = 0.0; ss += sp[0] * sp[0]; ss += sp[1] * sp[1]; ss += sp[2] * sp[2];
        const double inv = 1.0 / svm::sqrt_rn(ss);
        sp[0] = inv * sp[0]; sp[1] = inv * sp[1]; sp[2] = inv * sp[2];
    }
    const double* NX = d.nx[nin];
    const double* NY = d.ny[nin];
    const double* NZ = d.nz[nin];
    const int4 nb = d.vn[v];
    const double n0[3] = {NX[v], NY[v], NZ[v]};
    double nd[3] = {n0[0], n0[1], n0[2]};
    const int nbs[3] = {nb.x, nb.y, nb.z};
    for (int jj = 0; jj < 3; ++jj) {
        nd[0] += NX[nbs[jj]]; nd[1] += NY[nbs[jj]]; nd[2] += NZ[nbs[jj]];
    }
    const double mi = 1.0 / 4;
    const double eta = S->eta;
    for (int k = 0; k < 3; ++k) nd[k] = mi * nd[k] + eta * sp[k];
    project_normalize(nd, p);
    const double rate = S->dt / S->tau;
    double n[3];
    for (int k = 0; k < 3; ++k) n[k] = n0[k] + rate * (nd[k] - n0[k]);
    {   // projectToTangentPlane, then * (1 / norm)
        double dvn = 0.0; dvn += n[0] * p[0]; dvn += n[1] * p[1]; dvn += n[2] * p[2];
        double dnn = 0.0; dnn += p[0] * p[0]; dnn += p[1] * p[1]; dnn += p[2] * p[2];
        const double q = dvn / dnn;
        n[0] = n[0] - q * p[0]; n[1] = n[1] - q * p[1]; n[2] = n[2] - q * p[2];
        double ss = 0.0; ss += n[0] * n[0]; ss += n[1] * n[1]; ss += n[2] * n[2];
        const double inv = 1.0 / svm::sqrt_rn(ss);
        n[0] = inv * n[0]; n[1] = inv * n[1]; n[2] = inv * n[2];
    }
    d.nx[nout][v] = n[0]; d.ny[nout][v] = n[1]; d.nz[nout][v] = n[2];
}

// ------------------------------------------------------------- energies ---
// Potential energy per sphere (computeEnergy, sphericalVertexModel.cpp:150-160)
// and kinetic energy (simpleModel.cpp:63-73), tree-reduced.
__global__ __launch_bounds__(SVM_TILE) void k_energy_cells(Dev d) {
    __shared__ double red[SVM_TILE + 1];
    const int tile = blockIdx.x;
    SysParams* S = d.sys + d.ctile_sys[tile];
    const int c = tile * SVM_TILE + threadIdx.x;
    double e = 0.0;
    if (c - S->co < S->nc) {
        double da = d.A[c] - d.a0[c], dp = d.P[c] - d.p0[c];
        e = S->kr * da * da + dp * dp;
    }
    double part = block_tree256(e, red);
    if (threadIdx.x == 0) d.cpart[tile] = part;
}

__global__ __launch_bounds__(SVM_TILE) void k_energy_vertices(Dev d) {
    __shared__ double red[SVM_TILE + 1];
    const int tile = blockIdx.x;
    SysParams* S = d.sys + d.vtile_sys[tile];
    const int v = tile * SVM_TILE + threadIdx.x;
    double e = 0.0;
    if (v - S->vo < S->nv) {
        double vel[3] = {d.vx[v], d.vy[v], d.vz[v]};
        e = 0.5 * 1.0 * dot0(vel, vel);
    }
    double part = block_tree256(e, red);
    if (threadIdx.x == 0) d.vpart[3 * tile] = part;
}

__global__ __launch_bounds__(SVM_TILE) void k_energy_fold(Dev d) {
    __shared__ double buf[SVM_MAX_TILES];
    SysParams* S = d.sys + blockIdx.x;
    double pot = fold_lds(d.cpart + S->ctile0, 1, S->nctiles, buf);
    double kin = fold_lds(d.vpart + 3 * S->vtile0, 3, S->nvtiles, buf);
    if (threadIdx.x == 0) {
        S->energy_pot = pot;
        S->energy_kin = kin;
    }
}

}  // namespace svm

#include "svm_t1.h"
