// svm_t1.h -- T1 topology maintenance on the device, in exactly the order of
// the reference's sequential CPU loop (testAndPerformT1TransitionsCPU,
// src/model/sphericalVertexModel.cpp:439-604): vertex1 ascending, slots 0..2,
// every test seeing all earlier flips.
//
// Three launches per step:
//   k_t1_detect   thread per vertex: short-edge test on the moved positions
//                 (bit mask + compact candidate list), active-director update.
//   k_t1_cluster  one wavefront per candidate vertex u: a dry run of the
//                 sequential loop restricted to u -- u's visit, then the
//                 visits of every later vertex its flips disturbed, in index
//                 order -- on a private LDS copy of the data it touches,
//                 stamping each touched vertex and cell, and logging what it
//                 would write.  When no two candidates of a sphere touched a
//                 common item, their visits commute and any order gives the
//                 sequential result.
//   k_t1_resolve  block per sphere: applies the logs of a sphere whose
//                 candidates did not overlap; otherwise (or when a buffer was
//                 too small) runs the sequential loop itself over the
//                 candidates and disturbed vertices.
// No launch needs a device-scope fence: each kernel consumes what the
// previous one wrote.
// A vertex that is neither a candidate nor disturbed by an earlier flip has
// an unchanged neighbourhood since detection, so its visit is a no-op and
// skipping it is exact.
#pragma once

namespace svm {

#define SVM_T1_MAXCL 4096     // candidates handled by dry runs per step
#define SVM_T1_WV 64          // vertices cached per dry run
#define SVM_T1_WC 32          // cells cached per dry run
#define SVM_T1_RC 32          // ring entries per cached cell
#define SVM_T1_LV 32          // modified vertices logged per dry run
#define SVM_T1_LC 16          // modified cells logged per dry run
#define SVM_T1_WAVES 4        // dry-run wavefronts per block

struct T1Log {
    int next, nvert, ncell, pad;
    long long flips, skipped;
    int vid[SVM_T1_LV];
    int vflag[SVM_T1_LV];
    double vpos[SVM_T1_LV][3];
    int vvn[SVM_T1_LV][3];
    int vvcn[SVM_T1_LV][3];
    int cid[SVM_T1_LC];
    int ccnt[SVM_T1_LC];
    int cring[SVM_T1_LC][SVM_T1_RC];
};

__device__ inline int& ringref(const Dev& d, int c, int k) { return d.ring[(size_t)k * d.Cs + c]; }
__device__ inline int& vcnref(const Dev& d, int v, int k) { return ((int*)&d.vcn[v])[k]; }
__device__ inline int& vnref(const Dev& d, int v, int k) { return ((int*)&d.vn[v])[k]; }

// Min-heap of local vertex indices (the disturbed vertices still to visit).
struct Heap {
    int* a;
    int n;
    __device__ void push(int x) {
        int i = n++;
        while (i > 0) {
            int p = (i - 1) >> 1;
            if (a[p] <= x) break;
            a[i] = a[p];
            i = p;
        }
        a[i] = x;
    }
    __device__ int pop() {
        int top = a[0];
        int x = a[--n];
        int i = 0;
        while (true) {
            int l = 2 * i + 1;
            if (l >= n) break;
            int r = l + 1;
            int m = (r < n && a[r] < a[l]) ? r : l;
            if (a[m] >= x) break;
            a[i] = a[m];
            i = m;
        }
        if (n > 0) a[i] = x;
        return top;
    }
};

// ------------------------------------------------------------ accessors ---
// Direct access to the topology and positions in HBM.
struct GlobalAcc {
    const Dev* d;
    int pcur;
    __device__ void pos(int v, double* p) { p[0] = d->px[pcur][v]; p[1] = d->py[pcur][v]; p[2] = d->pz[pcur][v]; }
    __device__ void set_pos(int v, const double* p) {
        d->px[pcur][v] = p[0]; d->py[pcur][v] = p[1]; d->pz[pcur][v] = p[2];
    }
    __device__ int vn(int v, int k) { return vnref(*d, v, k); }
    __device__ void set_vn(int v, int k, int x) { vnref(*d, v, k) = x; }
    __device__ int vcn(int v, int k) { return vcnref(*d, v, k); }
    __device__ void set_vcn(int v, int k, int x) { vcnref(*d, v, k) = x; }
    __device__ int cnn(int c) { return d->cnn[c]; }
    __device__ void set_cnn(int c, int x) { d->cnn[c] = x; }
    __device__ int ring(int c, int k) { return ringref(*d, c, k); }
    __device__ void set_ring(int c, int k, int x) { ringref(*d, c, k) = x; }
    __device__ bool failed() { return false; }
};

// Private LDS copy of the vertices and cells one dry run touches (open
// addressing on the global row), filled from HBM on first touch.  Every first
// touch stamps the item with this dry run's (generation, id); meeting another
// dry run's stamp of the same generation marks the sphere as overlapping.
struct CacheAcc {
    const Dev* d;
    int pcur;
    unsigned long long stamp;
    int32_t* overlap;
    bool fail;
    int* vkey;        // [WV]
    int* vflag;       // [WV]  1: position written, 2: lists written
    double* vp;       // [WV][3]
    int* vvn;         // [WV][3]
    int* vvcn;        // [WV][3]
    int* ckey;        // [WC]
    int* cflag;       // [WC]
    int* ccnt;        // [WC]
    int* cring;       // [WC][RC]

    __device__ void claim(unsigned long long* own) {
        unsigned long long old = atomicMax(own, stamp);
        if ((old >> 32) == (stamp >> 32) && old != stamp) *overlap = 1;
    }
    __device__ int vslot(int v) {
        unsigned h = ((unsigned)v * 2654435761u) >> (32 - 6);
        for (int probe = 0; probe < SVM_T1_WV; ++probe, h = (h + 1) & (SVM_T1_WV - 1)) {
            if (vkey[h] == v) return h;
            if (vkey[h] < 0) {
                vkey[h] = v;
                vflag[h] = 0;
                vp[3 * h] = d->px[pcur][v];
                vp[3 * h + 1] = d->py[pcur][v];
                vp[3 * h + 2] = d->pz[pcur][v];
                int4 a = d->vn[v], b = d->vcn[v];
                vvn[3 * h] = a.x; vvn[3 * h + 1] = a.y; vvn[3 * h + 2] = a.z;
                vvcn[3 * h] = b.x; vvcn[3 * h + 1] = b.y; vvcn[3 * h + 2] = b.z;
                claim(d->vown + v);
                return h;
            }
        }
        fail = true;
        return 0;
    }
    __device__ int cslot(int c) {
        unsigned h = ((unsigned)c * 2654435761u) >> (32 - 5);
        for (int probe = 0; probe < SVM_T1_WC; ++probe, h = (h + 1) & (SVM_T1_WC - 1)) {
            if (ckey[h] == c) return h;
            if (ckey[h] < 0) {
                int n = d->cnn[c];
                if (n > SVM_T1_RC - 1) { fail = true; return 0; }
                ckey[h] = c;
                cflag[h] = 0;
                ccnt[h] = n;
                for (int k = 0; k < n; ++k) cring[h * SVM_T1_RC + k] = ringref(*d, c, k);
                claim(d->cown + c);
                return h;
            }
        }
        fail = true;
        return 0;
    }
    __device__ void pos(int v, double* p) {
        int h = vslot(v);
        p[0] = vp[3 * h]; p[1] = vp[3 * h + 1]; p[2] = vp[3 * h + 2];
    }
    __device__ void set_pos(int v, const double* p) {
        int h = vslot(v);
        vp[3 * h] = p[0]; vp[3 * h + 1] = p[1]; vp[3 * h + 2] = p[2];
        vflag[h] |= 1;
    }
    __device__ int vn(int v, int k) { return vvn[3 * vslot(v) + k]; }
    __device__ void set_vn(int v, int k, int x) { int h = vslot(v); vvn[3 * h + k] = x; vflag[h] |= 2; }
    __device__ int vcn(int v, int k) { return vvcn[3 * vslot(v) + k]; }
    __device__ void set_vcn(int v, int k, int x) { int h = vslot(v); vvcn[3 * h + k] = x; vflag[h] |= 2; }
    __device__ int cnn(int c) { return ccnt[cslot(c)]; }
    __device__ void set_cnn(int c, int x) {
        int h = cslot(c);
        if (x > SVM_T1_RC - 1) fail = true;
        else { ccnt[h] = x; cflag[h] = 1; }
    }
    __device__ int ring(int c, int k) { return cring[cslot(c) * SVM_T1_RC + (k & (SVM_T1_RC - 1))]; }
    __device__ void set_ring(int c, int k, int x) {
        int h = cslot(c);
        cring[h * SVM_T1_RC + (k & (SVM_T1_RC - 1))] = x;
        cflag[h] = 1;
    }
    __device__ bool failed() { return fail; }
};

// -------------------------------------------------------- T1 logic --------
// getCellVertexSetForT1 (src/model/sphericalVertexModel.cpp:611-731).
// Returns false when the four cells cannot all be determined (vertex1 and
// vertex2 share all three cells, so there is no cell l); the flip is then
// skipped, where the reference would read an uninitialised cell set.
template <class A>
__device__ bool cell_vertex_set(A& a, int vertex1, int vertex2, int* cs, int* vs) {
    cs[0] = cs[1] = cs[2] = cs[3] = -1;
    int cell1 = a.vcn(vertex1, 0), cell2 = a.vcn(vertex1, 1), cell3 = a.vcn(vertex1, 2);
    for (int ff = 0; ff < 3; ++ff) {
        int ctest = a.vcn(vertex2, ff);
        if (ctest != cell1 && ctest != cell2 && ctest != cell3) cs[3] = ctest;
    }
    int cells[3] = {cell1, cell2, cell3};
    int vlast, vcur, vnext = -1, cneigh;
    for (int t = 0; t < 3; ++t) {
        int cell = cells[t];
        cneigh = a.cnn(cell);
        vlast = a.ring(cell, cneigh - 2);
        vcur = a.ring(cell, cneigh - 1);
        for (int cn = 0; cn < cneigh; ++cn) {
            vnext = a.ring(cell, cn);
            if (vcur == vertex1) break;
            vlast = vcur;
            vcur = vnext;
        }
        if (vlast == vertex2) cs[0] = cell;
        else if (vnext == vertex2) cs[2] = cell;
        else cs[1] = cell;
    }
    if (cs[0] < 0 || cs[1] < 0 || cs[2] < 0 || cs[3] < 0) return false;
    cneigh = a.cnn(cs[1]);
    vlast = a.ring(cs[1], cneigh - 2);
    vcur = a.ring(cs[1], cneigh - 1);
    for (int cn = 0; cn < cneigh; ++cn) {
        vnext = a.ring(cs[1], cn);
        if (vcur == vertex1) break;
        vlast = vcur;
        vcur = vnext;
    }
    vs[0] = vlast;
    vs[1] = vnext;
    cneigh = a.cnn(cs[3]);
    vlast = a.ring(cs[3], cneigh - 2);
    vcur = a.ring(cs[3], cneigh - 1);
    for (int cn = 0; cn < cneigh; ++cn) {
        vnext = a.ring(cs[3], cn);
        if (vcur == vertex2) break;
        vlast = vcur;
        vcur = vnext;
    }
    vs[3] = vlast;
    vs[2] = vnext;
    return true;
}

// One T1 flip of edge (vertex1, vertex2): the body of
// testAndPerformT1TransitionsCPU (src/model/sphericalVertexModel.cpp:496-598).
// v1 is the caller's copy of vertex1's position.  Returns 0, or
// SVM_ERR_RING_OVERFLOW (nothing modified) when a cell would outgrow `cap`.
template <class A>
__device__ int t1_flip(A& a, int vertex1, int vertex2, const int* cs, const int* vs, double* v1, double R,
                       int cap, int* tmp) {
    if (a.cnn(cs[1]) + 1 > cap || a.cnn(cs[3]) + 1 > cap) return 4;
    double v2[3], vC[3];
    a.pos(vertex2, v2);
    double mid[3] = {0.5 * (v1[0] + v2[0]), 0.5 * (v1[1] + v2[1]), 0.5 * (v1[2] + v2[2])};
    double mm = 0.0; mm += mid[0] * mid[0]; mm += mid[1] * mid[1]; mm += mid[2] * mid[2];
    double im = 1.0 / __builtin_sqrt(mm);
    mid[0] = im * mid[0]; mid[1] = im * mid[1]; mid[2] = im * mid[2];
    a.pos(vs[2], vC);
    double determinant = vC[0] * (v1[1] * v2[2] - v1[2] * v2[1])
                       + vC[1] * (v1[2] * v2[0] - v1[0] * v2[2])
                       + vC[2] * (v1[0] * v2[1] - v1[1] * v2[0]);
    determinant = determinant > 0 ? 1. : -1.;
    rodrigues(v1, mid, -0.5 * determinant * SVM_PI);
    rodrigues(v2, mid, -0.5 * determinant * SVM_PI);
    double diff[3] = {0.5 * (v1[0] - v2[0]), 0.5 * (v1[1] - v2[1]), 0.5 * (v1[2] - v2[2])};
    for (int k = 0; k < 3; ++k) { v1[k] = v1[k] + diff[k]; v2[k] = v2[k] - diff[k]; }
    double n1 = 0.0; n1 += v1[0] * v1[0]; n1 += v1[1] * v1[1]; n1 += v1[2] * v1[2];
    double n2 = 0.0; n2 += v2[0] * v2[0]; n2 += v2[1] * v2[1]; n2 += v2[2] * v2[2];
    double i1 = 1.0 / __builtin_sqrt(n1), i2 = 1.0 / __builtin_sqrt(n2);
    for (int k = 0; k < 3; ++k) { v1[k] = R * (i1 * v1[k]); v2[k] = R * (i2 * v2[k]); }
    a.set_pos(vertex1, v1);
    a.set_pos(vertex2, v2);
    for (int vert = 0; vert < 3; ++vert) {
        if (a.vcn(vertex1, vert) == cs[2]) a.set_vcn(vertex1, vert, cs[3]);
        if (a.vcn(vertex2, vert) == cs[0]) a.set_vcn(vertex2, vert, cs[1]);
        if (a.vn(vs[1], vert) == vertex1) a.set_vn(vs[1], vert, vertex2);
        if (a.vn(vs[2], vert) == vertex2) a.set_vn(vs[2], vert, vertex1);
        if (a.vn(vertex1, vert) == vs[1]) a.set_vn(vertex1, vert, vs[2]);
        if (a.vn(vertex2, vert) == vs[2]) a.set_vn(vertex2, vert, vs[1]);
    }
    // cell i loses vertex2
    int cneigh = a.cnn(cs[0]);
    int cidx = 0;
    for (int cc = 0; cc < cneigh - 1; ++cc) {
        if (a.ring(cs[0], cc) == vertex2) cidx += 1;
        a.set_ring(cs[0], cc, a.ring(cs[0], cidx));
        cidx += 1;
    }
    a.set_cnn(cs[0], cneigh - 1);
    // cell j gains vertex2 between vertex1 and b
    cneigh = a.cnn(cs[1]);
    cidx = 0;
    for (int cc = 0; cc < cneigh; ++cc) {
        int e = a.ring(cs[1], cc);
        tmp[cidx++] = e;
        if (e == vertex1) tmp[cidx++] = vertex2;
    }
    for (int cc = 0; cc < cneigh + 1; ++cc) a.set_ring(cs[1], cc, tmp[cc]);
    a.set_cnn(cs[1], cneigh + 1);
    // cell k loses vertex1
    cneigh = a.cnn(cs[2]);
    cidx = 0;
    for (int cc = 0; cc < cneigh - 1; ++cc) {
        if (a.ring(cs[2], cc) == vertex1) cidx += 1;
        a.set_ring(cs[2], cc, a.ring(cs[2], cidx));
        cidx += 1;
    }
    a.set_cnn(cs[2], cneigh - 1);
    // cell l gains vertex1 between vertex2 and a
    cneigh = a.cnn(cs[3]);
    cidx = 0;
    for (int cc = 0; cc < cneigh; ++cc) {
        int e = a.ring(cs[3], cc);
        tmp[cidx++] = e;
        if (e == vertex2) tmp[cidx++] = vertex1;
    }
    for (int cc = 0; cc < cneigh + 1; ++cc) a.set_ring(cs[3], cc, tmp[cc]);
    a.set_cnn(cs[3], cneigh + 1);
    return 0;
}

struct VisitStats {
    long long flips, skipped;
};

// The sequential loop's visit of vertex u (row gu = vo + u): slots 0..2, each
// edge (u, w) with u < w shorter than thr flipped unless protected.  `push`
// receives every later vertex (local index > u) a flip disturbed; `on_flip`
// receives each flip's (vertex1, vertex2, cell set).  Returns 0 or an error.
template <class A, class Push, class OnFlip>
__device__ int visit_vertex(A& a, int gu, int u, int vo, int nv, int co, int nc, double R, double thr, int cap,
                            int* tmp, VisitStats& st, Push push, OnFlip on_flip) {
    double v1[3];
    a.pos(gu, v1);
    for (int vv = 0; vv < 3; ++vv) {
        int vertex2 = a.vn(gu, vv);
        if (gu < vertex2) {
            double q[3];
            a.pos(vertex2, q);
            double arcLength = geodesic(v1, q, pnorm(v1), pnorm(q), R);
            if (arcLength < thr) {
                int cs[4], vs[4] = {0, 0, 0, 0};
                if (!cell_vertex_set(a, gu, vertex2, cs, vs)) {
                    ++st.skipped;
                    continue;
                }
                bool inside = true;
                for (int k = 0; k < 4; ++k) {
                    inside &= cs[k] >= co && cs[k] < co + nc;
                    inside &= vs[k] >= vo && vs[k] < vo + nv;
                }
                if (!inside) return 5;
                if (a.cnn(cs[0]) == 3 || a.cnn(cs[2]) == 3) continue;
                int r = t1_flip(a, gu, vertex2, cs, vs, v1, R, cap, tmp);
                if (r) return r;
                ++st.flips;
                on_flip(gu, vertex2, cs);
                int dirty[6] = {u, vertex2 - vo, vs[0] - vo, vs[1] - vo, vs[2] - vo, vs[3] - vo};
                for (int k = 0; k < 6; ++k)
                    if (dirty[k] > u && dirty[k] < nv) push(dirty[k]);
            }
        }
        if (a.failed()) return 0;
    }
    return 0;
}

// Rebuild vln for vertex x from the current rings; returns false if a ring
// neighbour is not in x's vertex-neighbour list.
__device__ inline bool build_vln(const Dev& d, int x) {
    int4 cl = d.vcn[x];
    int4 nb = d.vn[x];
    int cells[3] = {cl.x, cl.y, cl.z};
    int out = 0;
    bool ok = true;
    for (int sl = 0; sl < 3; ++sl) {
        int c = cells[sl];
        int n = d.cnn[c];
        int k = -1;
        for (int j = 0; j < n; ++j)
            if (ringref(d, c, j) == x) { k = j; break; }
        if (k < 0) { ok = false; continue; }
        int prev = ringref(d, c, (k + n - 1) % n);
        int next = ringref(d, c, (k + 1) % n);
        int ls = prev == nb.x ? 0 : (prev == nb.y ? 1 : (prev == nb.z ? 2 : -1));
        int ns = next == nb.x ? 0 : (next == nb.y ? 1 : (next == nb.z ? 2 : -1));
        if (ls < 0 || ns < 0) { ok = false; continue; }
        out |= (ls << (4 * sl)) | (ns << (4 * sl + 2));
    }
    d.vln[x] = out;
    return ok;
}

__global__ __launch_bounds__(SVM_TILE) void k_build_vln(Dev d, int only) {
    const int tile = blockIdx.x;
    const int s = d.vtile_sys[tile];
    SysParams* S = d.sys + s;
    if (only >= 0 && s != only) return;
    const int v = tile * SVM_TILE + threadIdx.x;
    if (v - S->vo >= S->nv) return;
    if (!build_vln(d, v)) atomicMax(&S->err, 5);
}

// ------------------------------------------------------------- detect -----
// Short-edge test of testAndPerformT1TransitionsCPU
// (src/model/sphericalVertexModel.cpp:469-481): vertex v is a candidate when
// it has an edge (v, w), v < w, shorter than the threshold.  Candidates go to
// the bit mask (for the sequential path) and to the compact list (for the dry
// runs); the active updater's director step is fused here.
__global__ __launch_bounds__(SVM_TILE) void k_t1_detect(Dev d, int pcur, int nin, int nout, int dir_update,
                                                        int skip_frozen) {
    const int tile = blockIdx.x;
    const int s = d.vtile_sys[tile];
    SysParams* S = d.sys + s;
    if (skip_frozen && S->f_frozen) return;
    const int v = tile * SVM_TILE + threadIdx.x;
    const int local = v - S->vo;
    const double* X = d.px[pcur];
    const double* Y = d.py[pcur];
    const double* Z = d.pz[pcur];
    bool flag = false;
    if (local < S->nv) {
        double p[3] = {X[v], Y[v], Z[v]};
        double n1 = pnorm(p);
        int4 nb = d.vn[v];
        int w3[3] = {nb.x, nb.y, nb.z};
        for (int k = 0; k < 3; ++k) {
            int w = w3[k];
            if (v < w) {
                double q[3] = {X[w], Y[w], Z[w]};
                double arc = geodesic(p, q, n1, pnorm(q), S->R);
                if (arc < S->thr) flag = true;
            }
        }
        if (dir_update) {
            double nI[3] = {d.nx[nin][v], d.ny[nin][v], d.nz[nin][v]};
            double nO[3];
            director_update(S, local, p, nI, nO);
            d.nx[nout][v] = nO[0];
            d.ny[nout][v] = nO[1];
            d.nz[nout][v] = nO[2];
        }
    }
    unsigned long long m = __ballot(flag);
    const int lane = threadIdx.x & 63;
    int base = 0;
    if (lane == 0) {
        d.t1mask[v >> 6] = m;
        if (m) {
            atomicOr(&S->t1_flag, 1);
            int cnt = __popcll(m);
            base = atomicAdd(d.t1_count, cnt);
            if (base + cnt > SVM_T1_MAXCL) atomicOr(&S->t1_overlap, 1);
        }
    }
    base = __shfl(base, 0, 64);
    if (flag) {
        int pos = base + __popcll(m & ((1ull << lane) - 1ull));
        if (pos < SVM_T1_MAXCL) d.t1_list[pos] = v;
    }
}

// ------------------------------------------------------------ dry runs ----
// Apply the logged results of every dry run of one sphere (whole block),
// then rebuild vln of the vertices of modified cells and, for the active
// updater, the directors of moved vertices.
__device__ inline void commit_sphere(const Dev& d, SysParams* S, int pcur, int nin, int nout, int dir_mode) {
    const int t = threadIdx.x, nt = blockDim.x;
    for (int idx = S->t1_head; idx >= 0; idx = d.t1_log[idx].next) {
        const T1Log& L = d.t1_log[idx];
        for (int i = t; i < L.nvert; i += nt) {
            int v = L.vid[i];
            if (L.vflag[i] & 1) {
                d.px[pcur][v] = L.vpos[i][0];
                d.py[pcur][v] = L.vpos[i][1];
                d.pz[pcur][v] = L.vpos[i][2];
            }
            if (L.vflag[i] & 2) {
                d.vn[v] = make_int4(L.vvn[i][0], L.vvn[i][1], L.vvn[i][2], d.vn[v].w);
                d.vcn[v] = make_int4(L.vvcn[i][0], L.vvcn[i][1], L.vvcn[i][2], d.vcn[v].w);
            }
        }
        for (int i = t; i < L.ncell; i += nt) {
            int c = L.cid[i];
            d.cnn[c] = L.ccnt[i];
            for (int k = 0; k < L.ccnt[i]; ++k) ringref(d, c, k) = L.cring[i][k];
        }
    }
    if (t == 0) {
        long long flips = 0, skipped = 0;
        for (int idx = S->t1_head; idx >= 0; idx = d.t1_log[idx].next) {
            flips += d.t1_log[idx].flips;
            skipped += d.t1_log[idx].skipped;
        }
        S->n_t1 += flips;
        S->n_t1_skipped += skipped;
    }
    __threadfence_block();
    __syncthreads();
    for (int idx = S->t1_head; idx >= 0; idx = d.t1_log[idx].next) {
        const T1Log& L = d.t1_log[idx];
        for (int i = t; i < L.ncell; i += nt) {
            int c = L.cid[i];
            int n = d.cnn[c];
            for (int k = 0; k < n; ++k)
                if (!build_vln(d, ringref(d, c, k))) atomicMax(&S->err, 5);
        }
        if (dir_mode) {
            for (int i = t; i < L.nvert; i += nt) {
                if (!(L.vflag[i] & 1)) continue;
                int x = L.vid[i];
                double p[3] = {d.px[pcur][x], d.py[pcur][x], d.pz[pcur][x]};
                double nI[3] = {d.nx[nin][x], d.ny[nin][x], d.nz[nin][x]}, nO[3];
                director_update(S, x - S->vo, p, nI, nO);
                d.nx[nout][x] = nO[0]; d.ny[nout][x] = nO[1]; d.nz[nout][x] = nO[2];
            }
        }
    }
    __syncthreads();
}

__global__ __launch_bounds__(64 * SVM_T1_WAVES) void k_t1_cluster(Dev d, int pcur) {
    __shared__ int s_vkey[SVM_T1_WAVES][SVM_T1_WV], s_vflag[SVM_T1_WAVES][SVM_T1_WV];
    __shared__ double s_vp[SVM_T1_WAVES][SVM_T1_WV * 3];
    __shared__ int s_vvn[SVM_T1_WAVES][SVM_T1_WV * 3], s_vvcn[SVM_T1_WAVES][SVM_T1_WV * 3];
    __shared__ int s_ckey[SVM_T1_WAVES][SVM_T1_WC], s_cflag[SVM_T1_WAVES][SVM_T1_WC], s_ccnt[SVM_T1_WAVES][SVM_T1_WC];
    __shared__ int s_cring[SVM_T1_WAVES][SVM_T1_WC * SVM_T1_RC];
    __shared__ int s_heap[SVM_T1_WAVES][64];
    __shared__ int s_tmp[SVM_T1_WAVES][SVM_T1_RC + 8];
    const int w = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int nwaves = gridDim.x * SVM_T1_WAVES;
    const int total = min(*d.t1_count, SVM_T1_MAXCL);
    const unsigned long long gen = (unsigned long long)(*d.t1_gen) << 32;
    for (int idx = blockIdx.x * SVM_T1_WAVES + w; idx < total; idx += nwaves) {
        const int row = d.t1_list[idx];
        const int s = d.vtile_sys[row / SVM_TILE];
        SysParams* S = d.sys + s;
        // reset the private cache
        s_vkey[w][lane] = -1;
        if (lane < SVM_T1_WC) s_ckey[w][lane] = -1;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        if (lane == 0 && !__hip_atomic_load(&S->t1_overlap, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            CacheAcc a{&d, pcur, gen | (unsigned long long)(idx + 1), &S->t1_overlap, false,
                       s_vkey[w], s_vflag[w], s_vp[w], s_vvn[w], s_vvcn[w],
                       s_ckey[w], s_cflag[w], s_ccnt[w], s_cring[w]};
            const int vo = S->vo, nv = S->nv;
            Heap heap{s_heap[w], 0};
            bool heap_full = false;
            VisitStats st{0, 0};
            int err = 0, last = -1;
            int u = row - vo;
            while (true) {
                last = u;
                err = visit_vertex(a, vo + u, u, vo, nv, S->co, S->nc, S->R, S->thr, d.cap, s_tmp[w], st,
                                   [&](int x) {
                                       if (heap.n < 64) heap.push(x);
                                       else heap_full = true;
                                   },
                                   [&](int, int, const int*) {});
                if (err || a.failed() || heap_full) break;
                int nxt = -1;
                while (heap.n > 0) {
                    int x = heap.pop();
                    if (x > last) { nxt = x; break; }
                }
                if (nxt < 0) break;
                u = nxt;
            }
            if (err == 4) atomicMax(&S->err, 4);
            if (err || a.failed() || heap_full) {
                atomicOr(&S->t1_overlap, 1);
            } else if (st.flips || st.skipped) {
                // log the modified items
                T1Log& L = d.t1_log[idx];
                int nvl = 0, ncl = 0;
                bool too_many = false;
                for (int h = 0; h < SVM_T1_WV; ++h) {
                    if (s_vkey[w][h] < 0 || !s_vflag[w][h]) continue;
                    if (nvl == SVM_T1_LV) { too_many = true; break; }
                    L.vid[nvl] = s_vkey[w][h];
                    L.vflag[nvl] = s_vflag[w][h];
                    for (int k = 0; k < 3; ++k) {
                        L.vpos[nvl][k] = s_vp[w][3 * h + k];
                        L.vvn[nvl][k] = s_vvn[w][3 * h + k];
                        L.vvcn[nvl][k] = s_vvcn[w][3 * h + k];
                    }
                    ++nvl;
                }
                for (int h = 0; h < SVM_T1_WC && !too_many; ++h) {
                    if (s_ckey[w][h] < 0 || !s_cflag[w][h]) continue;
                    if (ncl == SVM_T1_LC) { too_many = true; break; }
                    L.cid[ncl] = s_ckey[w][h];
                    L.ccnt[ncl] = s_ccnt[w][h];
                    for (int k = 0; k < s_ccnt[w][h]; ++k) L.cring[ncl][k] = s_cring[w][h * SVM_T1_RC + k];
                    ++ncl;
                }
                if (too_many) {
                    atomicOr(&S->t1_overlap, 1);
                } else {
                    L.nvert = nvl;
                    L.ncell = ncl;
                    L.flips = st.flips;
                    L.skipped = st.skipped;
                    L.next = atomicExch(&S->t1_head, idx);
                }
            }
        }
    }
}

// ------------------------------------------------------ sequential pass ----
// Closes the T1 step of each sphere (block per sphere): commits the dry-run
// logs of spheres whose candidates did not overlap; otherwise applies the
// flips in the order of the reference's sequential loop, over the candidates
// of the bit mask plus every later vertex a flip disturbed (a min-heap).  If
// the heap fills up, every remaining vertex is visited in order (still
// exact).
__global__ __launch_bounds__(SVM_TILE) void k_t1_resolve(Dev d, int pcur, int nin, int nout, int dir_mode,
                                                         int advance_step, int fire_mode) {
    __shared__ int heapbuf[SVM_HEAP_CAP];
    __shared__ int moved[SVM_LIST_CAP];
    __shared__ int touched[SVM_LIST_CAP];
    __shared__ int scan[SVM_TILE];
    __shared__ int tmp[72];
    __shared__ int sh_run, sh_commit, sh_nmoved, sh_ntouched;
    const int s = blockIdx.x;
    SysParams* S = d.sys + s;
    const int t = threadIdx.x;
    if (t == 0) {
        bool frozen = fire_mode && S->f_frozen;
        int any = S->t1_flag;
        sh_run = !frozen && any && S->t1_overlap;
        sh_commit = !frozen && any && !S->t1_overlap && S->t1_head >= 0;
        if (!frozen) {
            S->t1_flag = 0;
            if (fire_mode) S->mix_pending = 0;
        }
        sh_nmoved = 0;
        sh_ntouched = 0;
        if (s == 0) {
            // the dry-run list and generation of this step are used up
            *d.t1_count = 0;
            *d.t1_gen += 1;
        }
    }
    __syncthreads();
    if (!sh_run) {
        if (sh_commit) commit_sphere(d, S, pcur, nin, nout, dir_mode);
        if (t == 0) {
            S->t1_overlap = 0;
            S->t1_head = -1;
            if (advance_step) S->step += 1;
        }
        return;
    }
    const int vo = S->vo, nv = S->nv;
    const int nw = (nv + 63) >> 6;
    const int wbase = vo >> 6;
    int ncand = 0;
    for (int base = 0; base < nw; base += SVM_TILE) {
        unsigned long long w = (base + t < nw) ? d.t1mask[wbase + base + t] : 0ull;
        int cnt = __popcll(w);
        scan[t] = cnt;
        __syncthreads();
        for (int off = 1; off < SVM_TILE; off <<= 1) {
            int add = t >= off ? scan[t - off] : 0;
            __syncthreads();
            scan[t] += add;
            __syncthreads();
        }
        int pos = ncand + scan[t] - cnt;
        int total = scan[SVM_TILE - 1];
        while (w) {
            int b = __ffsll((long long)w) - 1;
            d.cand[vo + pos++] = (base + t) * 64 + b;
            w &= w - 1;
        }
        __syncthreads();
        ncand += total;
    }
    __threadfence_block();
    __syncthreads();
    if (t == 0) {
        S->n_t1_sequential += 1;
        GlobalAcc a{&d, pcur};
        Heap heap{heapbuf, 0};
        int ci = 0, last = -1, nm = 0, nt = 0, err = 0;
        bool full_scan = false;
        VisitStats st{0, 0};
        auto push = [&](int x) {
            if (full_scan) return;
            if (heap.n < SVM_HEAP_CAP) heap.push(x);
            else full_scan = true;
        };
        auto on_flip = [&](int v1, int v2, const int* cs) {
            if (nm + 2 <= SVM_LIST_CAP) { moved[nm++] = v1; moved[nm++] = v2; }
            else nm = SVM_LIST_CAP + 1;
            for (int k = 0; k < 4; ++k) {
                if (nt < SVM_LIST_CAP) touched[nt++] = cs[k];
                else nt = SVM_LIST_CAP + 1;
            }
        };
        while (true) {
            int u;
            if (!full_scan) {
                int cu = ci < ncand ? d.cand[vo + ci] : 0x7fffffff;
                int hu = heap.n > 0 ? heap.a[0] : 0x7fffffff;
                if (cu == 0x7fffffff && hu == 0x7fffffff) break;
                if (hu < cu) { u = hu; heap.pop(); }
                else { u = cu; ++ci; }
                if (u <= last) continue;
            } else {
                u = last + 1;
                if (u >= nv) break;
            }
            last = u;
            err = visit_vertex(a, vo + u, u, vo, nv, S->co, S->nc, S->R, S->thr, d.cap, tmp, st, push, on_flip);
            if (err) break;
        }
        S->n_t1 += st.flips;
        S->n_t1_skipped += st.skipped;
        sh_nmoved = nm;
        sh_ntouched = nt;
        if (err) atomicMax(&S->err, err);
    }
    __threadfence_block();
    __syncthreads();
    const int ntouch = sh_ntouched;
    if (ntouch > SVM_LIST_CAP) {
        for (int x = vo + t; x < vo + nv; x += SVM_TILE)
            if (!build_vln(d, x)) atomicMax(&S->err, 5);
    } else {
        for (int i = t; i < ntouch; i += SVM_TILE) {
            int c = touched[i];
            int n = d.cnn[c];
            for (int k = 0; k < n; ++k)
                if (!build_vln(d, ringref(d, c, k))) atomicMax(&S->err, 5);
        }
    }
    if (dir_mode) {
        const int nmv = sh_nmoved;
        const double* X = d.px[pcur];
        const double* Y = d.py[pcur];
        const double* Z = d.pz[pcur];
        const int hi = nmv > SVM_LIST_CAP ? nv : nmv;
        for (int i = t; i < hi; i += SVM_TILE) {
            int x = nmv > SVM_LIST_CAP ? vo + i : moved[i];
            double p[3] = {X[x], Y[x], Z[x]};
            double nI[3] = {d.nx[nin][x], d.ny[nin][x], d.nz[nin][x]}, nO[3];
            director_update(S, x - vo, p, nI, nO);
            d.nx[nout][x] = nO[0]; d.ny[nout][x] = nO[1]; d.nz[nout][x] = nO[2];
        }
    }
    __syncthreads();
    if (t == 0) {
        S->t1_overlap = 0;
        S->t1_head = -1;
        if (advance_step) S->step += 1;
    }
}

}  // namespace svm
