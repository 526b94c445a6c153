// svm_voronoi.h -- the spherical Delaunay triangulation of cell points on the
// GPU, and the vertex-model topology built from it.
//
// Reference behaviour: convexHullCGALInterface::sphericalConvexHullForVertexModel
// (src/model/convexHullCGAL.cpp:82-186) takes the convex hull of the cell
// points (for points on a sphere, the Delaunay triangulation); every hull
// facet becomes a Voronoi vertex at R * normalize(a + b + c); cells list their
// vertices clockwise seen from outside.  sphericalVoronoi::convexHull
// (src/model/sphericalVoronoi.cpp:8-12) is the same triangulation, redone at
// every move.  The neighbour search follows the role of hyperrectangularCellList
// (src/utilities/hyperrectangularCellList.cpp:14-51, 165-342): points binned in
// a uniform cube grid, candidates read from the bins around a point.
//
// Method (own design):
//  * cell list: bin = floor((x - lo) / h) per axis; counting sort; each bin's
//    points sorted by index, so the list does not depend on atomic order.
//  * star of point i, one wavefront per point: the candidates within chord
//    distance r; the nearest is a Delaunay neighbour; then gift wrapping
//    around i: the next neighbour after j is the candidate k with no other
//    candidate beyond the plane (i, j, k).  It is found by a floating-point
//    angle key about the axis (i, j) and confirmed with the exact orient3d
//    predicate (svm_exact.h) against every candidate; when the confirmation
//    fails (or meets an exact zero), by a tournament across the lanes whose
//    comparison is the exact predicate.
//    Every facet is then checked against every candidate (exactly), and its
//    circumcircle must lie well inside the search ball, so no point outside
//    the candidates can lie in it; otherwise the point is redone with a
//    larger r.  Four co-circular points (an exact zero) are reported as
//    degenerate input.
//  * triangles numbered as the host canonical form does: CCW seen from
//    outside, smallest cell first, sorted lexicographically; rings walk the
//    triangles around a cell starting from its smallest triangle.
#pragma once
#include "svm_exact.h"
#include "svm_device.h"
#include "svm_geom.h"

namespace svm {

#define SVM_DL_MAXC 512        // candidates one star can hold (8 per lane)
#define SVM_DL_WAVES 4         // stars per block

enum {
    DL_OK = 0,
    DL_RETRY = 1,        // circumcircle not inside the search ball: redo with another radius
    DL_CROWDED = 2,      // more than SVM_DL_MAXC candidates at every radius that could prove the star
    DL_OPEN = 3,         // the star did not close within the ring capacity
    DL_INVALID = 4,      // a facet failed the exact check
    DL_DEGENERATE = 5,   // four co-circular points
    DL_PENDING = 6,      // not computed yet
};

struct DlGrid {
    double lo, h;
    int G;
};

// ------------------------------------------------------------- cell list --
__device__ inline int dl_axis(double x, const DlGrid& g) {
    int k = (int)__builtin_floor((x - g.lo) / g.h);
    return k < 0 ? 0 : (k >= g.G ? g.G - 1 : k);
}

__global__ __launch_bounds__(256) void k_dl_bin(const double* P, int n, DlGrid g, int* bin_of, int* count) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int b = (dl_axis(P[3 * (size_t)i], g) * g.G + dl_axis(P[3 * (size_t)i + 1], g)) * g.G +
            dl_axis(P[3 * (size_t)i + 2], g);
    bin_of[i] = b;
    atomicAdd(count + b, 1);
}

__global__ __launch_bounds__(256) void k_dl_fill(int n, const int* bin_of, const int* start, int* fill, int* list) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int b = bin_of[i];
    list[start[b] + atomicAdd(fill + b, 1)] = i;
}

// each bin's points in index order
__global__ __launch_bounds__(256) void k_dl_sort_bins(int nbins, const int* start, int* list) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbins) return;
    int lo = start[b], hi = start[b + 1];
    for (int i = lo + 1; i < hi; ++i) {
        int x = list[i], j = i - 1;
        while (j >= lo && list[j] > x) { list[j + 1] = list[j]; --j; }
        list[j + 1] = x;
    }
}

// exclusive scan in blocks of 1024: out[i] = sum of in[<i] within the block,
// block totals to sums[block]
__global__ __launch_bounds__(1024) void k_scan_block(const int* in, int n, int* out, int* sums) {
    __shared__ int s[1024];
    int i = blockIdx.x * 1024 + threadIdx.x;
    int x = i < n ? in[i] : 0;
    s[threadIdx.x] = x;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int add = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += add;
        __syncthreads();
    }
    if (i < n) out[i] = s[threadIdx.x] - x;
    if (threadIdx.x == 1023) sums[blockIdx.x] = s[1023];
}

__global__ __launch_bounds__(1024) void k_scan_add(int* out, int n, const int* offs) {
    int i = blockIdx.x * 1024 + threadIdx.x;
    if (i < n) out[i] += offs[blockIdx.x];
}

// ------------------------------------------------------------------ stars --
__device__ inline void dl_load(const double* P, int i, double* p) {
    p[0] = P[3 * (size_t)i];
    p[1] = P[3 * (size_t)i + 1];
    p[2] = P[3 * (size_t)i + 2];
}

// Does candidate x beat candidate k as the wrap point after edge (pi, pj)?
// Exactly: x lies beyond the plane (pi, pj, pk).  A tie (co-planar) is
// broken by index: if both lose to a third candidate the tie does not matter,
// and if they win, their facet's circle holds four points, which the exact
// check of every facet against every candidate reports as degenerate input.
__device__ inline bool dl_beats(const double* pi, const double* pj, const double* pk, int k, const double* px, int x) {
    int o = orient3d_sign(pi, pj, pk, px);
    if (o == 0) return x < k;
    return o > 0;
}

// Star of every point i with status[i] == DL_PENDING or DL_RETRY.  Writes the
// CCW ring of Delaunay neighbours (nbr[i*cap ..], deg[i]) and the status.
//
// Latency is what bounds a star (a handful of FP64 predicates between
// dependent memory round trips), so the round trips are kept few: the bins
// read are those that meet the box [p - r, p + r], their entries are dealt
// evenly over the lanes (a wave scan of the bin sizes, compaction by ballot,
// so the candidates come in bin order), each lane keeps its first candidate
// in registers for the tournament and the checks, and lane k keeps ring
// point k.  At least 4 waves per SIMD (128 VGPRs): measured faster than the
// compiler's own choice of 132 registers and 3 waves.
__global__ __launch_bounds__(64 * SVM_DL_WAVES, 4) void k_dl_star(const double* P, int n, DlGrid g, const int* start,
                                                               const int* list, const double* rad, double R,
                                                               int cap, int* nbr, int* deg, int* status) {
    __shared__ int s_idx[SVM_DL_WAVES][SVM_DL_MAXC];
    __shared__ int s_ring[SVM_DL_WAVES][64];
    __shared__ int s_bstart[SVM_DL_WAVES][64];   // first entry of each bin of a chunk
    __shared__ int s_bpre[SVM_DL_WAVES][64];     // entries of the chunk's earlier bins
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = blockIdx.x * SVM_DL_WAVES + w;
    if (i >= n) return;
    const int st0 = status[i];
    if (st0 != DL_PENDING && st0 != DL_RETRY) return;
    const double r = rad[i];   // this star's search radius
    double pi[3];
    dl_load(P, i, pi);
    // ---- gather the candidates within chord distance r.  Binning is monotone
    // per axis and so is rounding, so every point within r of p lies in a bin
    // between those of p - r and p + r on each axis; the 1e-9 pad covers the
    // rounding of the distance test itself.
    const double rb = r * (1.0 + 1e-9);
    const int x0 = dl_axis(pi[0] - rb, g), y0 = dl_axis(pi[1] - rb, g), z0 = dl_axis(pi[2] - rb, g);
    const int nx = dl_axis(pi[0] + rb, g) - x0 + 1, ny = dl_axis(pi[1] + rb, g) - y0 + 1,
              nz = dl_axis(pi[2] + rb, g) - z0 + 1;
    const int nbin = nx * ny * nz;
    const double r2 = r * r;
    int* bs = s_bstart[w];
    int* bp = s_bpre[w];
    int nc = 0;   // wave-uniform: candidates found (also those beyond the capacity)
    for (int c0 = 0; c0 < nbin; c0 += 64) {
        const int b = c0 + lane;
        int s0 = 0, cnt = 0;
        if (b < nbin) {
            const int x = x0 + b % nx, y = y0 + (b / nx) % ny, z = z0 + b / (nx * ny);
            const int bin = (x * g.G + y) * g.G + z;
            s0 = start[bin];
            cnt = start[bin + 1] - s0;
        }
        int inc = cnt;   // inclusive scan of the bin sizes over the lanes
        for (int m = 1; m < 64; m <<= 1) {
            const int o = __shfl_up(inc, m, 64);
            if (lane >= m) inc += o;
        }
        const int total = __shfl(inc, 63, 64);
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        bs[lane] = s0;
        bp[lane] = inc - cnt;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        for (int t0 = 0; t0 < total; t0 += 64) {
            const int t = t0 + lane;
            bool hit = false;
            int q = -1;
            if (t < total) {
                // the last bin whose prefix is <= t holds entry t (it has entries:
                // the next bin's prefix exceeds t)
                int lo = 0, hi = 63;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (bp[mid] <= t) lo = mid;
                    else hi = mid - 1;
                }
                q = list[bs[lo] + (t - bp[lo])];
                if (q != i) {
                    const double dx = P[3 * (size_t)q] - pi[0], dy = P[3 * (size_t)q + 1] - pi[1],
                                 dz = P[3 * (size_t)q + 2] - pi[2];
                    hit = dx * dx + dy * dy + dz * dz < r2;
                }
            }
            const unsigned long long hits = __ballot(hit);
            if (hit) {
                const int k = nc + __popcll(hits & ((1ull << lane) - 1ull));
                if (k < SVM_DL_MAXC) s_idx[w][k] = q;
            }
            nc += __popcll(hits);
        }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (nc > SVM_DL_MAXC || nc < 3) {
        if (lane == 0) status[i] = nc < 3 ? (r < 2.0 * R ? DL_RETRY : DL_INVALID) : DL_CROWDED;
        return;
    }
    const int* cand = s_idx[w];
    // candidate c's point: lane c holds it for c < 64, later ones are loaded
    double pm[3] = {0.0, 0.0, 0.0};
    if (lane < nc) dl_load(P, cand[lane], pm);
    auto cand_point = [&](int c, double* pq) {
        if (c < 64) {
            pq[0] = pm[0]; pq[1] = pm[1]; pq[2] = pm[2];
        } else {
            dl_load(P, cand[c], pq);
        }
    };
    // ---- nearest candidate (squared distance, then index)
    double bd = 1e300;
    int bj = 0x7fffffff;
    for (int c = lane; c < nc; c += 64) {
        const int q = cand[c];
        double pq[3];
        cand_point(c, pq);
        const double dx = pq[0] - pi[0], dy = pq[1] - pi[1], dz = pq[2] - pi[2];
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 < bd || (d2 == bd && q < bj)) { bd = d2; bj = q; }
    }
    for (int m = 32; m >= 1; m >>= 1) {
        double od = __shfl_xor(bd, m, 64);
        int oj = __shfl_xor(bj, m, 64);
        if (od < bd || (od == bd && oj < bj)) { bd = od; bj = oj; }
    }
    const int j0 = bj;
    // ---- gift wrapping around i
    bool degenerate = false;
    int status_out = DL_OK;
    int cur = j0, d = 0;
    int* ring = s_ring[w];   // cap <= 64
    double pc[3];
    dl_load(P, cur, pc);
    double rp[3] = {0.0, 0.0, 0.0};   // lane k: ring point k
    unsigned long long confirmed = 0;   // facets whose candidate checks were done during the wrap
    while (true) {
        if (d >= cap) { status_out = DL_OPEN; break; }
        if (lane == 0) ring[d] = cur;
        if (lane == d) { rp[0] = pc[0]; rp[1] = pc[1]; rp[2] = pc[2]; }
        ++d;
        // Fast path: the next neighbour is the candidate of largest angle
        // about the axis u = pc - pi (angle measured from e1, the direction to
        // the sphere's centre normal to u, towards e2 = u x e1), found with a
        // floating-point key and then confirmed exactly: no other candidate
        // on or beyond the plane (pi, pc, pk).  A confirmed winner beats every
        // candidate, so the exact tournament below would return it too; and
        // the confirmation is that facet's candidate check.
        int bk = -1;
        double pk[3] = {0.0, 0.0, 0.0};
        {
            const double ux = pc[0] - pi[0], uy = pc[1] - pi[1], uz = pc[2] - pi[2];
            const double mx = -(pi[0] + pc[0]), my = -(pi[1] + pc[1]), mz = -(pi[2] + pc[2]);
            const double t = (mx * ux + my * uy + mz * uz) / (ux * ux + uy * uy + uz * uz);
            const double e1x = mx - t * ux, e1y = my - t * uy, e1z = mz - t * uz;
            const double e2x = uy * e1z - uz * e1y, e2y = uz * e1x - ux * e1z, e2z = ux * e1y - uy * e1x;
            double key = -1e300;
            int kq = 0x7fffffff;
            for (int c = lane; c < nc; c += 64) {
                const int q = cand[c];
                if (q == cur) continue;
                double pq[3];
                cand_point(c, pq);
                const double dx = pq[0] - pi[0], dy = pq[1] - pi[1], dz = pq[2] - pi[2];
                const double a = dx * e1x + dy * e1y + dz * e1z, b = dx * e2x + dy * e2y + dz * e2z;
                const double sab = __builtin_fabs(a) + __builtin_fabs(b);
                if (!(sab > 0.0)) continue;   // no direction (or not finite): left to the exact path
                const double ps = b / sab;    // pseudo-angle, increasing with atan2(b, a)
                const double kx = a >= 0.0 ? ps : (b >= 0.0 ? 2.0 - ps : -2.0 - ps);
                if (kx > key || (kx == key && q < kq)) {
                    key = kx;
                    kq = q;
                    pk[0] = pq[0]; pk[1] = pq[1]; pk[2] = pq[2];
                }
            }
            for (int m = 1; m < 64; m <<= 1) {
                const double ok = __shfl_xor(key, m, 64);
                const int oq = __shfl_xor(kq, m, 64);
                const double o0 = __shfl_xor(pk[0], m, 64), o1 = __shfl_xor(pk[1], m, 64),
                             o2 = __shfl_xor(pk[2], m, 64);
                if (ok > key || (ok == key && oq < kq)) {
                    key = ok;
                    kq = oq;
                    pk[0] = o0; pk[1] = o1; pk[2] = o2;
                }
            }
            if (kq != 0x7fffffff) {
                bool fail = false;
                for (int c = lane; c < nc; c += 64) {
                    const int q = cand[c];
                    if (q == cur || q == kq) continue;
                    double pq[3];
                    cand_point(c, pq);
                    if (orient3d_sign(pi, pc, pk, pq) >= 0) fail = true;
                }
                if (__ballot(fail) == 0ull) {
                    bk = kq;
                    confirmed |= 1ull << (d - 1);
                }
            }
        }
        if (bk >= 0) {
            if (bk == j0) {
                if (d < 3) status_out = DL_INVALID;
                break;
            }
            cur = bk;
            pc[0] = pk[0]; pc[1] = pk[1]; pc[2] = pk[2];
            continue;
        }
        // exact tournament: each lane its candidates, then across lanes
        pk[0] = 0.0; pk[1] = 0.0; pk[2] = 0.0;
        for (int c = lane; c < nc; c += 64) {
            int q = cand[c];
            if (q == cur) continue;
            double pq[3];
            cand_point(c, pq);
            if (bk < 0 || dl_beats(pi, pc, pk, bk, pq, q)) {
                bk = q;
                pk[0] = pq[0]; pk[1] = pq[1]; pk[2] = pq[2];
            }
        }
        for (int m = 1; m < 64; m <<= 1) {
            int ok = __shfl_xor(bk, m, 64);
            double o0 = __shfl_xor(pk[0], m, 64), o1 = __shfl_xor(pk[1], m, 64), o2 = __shfl_xor(pk[2], m, 64);
            if (ok < 0) continue;
            double po[3] = {o0, o1, o2};
            bool take;
            if (bk < 0) take = true;
            else if (ok == bk) take = false;
            else {
                // both lanes of the pair decide the same way: compare with the smaller index first
                if (bk < ok) take = dl_beats(pi, pc, pk, bk, po, ok);
                else take = !dl_beats(pi, pc, po, ok, pk, bk);
            }
            if (take) { bk = ok; pk[0] = o0; pk[1] = o1; pk[2] = o2; }
        }
        if (bk < 0) { status_out = DL_INVALID; break; }
        if (bk == j0) {
            if (d < 3) status_out = DL_INVALID;
            break;
        }
        cur = bk;
        pc[0] = pk[0]; pc[1] = pk[1]; pc[2] = pk[2];
    }
    // ---- verification of every facet (i, ring[s], ring[s+1])
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    bool bad = false, far = false;
    if (status_out == DL_OK) {
        // lane s: facet s's orientation and circumcircle size (d >= 3 here)
        {
            const int s1 = (lane + 1) % d;
            double pb[3];
            for (int k = 0; k < 3; ++k) pb[k] = __shfl(rp[k], s1, 64);
            if (lane < d) {
                if (orient_origin_sign(pi, rp, pb) <= 0) bad = true;
                // circumradius of (pi, pa, pb); the cap beyond the facet lies within 2.25 rho of pi
                double u[3] = {rp[0] - pi[0], rp[1] - pi[1], rp[2] - pi[2]};
                double v[3] = {pb[0] - pi[0], pb[1] - pi[1], pb[2] - pi[2]};
                double wv[3] = {pb[0] - rp[0], pb[1] - rp[1], pb[2] - rp[2]};
                double cx = u[1] * v[2] - u[2] * v[1], cy = u[2] * v[0] - u[0] * v[2], cz = u[0] * v[1] - u[1] * v[0];
                double area2 = svm::sqrt_rn(cx * cx + cy * cy + cz * cz);
                double rho = svm::sqrt_rn(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) *
                             svm::sqrt_rn(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) *
                             svm::sqrt_rn(wv[0] * wv[0] + wv[1] * wv[1] + wv[2] * wv[2]) / (2.0 * area2);
                if (!(2.25 * rho * (1.0 + 1e-6) < r) && r < 2.0 * R) far = true;
            }
        }
        // every candidate against the facets the wrap did not confirm
        for (int s = 0; s < d; ++s) {
            if ((confirmed >> s) & 1ull) continue;
            const int s1 = (s + 1) % d;
            int a = ring[s], b = ring[s1];
            double pa[3], pb[3];
            for (int k = 0; k < 3; ++k) {
                pa[k] = __shfl(rp[k], s, 64);
                pb[k] = __shfl(rp[k], s1, 64);
            }
            for (int c = lane; c < nc; c += 64) {
                int q = cand[c];
                if (q == a || q == b) continue;
                double pq[3];
                cand_point(c, pq);
                int o = orient3d_sign(pi, pa, pb, pq);
                if (o > 0) bad = true;
                if (o == 0) degenerate = true;
            }
        }
    }
    unsigned long long any_bad = __ballot(bad), any_far = __ballot(far), any_deg = __ballot(degenerate);
    if (lane == 0) {
        // A star is proven only when every facet passes the exact checks and
        // its circumcircle lies inside the search ball.  Any failure while
        // the search ball does not cover the sphere may come from a missing
        // candidate: search wider.
        const bool partial = r < 2.0 * R;
        int st = status_out;
        if (st == DL_OK) {
            if (any_bad) st = partial ? DL_RETRY : (any_deg ? DL_DEGENERATE : DL_INVALID);
            else if (any_far) st = DL_RETRY;
            else if (any_deg) st = DL_DEGENERATE;
        } else if (partial) {
            st = DL_RETRY;
        }
        status[i] = st;
        deg[i] = d;
    }
    for (int s = lane; s < d; s += 64) nbr[(size_t)i * cap + s] = ring[s];
}

// Search radius of every star for the next pass, from this pass's status:
// a star that must search wider (DL_RETRY) doubles its radius, or bisects
// between the largest radius that was too small and the smallest one that
// held too many candidates; a crowded star (DL_CROWDED) bisects down the same
// way and is redone, until the two bounds are within a thousandth of each
// other (then it stays DL_CROWDED: no radius both holds SVM_DL_MAXC
// candidates and proves the star).  Radii never exceed `whole` (a ball that
// covers the sphere).
__global__ __launch_bounds__(256) void k_dl_radius(int* status, double* rad, double* rlo, double* rhi, int n,
                                                   double whole) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int st = status[i];
    const double r = rad[i];
    if (st == DL_RETRY) {
        rlo[i] = r;
        const double hi = rhi[i];
        const double nr = hi > 0.0 ? 0.5 * (r + hi) : 2.0 * r;
        rad[i] = nr < whole ? nr : whole;
    } else if (st == DL_CROWDED) {
        rhi[i] = r;
        const double lo = rlo[i];
        if (r - lo > 1e-3 * r) {
            rad[i] = lo > 0.0 ? 0.5 * (lo + r) : 0.5 * r;
            status[i] = DL_RETRY;
        }
    }
}

// Outcome of a star pass for the host: out[0] = stars to redo wider,
// out[1] = smallest point whose star failed (n if none).  out = {0, n} on entry.
__global__ __launch_bounds__(256) void k_dl_status(const int* status, int n, int* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int st = i < n ? status[i] : DL_OK;
    const unsigned long long retry = __ballot(st == DL_RETRY);
    if ((threadIdx.x & 63) == 0 && retry) atomicAdd(out, (int)__popcll(retry));
    if (st != DL_OK && st != DL_RETRY) atomicMin(out + 1, i);
}

__global__ __launch_bounds__(256) void k_fill_int(int* a, int n, int value) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = value;
}

__global__ __launch_bounds__(256) void k_fill_double(double* a, int n, double value) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = value;
}

// ---------------------------------------------------------- triangles ----
// Canonical triangles of point m: (m, a, b) for consecutive ring neighbours
// a, b both above m (CCW order kept), sorted by a.
__global__ __launch_bounds__(256) void k_dl_tri_count(int n, int cap, const int* nbr, const int* deg, int* cnt) {
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const int* rg = nbr + (size_t)m * cap;
    int d = deg[m], c = 0;
    for (int j = 0; j < d; ++j) c += (rg[j] > m && rg[(j + 1) % d] > m);
    cnt[m] = c;
}

__global__ __launch_bounds__(256) void k_dl_tri_write(int n, int cap, const int* nbr, const int* deg, const int* off,
                                                      int* T) {
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const int* rg = nbr + (size_t)m * cap;
    int d = deg[m];
    int base = off[m], c = 0;
    for (int j = 0; j < d; ++j) {
        int a = rg[j], b = rg[(j + 1) % d];
        if (a > m && b > m) {
            // insertion by a into this point's slice
            int k = c++;
            while (k > 0 && T[3 * (size_t)(base + k - 1) + 1] > a) {
                for (int e = 0; e < 3; ++e) T[3 * (size_t)(base + k) + e] = T[3 * (size_t)(base + k - 1) + e];
                --k;
            }
            T[3 * (size_t)(base + k)] = m;
            T[3 * (size_t)(base + k) + 1] = a;
            T[3 * (size_t)(base + k) + 2] = b;
        }
    }
}

// index of the canonical triangle (m, a, .) in m's slice; -1 if absent
__device__ inline int dl_tri_index(const int* T, const int* off, int m, int a) {
    for (int t = off[m]; t < off[m + 1]; ++t)
        if (T[3 * (size_t)t + 1] == a) return t;
    return -1;
}

// Per point c: the global index of each triangle (c, r_j, r_{j+1}) of its star.
__global__ __launch_bounds__(256) void k_dl_tid(int n, int cap, const int* nbr, const int* deg, const int* T,
                                                const int* off, int* tid, int* err) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int* rg = nbr + (size_t)c * cap;
    int d = deg[c];
    for (int j = 0; j < d; ++j) {
        int a = rg[j], b = rg[(j + 1) % d];
        int m, s;
        if (c < a && c < b) { m = c; s = a; }
        else if (a < b) { m = a; s = b; }
        else { m = b; s = c; }
        int t = dl_tri_index(T, off, m, s);
        if (t < 0) atomicOr(err, 1);   // stars of neighbouring points disagree
        tid[(size_t)c * cap + j] = t;
    }
}

// Topology of sphere rows: per triangle t (vertex vo + t) the position
// R_hull * normalize(centroid) (convexHullCGAL.cpp:118-123; same operation
// order as the host path), vertex->cell and vertex->vertex lists; per cell
// its ring, starting at its smallest triangle and walking against the CCW
// star order (clockwise seen from outside).
// circum 0: the vertex of triangle t is R normalize(centroid), as the vertex
// model places it (convexHullCGAL.cpp:67-80); circum 1: its circumcentre on
// the sphere, R normalize((b - a) x (c - a)) for the CCW triangle (a, b, c),
// the Voronoi vertex of the three generators.
__global__ __launch_bounds__(256) void k_dl_vertices(Dev d, int sys, int pout, const double* P, int ntri, int cap,
                                                     const int* T, const int* nbr, const int* deg, const int* tid,
                                                     double radius, int* err, int circum) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntri) return;
    const SysParams& S = d.sys[sys];
    int c3[3] = {T[3 * (size_t)t], T[3 * (size_t)t + 1], T[3 * (size_t)t + 2]};
    double c[3] = {0.0, 0.0, 0.0};
    if (circum) {
        const double* a = P + 3 * (size_t)c3[0];
        const double* b = P + 3 * (size_t)c3[1];
        const double* e = P + 3 * (size_t)c3[2];
        const double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        const double w[3] = {e[0] - a[0], e[1] - a[1], e[2] - a[2]};
        c[0] = u[1] * w[2] - u[2] * w[1];
        c[1] = u[2] * w[0] - u[0] * w[2];
        c[2] = u[0] * w[1] - u[1] * w[0];
    } else {
        for (int k = 0; k < 3; ++k) {
            const double* p = P + 3 * (size_t)c3[k];
            c[0] += p[0];
            c[1] += p[1];
            c[2] += p[2];
        }
        for (int k = 0; k < 3; ++k) c[k] = (1.0 / 3) * c[k];
    }
    double nrm = svm::sqrt_rn(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    const int v = S.vo + t;
    d.px[pout][v] = (radius / nrm) * c[0];
    d.py[pout][v] = (radius / nrm) * c[1];
    d.pz[pout][v] = (radius / nrm) * c[2];
    int vn3[3];
    for (int k = 0; k < 3; ++k) {
        int cur = c3[k], prev = c3[(k + 2) % 3];
        // the triangle with directed edge cur -> prev: (cur, prev, .) in cur's star
        const int* rg = nbr + (size_t)cur * cap;
        int j = -1;
        for (int e = 0; e < deg[cur]; ++e)
            if (rg[e] == prev) { j = e; break; }
        if (j < 0) { atomicOr(err, 2); vn3[k] = 0; continue; }
        vn3[k] = tid[(size_t)cur * cap + j];
    }
    d.vcn[v] = make_int4(c3[0] + S.co, c3[1] + S.co, c3[2] + S.co, sys);
    d.vn[v] = make_int4(vn3[0] + S.vo, vn3[1] + S.vo, vn3[2] + S.vo, sys);
}

__global__ __launch_bounds__(256) void k_dl_rings(Dev d, int sys, int n, int cap, const int* deg, const int* tid) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const SysParams& S = d.sys[sys];
    const int dg = deg[c];
    const int* td = tid + (size_t)c * cap;
    int js = 0;
    for (int j = 1; j < dg; ++j)
        if (td[j] < td[js]) js = j;
    for (int k = 0; k < d.cap; ++k) {
        int x = -1;
        if (k < dg) x = S.vo + td[((js - k) % dg + dg) % dg];
        d.ring[(size_t)k * d.Cs + S.co + c] = x;
    }
    d.cnn[S.co + c] = dg;
}

}  // namespace svm


namespace svm {

// Generator points of sphere `sys` (AoS [n][3]) into the SoA planes d.gen.
__global__ __launch_bounds__(256) void k_store_generators(Dev d, int sys, const double* P, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = d.sys[sys].co + i;
    const size_t Cs = (size_t)d.Cs;
    d.gen[c] = P[3 * (size_t)i];
    d.gen[Cs + c] = P[3 * (size_t)i + 1];
    d.gen[2 * Cs + c] = P[3 * (size_t)i + 2];
}

// Generators of sphere `sys` (the three planes of d.gen) as AoS points for
// the triangulation.
__global__ __launch_bounds__(256) void k_gen_aos(Dev d, int sys, double* P) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const SysParams& S = d.sys[sys];
    if (i >= S.nc) return;
    const int c = S.co + i;
    const size_t Cs = (size_t)d.Cs;
    P[3 * (size_t)i] = d.gen[c];
    P[3 * (size_t)i + 1] = d.gen[Cs + c];
    P[3 * (size_t)i + 2] = d.gen[2 * Cs + c];
}

// Brownian move of the generators of a Voronoi-model sphere
// (brownianDynamics::integrateEOMCPU, src/updaters/brownianDynamics.cpp:29-48,
// on the cell degrees of freedom, then sphericalDomain::move,
// inc/sphericalDomain.h:203-209): d = sqrt(2 dt mu T) xi + dt mu F with F the
// generator force (k_voronoi_chain) and xi the step's Philox normals of the
// cell, moved on the sphere of the model's radius.  Thread per cell.
__global__ __launch_bounds__(SVM_TILE) void k_voronoi_move(Dev d) {
    const int tile = blockIdx.x;
    const SysParams* S = d.sys + d.ctile_sys[tile];
    const int c = tile * SVM_TILE + threadIdx.x;
    const int local = c - S->co;
    if (local >= S->nc) return;
    const size_t Cs = (size_t)d.Cs;
    const double forcePrefactor = S->dt * S->mu;
    const double noisePrefactor = svm::sqrt_rn(2.0 * forcePrefactor * S->T);
    double xi[3] = {0.0, 0.0, 0.0};
    if (S->T != 0.0) normal3(S->seed, S->step, (uint32_t)local, xi);
    double disp[3];
    for (int k = 0; k < 3; ++k) disp[k] = xi[k] * noisePrefactor + d.cf[k * Cs + c] * forcePrefactor;
    double p[3] = {d.gen[c], d.gen[Cs + c], d.gen[2 * Cs + c]};
    move_on_sphere(p, disp, S->R);
    d.gen[c] = p[0];
    d.gen[Cs + c] = p[1];
    d.gen[2 * Cs + c] = p[2];
}

// One block per sphere: the step counter that keys the noise advances.
__global__ void k_advance_step(Dev d) {
    if (threadIdx.x == 0) d.sys[blockIdx.x].step += 1;
}

// Voronoi-model force on each generator: the chain rule from the vertex
// forces F_v = -dE/dv (k_vertex<VM_FORCE>) through the circumcentre vertices
// v = s R n / |n|, n = (b - a) x (c - a) of each Delaunay triangle (a, b, c)
// (s = +1 for an outward n).  With dn/da = [c - b]x, dn/db = [a - c]x,
// dn/dc = [b - a]x and y = s (R / |n|) (F_v - n^ (n^ . F_v)), generator a
// receives (b - c) x y, b receives (c - a) x y and c receives (a - b) x y from
// every vertex of its cell; the sum is projected on the tangent plane.
__global__ __launch_bounds__(SVM_TILE) void k_voronoi_chain(Dev d) {
    const int tile = blockIdx.x;
    const SysParams* S = d.sys + d.ctile_sys[tile];
    const int c = tile * SVM_TILE + threadIdx.x;
    if (c - S->co >= S->nc) return;
    const size_t Cs = (size_t)d.Cs;
    const double R = S->R;
    auto gen = [&](int i, double* p) { p[0] = d.gen[i]; p[1] = d.gen[Cs + i]; p[2] = d.gen[2 * Cs + i]; };
    double F[3] = {0.0, 0.0, 0.0};
    const int n = d.cnn[c];
    for (int k = 0; k < n; ++k) {
        const int v = d.ring[(size_t)k * Cs + c];
        const int4 t = d.vcn[v];
        double pa[3], pb[3], pc[3];
        gen(t.x, pa);
        gen(t.y, pb);
        gen(t.z, pc);
        const double u[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
        const double w[3] = {pc[0] - pa[0], pc[1] - pa[1], pc[2] - pa[2]};
        const double nv[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
        const double nn = svm::sqrt_rn(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
        const double nh[3] = {nv[0] / nn, nv[1] / nn, nv[2] / nn};
        const double out = nv[0] * (pa[0] + pb[0] + pc[0]) + nv[1] * (pa[1] + pb[1] + pc[1]) +
                           nv[2] * (pa[2] + pb[2] + pc[2]);
        const double sc = (out < 0.0 ? -1.0 : 1.0) * R / nn;
        const double Fv[3] = {d.fx[v], d.fy[v], d.fz[v]};
        const double fn = nh[0] * Fv[0] + nh[1] * Fv[1] + nh[2] * Fv[2];
        const double y[3] = {sc * (Fv[0] - nh[0] * fn), sc * (Fv[1] - nh[1] * fn), sc * (Fv[2] - nh[2] * fn)};
        double q[3];
        if (c == t.x) { q[0] = pb[0] - pc[0]; q[1] = pb[1] - pc[1]; q[2] = pb[2] - pc[2]; }
        else if (c == t.y) { q[0] = pc[0] - pa[0]; q[1] = pc[1] - pa[1]; q[2] = pc[2] - pa[2]; }
        else { q[0] = pa[0] - pb[0]; q[1] = pa[1] - pb[1]; q[2] = pa[2] - pb[2]; }
        F[0] += q[1] * y[2] - q[2] * y[1];
        F[1] += q[2] * y[0] - q[0] * y[2];
        F[2] += q[0] * y[1] - q[1] * y[0];
    }
    double p[3];
    gen(c, p);
    const double pf = (p[0] * F[0] + p[1] * F[1] + p[2] * F[2]) / (p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    d.cf[c] = F[0] - pf * p[0];
    d.cf[Cs + c] = F[1] - pf * p[1];
    d.cf[2 * Cs + c] = F[2] - pf * p[2];
}

}  // namespace svm
