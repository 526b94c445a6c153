// svm_neighbors.h -- metric neighbour lists of points on the GPU.
//
// Reference behaviour: neighborList::computeNeighborLists
// (src/utilities/neighborList.cpp:102-232, kernel neighborList.cu:108-236):
// each point's neighbours are the other points j with
// |minDist(p_i, p_j)| < range, where minDist is the minimum-image difference
// p_i - p_j in a periodic cube of side 2 * sphereRadius
// (periodicBoundaryConditions::minDist, inc/periodicBoundaryConditions.h:124-133)
// and |d| = sqrt(dot(d, d)) (inc/functions.h:131-138, 170-173); per entry it
// keeps the neighbour's index, d and |d|.
//
// Method (own design): the points are binned in the uniform cube grid of the
// Delaunay cell list (svm_voronoi.h, bin side at least `range`); a thread per
// point scans the bins that meet the box [p - range, p + range] (binning is
// monotone, so no point within range is missed), in a counting pass and then
// a fill pass; each point's entries are sorted by neighbour index, so the
// lists do not depend on the order of the bins or of any atomics.
#pragma once
#include "svm_voronoi.h"

namespace svm {

// d = p_i - p_j and |d| as the reference forms them (no periodic image
// changes d: the host admits only boxes where none can, see svm.h)
__device__ inline double nl_diff(const double* P, int i, int j, double* d) {
    d[0] = P[3 * (size_t)i] - P[3 * (size_t)j];
    d[1] = P[3 * (size_t)i + 1] - P[3 * (size_t)j + 1];
    d[2] = P[3 * (size_t)i + 2] - P[3 * (size_t)j + 2];
    double s = 0.0;
    s += d[0] * d[0];
    s += d[1] * d[1];
    s += d[2] * d[2];
    return __builtin_sqrt(s);
}

// FILL = false: count[i] = number of neighbours of i, *maxc = the largest.
// FILL = true: the entries of point i at [i * M, i * M + count[i]), sorted by
// neighbour index (the caller made M >= every count).
template <bool FILL>
__global__ __launch_bounds__(256) void k_neighbors(const double* P, int n, DlGrid g, const int* start,
                                                   const int* list, double range, int M, int* count, int* maxc,
                                                   int* idx, double* vec, double* dist) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
        const double x = P[3 * (size_t)i + k];
        lo[k] = dl_axis(x - range, g);
        hi[k] = dl_axis(x + range, g);
    }
    int cnt = 0;
    const size_t base = (size_t)i * M;
    for (int gx = lo[0]; gx <= hi[0]; ++gx)
        for (int gy = lo[1]; gy <= hi[1]; ++gy)
            for (int gz = lo[2]; gz <= hi[2]; ++gz) {
                const int b = (gx * g.G + gy) * g.G + gz;
                for (int e = start[b]; e < start[b + 1]; ++e) {
                    const int j = list[e];
                    if (j == i) continue;
                    double d[3];
                    const double r = nl_diff(P, i, j, d);
                    if (r >= range) continue;   // neighborList.cpp:139 (a NaN distance is kept, as there)
                    if (FILL) {
                        idx[base + cnt] = j;
                        vec[3 * (base + cnt)] = d[0];
                        vec[3 * (base + cnt) + 1] = d[1];
                        vec[3 * (base + cnt) + 2] = d[2];
                        dist[base + cnt] = r;
                    }
                    ++cnt;
                }
            }
    if (!FILL) {
        count[i] = cnt;
        atomicMax(maxc, cnt);
        return;
    }
    // insertion sort of the point's entries by neighbour index
    for (int a = 1; a < cnt; ++a) {
        const int j = idx[base + a];
        const double r = dist[base + a];
        const double d0 = vec[3 * (base + a)], d1 = vec[3 * (base + a) + 1], d2 = vec[3 * (base + a) + 2];
        int b = a - 1;
        while (b >= 0 && idx[base + b] > j) {
            idx[base + b + 1] = idx[base + b];
            dist[base + b + 1] = dist[base + b];
            vec[3 * (base + b + 1)] = vec[3 * (base + b)];
            vec[3 * (base + b + 1) + 1] = vec[3 * (base + b) + 1];
            vec[3 * (base + b + 1) + 2] = vec[3 * (base + b) + 2];
            --b;
        }
        idx[base + b + 1] = j;
        dist[base + b + 1] = r;
        vec[3 * (base + b + 1)] = d0;
        vec[3 * (base + b + 1) + 1] = d1;
        vec[3 * (base + b + 1) + 2] = d2;
    }
}

}  // namespace svm
