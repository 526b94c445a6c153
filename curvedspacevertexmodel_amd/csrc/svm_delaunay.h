// svm_delaunay.h -- seeded inputs and the spherical Delaunay triangulation
// that gives the vertex model its initial topology.
#pragma once
#include <stdint.h>
#include "svm_fp64.h"
#include "svm_exact.h"

namespace svm {

// Cell centres of sphericalModel::setParticlePositionsRandomly followed by
// setRadius(sqrt(N/4pi)) (src/model/sphericalModel.cpp:11-51,
// src/model/sphericalVertexModel.cpp:12): theta = acos(2v-1), phi = 2 pi u,
// put on the unit sphere, then rescaled onto radius R.  Draw i is Philox
// stream CELL_POINTS, counter i.
inline void generate_cell_points(int n, uint64_t seed, double* out) {
    const double R = svm::sqrt_rn(n / (4.0 * SVM_PI));
    for (int i = 0; i < n; ++i) {
        double u, v;
        uniform2(seed, STREAM_CELL_POINTS, 0, (uint32_t)i, &u, &v);
        double phi = 2.0 * SVM_PI * u;
        double theta = svm::acos(2.0 * v - 1);
        double p[3] = {1.0 * svm::sin(theta) * svm::cos(phi), 1.0 * svm::sin(theta) * svm::sin(phi),
                       1.0 * svm::cos(theta)};
        for (int pass = 0; pass < 2; ++pass) {
            double radius = pass == 0 ? 1.0 : R;
            double nn = 0.0;
            nn += p[0] * p[0]; nn += p[1] * p[1]; nn += p[2] * p[2];
            double inv = 1.0 / svm::sqrt_rn(nn);
            for (int k = 0; k < 3; ++k) p[k] = radius * (inv * p[k]);
        }
        for (int k = 0; k < 3; ++k) out[3 * (size_t)i + k] = p[k];
    }
}

// Unit tangent directors of the vertex-model constructor
// (src/model/sphericalVertexModel.cpp:58-73); draw i is Philox stream
// DIRECTORS, counter i.
inline void generate_directors(int nv, const double* pos, uint64_t seed, double* out) {
    for (int i = 0; i < nv; ++i) {
        double u2, v2;
        uniform2(seed, STREAM_DIRECTORS, 0, (uint32_t)i, &u2, &v2);
        double phi = 2.0 * SVM_PI * u2;
        double theta = svm::acos(2.0 * v2 - 1);
        double n[3] = {1.0 * svm::sin(theta) * svm::cos(phi), 1.0 * svm::sin(theta) * svm::sin(phi),
                       1.0 * svm::cos(theta)};
        const double* p = pos + 3 * (size_t)i;
        double dvn = 0.0; dvn += n[0] * p[0]; dvn += n[1] * p[1]; dvn += n[2] * p[2];
        double dnn = 0.0; dnn += p[0] * p[0]; dnn += p[1] * p[1]; dnn += p[2] * p[2];
        double s = dvn / dnn;
        double a[3] = {n[0] - s * p[0], n[1] - s * p[1], n[2] - s * p[2]};
        double nn = 0.0; nn += a[0] * a[0]; nn += a[1] * a[1]; nn += a[2] * a[2];
        double inv = 1.0 / svm::sqrt_rn(nn);
        for (int k = 0; k < 3; ++k) out[3 * (size_t)i + k] = inv * a[k];
    }
}

inline int orient_origin(const double* a, const double* b, const double* c) {
    return orient_origin_sign(a, b, c);
}

}  // namespace svm
