// svm_exact.h -- exact orientation predicates for points on a sphere.
//
// Method: a floating-point evaluation with a forward error bound decides the
// sign when it can (the overwhelmingly common case); otherwise the
// determinant is evaluated exactly as a non-overlapping expansion of doubles
// (J. R. Shewchuk, "Adaptive precision floating-point arithmetic and fast
// robust geometric predicates", Discrete Comput. Geom. 18, 1997): products by
// fused multiply-add error terms, sums by two-sum, and the sign of an
// expansion is the sign of its largest component.  Inputs must be finite and
// far from overflow/underflow (true for cell centres on spheres of radius
// 1e-3 .. 1e5).  Exactly zero determinants (four co-circular points) are
// reported as 0; callers break such ties symbolically.
#pragma once
#include "svm_fp64.h"

namespace svm {

SVM_HD void two_sum(double a, double b, double& x, double& y) {
    x = a + b;
    double bv = x - a;
    double av = x - bv;
    double br = b - bv;
    double ar = a - av;
    y = ar + br;
}

SVM_HD void two_prod(double a, double b, double& x, double& y) {
    x = a * b;
    y = __builtin_fma(a, b, -x);
}

// h = e + b for a non-overlapping expansion e (increasing magnitude), zero
// components dropped.  h may not alias e.  Returns the length of h.
SVM_HD int grow_expansion(int elen, const double* e, double b, double* h) {
    double q = b;
    int hlen = 0;
    for (int i = 0; i < elen; ++i) {
        double s, r;
        two_sum(q, e[i], s, r);
        q = s;
        if (r != 0.0) h[hlen++] = r;
    }
    if (q != 0.0 || hlen == 0) h[hlen++] = q;
    return hlen;
}

// Add the exact product a*b*c (c may be 1) into expansion (len, buf): the
// caller passes a scratch buffer tmp of the same capacity.
SVM_HD int add_product3(int len, double* buf, double* tmp, double a, double b, double c) {
    double p, pe;
    two_prod(a, b, p, pe);
    double x0, x1, y0, y1;
    two_prod(p, c, x0, x1);     // p*c = x0 + x1
    two_prod(pe, c, y0, y1);    // pe*c = y0 + y1 (exact parts)
    double parts[4] = {x1, y1, y0, x0};
    for (int k = 0; k < 4; ++k) {
        if (parts[k] == 0.0) continue;
        int n = grow_expansion(len, buf, parts[k], tmp);
        for (int i = 0; i < n; ++i) buf[i] = tmp[i];
        len = n;
    }
    return len;
}

SVM_HD int expansion_sign(int len, const double* e) {
    for (int i = len - 1; i >= 0; --i) {
        if (e[i] > 0) return 1;
        if (e[i] < 0) return -1;
    }
    return 0;
}

// Exact sign of sum_k sgn_k * det3(rows) over the listed raw coordinate
// triples; det3(x, y, z) = x . (y x z) is six signed triple products.
SVM_HD int exact_det3_sum(const double* const* X, const double* const* Y, const double* const* Z,
                          const double* sg, int nterms) {
    double buf[256], tmp[256];
    int len = 0;
    for (int t = 0; t < nterms; ++t) {
        const double* x = X[t];
        const double* y = Y[t];
        const double* z = Z[t];
        double s = sg[t];
        len = add_product3(len, buf, tmp, s * x[0], y[1], z[2]);
        len = add_product3(len, buf, tmp, -s * x[0], y[2], z[1]);
        len = add_product3(len, buf, tmp, s * x[1], y[2], z[0]);
        len = add_product3(len, buf, tmp, -s * x[1], y[0], z[2]);
        len = add_product3(len, buf, tmp, s * x[2], y[0], z[1]);
        len = add_product3(len, buf, tmp, -s * x[2], y[1], z[0]);
    }
    return expansion_sign(len, buf);
}

// The exact evaluations take their points by value and are kept out of line:
// the expansion code indexes arrays of point pointers, which would otherwise
// force the callers' coordinates out of registers into private memory on
// every call, the common filtered case included.
SVM_HD_OUTLINE int orient_origin_exact(double a0, double a1, double a2, double b0, double b1, double b2,
                                       double c0, double c1, double c2) {
    const double a[3] = {a0, a1, a2}, b[3] = {b0, b1, b2}, c[3] = {c0, c1, c2};
    const double* X[1] = {a};
    const double* Y[1] = {b};
    const double* Z[1] = {c};
    double sg[1] = {1.0};
    return exact_det3_sum(X, Y, Z, sg, 1);
}

// det[b-a, c-a, d-a] = det3(b,c,d) - det3(a,c,d) + det3(a,b,d) - det3(a,b,c)
SVM_HD_OUTLINE int orient3d_exact(double a0, double a1, double a2, double b0, double b1, double b2, double c0,
                                  double c1, double c2, double d0, double d1, double d2) {
    const double a[3] = {a0, a1, a2}, b[3] = {b0, b1, b2}, c[3] = {c0, c1, c2}, d[3] = {d0, d1, d2};
    const double* X[4] = {b, a, a, a};
    const double* Y[4] = {c, c, b, b};
    const double* Z[4] = {d, d, d, c};
    double sg[4] = {1.0, -1.0, 1.0, -1.0};
    return exact_det3_sum(X, Y, Z, sg, 4);
}

// Sign of det[a, b, c] = a . (b x c): +1 when (a, b, c) is counter-clockwise
// seen from outside a sphere centred at the origin.
SVM_HD int orient_origin_sign(const double* a, const double* b, const double* c) {
    double m0 = b[1] * c[2], m1 = b[2] * c[1];
    double m2 = b[2] * c[0], m3 = b[0] * c[2];
    double m4 = b[0] * c[1], m5 = b[1] * c[0];
    double det = a[0] * (m0 - m1) + a[1] * (m2 - m3) + a[2] * (m4 - m5);
    double perm = __builtin_fabs(a[0]) * (__builtin_fabs(m0) + __builtin_fabs(m1)) +
                  __builtin_fabs(a[1]) * (__builtin_fabs(m2) + __builtin_fabs(m3)) +
                  __builtin_fabs(a[2]) * (__builtin_fabs(m4) + __builtin_fabs(m5));
    double bound = 1e-14 * perm;
    if (det > bound) return 1;
    if (det < -bound) return -1;
    return orient_origin_exact(a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2]);
}

// Sign of det[b-a, c-a, d-a].  For (a, b, c) counter-clockwise seen from
// outside, +1 means d lies beyond the plane of a, b, c (inside their
// circumcircle on the sphere), -1 on the origin's side.
SVM_HD int orient3d_sign(const double* a, const double* b, const double* c, const double* d) {
    double bx = b[0] - a[0], by = b[1] - a[1], bz = b[2] - a[2];
    double cx = c[0] - a[0], cy = c[1] - a[1], cz = c[2] - a[2];
    double dx = d[0] - a[0], dy = d[1] - a[1], dz = d[2] - a[2];
    double m0 = cy * dz, m1 = cz * dy, m2 = cz * dx, m3 = cx * dz, m4 = cx * dy, m5 = cy * dx;
    double det = bx * (m0 - m1) + by * (m2 - m3) + bz * (m4 - m5);
    double perm = __builtin_fabs(bx) * (__builtin_fabs(m0) + __builtin_fabs(m1)) +
                  __builtin_fabs(by) * (__builtin_fabs(m2) + __builtin_fabs(m3)) +
                  __builtin_fabs(bz) * (__builtin_fabs(m4) + __builtin_fabs(m5));
    double bound = 1e-14 * perm;
    if (det > bound) return 1;
    if (det < -bound) return -1;
    return orient3d_exact(a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2], d[0], d[1], d[2]);
}

}  // namespace svm
