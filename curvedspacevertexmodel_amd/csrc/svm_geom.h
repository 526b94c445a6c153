// svm_geom.h -- spherical-geometry building blocks of the vertex model in
// "parity mode": every expression tree is the reference's
// (inc/sphericalDomain.h), evaluated with -ffp-contract=off, so results are
// bit-identical to the reference's CPU arithmetic.  The one change is
// bit-preserving common-subexpression reuse: the spherical coordinates of a
// point (r, cos/sin theta, cos/sin phi) are computed once per point instead of
// once per call, and consecutive cross products in a ring walk are reused.
// Identical inputs through identical IEEE operations give identical bits, so
// the reuse cannot change any result.
#pragma once
#include "svm_fp64.h"

namespace svm {

// Spherical coordinates of a point, exactly as each reference gradient
// recomputes them (sphericalDomain.h:250-261, 350-365, 218-225).
struct PQ {
    double x, y, z;
    double cosT, sinT, cosP, sinP;
};

SVM_HD PQ make_pq(double x, double y, double z) {
    PQ q;
    q.x = x; q.y = y; q.z = z;
    double r = __builtin_sqrt(x * x + y * y + z * z);
    q.cosT = z / r;
    q.sinT = __builtin_sqrt(1.0 - q.cosT * q.cosT);
    double s = __builtin_sqrt(x * x + y * y);
    q.cosP = x / s;
    q.sinP = y / s;
    return q;
}

// thetaHat, phiHat at q (sphericalDomain.h:226-231).
SVM_HD void basis(const PQ& q, double* tH, double* pH) {
    tH[0] = q.cosT * q.cosP;
    tH[1] = q.cosT * q.sinP;
    tH[2] = -q.sinT;
    pH[0] = -q.sinP;
    pH[1] = q.cosP;
    pH[2] = 0;
}

// d(arc length)/d v1 (sphericalDomain.h:248-273), v1 = a, v2 = b.
SVM_HD void grad_geodesic(const PQ& a, const PQ& b, const double* tH, const double* pH, double* d) {
    double sinP1MinusP2 = a.sinP * b.cosP - a.cosP * b.sinP;
    double cosP1MinusP2 = a.cosP * b.cosP + a.sinP * b.sinP;
    double denomPart = a.cosT * b.cosT + cosP1MinusP2 * a.sinT * b.sinT;
    double denomInverse = 1. / __builtin_sqrt(1 - denomPart * denomPart);
    double gradTheta = 1.0 * (b.cosT * a.sinT - a.cosT * cosP1MinusP2 * b.sinT) * denomInverse;
    double gradPhi = b.sinT * sinP1MinusP2 * denomInverse;
    d[0] = gradTheta * tH[0] + gradPhi * pH[0];
    d[1] = gradTheta * tH[1] + gradPhi * pH[1];
    d[2] = gradTheta * tH[2] + gradPhi * pH[2];
}

// d(spherical triangle area)/d v1 (sphericalDomain.h:347-411), v1=a, v2=b, v3=c.
SVM_HD void grad_triangle(const PQ& a, const PQ& b, const PQ& c, const double* tH, const double* pH,
                          double radius, double* d) {
    double cosT1 = a.cosT, cosT2 = b.cosT, cosT3 = c.cosT;
    double sinT1 = a.sinT, sinT2 = b.sinT, sinT3 = c.sinT;
    double cosP1 = a.cosP, cosP2 = b.cosP, cosP3 = c.cosP;
    double sinP1 = a.sinP, sinP2 = b.sinP, sinP3 = c.sinP;
    double sinP1MinusP2 = sinP1 * cosP2 - cosP1 * sinP2;
    double sinP1MinusP3 = sinP1 * cosP3 - cosP1 * sinP3;
    double cosP1MinusP2 = cosP1 * cosP2 + sinP1 * sinP2;
    double cosP1MinusP3 = cosP1 * cosP3 + sinP1 * sinP3;
    double cosP2MinusP3 = cosP2 * cosP3 + sinP2 * sinP3;
    double s12 = cosT1 * cosT2 + cosP1MinusP2 * sinT1 * sinT2;
    double s13 = cosT1 * cosT3 + cosP1MinusP3 * sinT1 * sinT3;
    double s23 = cosT2 * cosT3 + cosP2MinusP3 * sinT2 * sinT3;
    double d12 = cosT1 * cosP1MinusP2 * sinT2 - cosT2 * sinT1;
    double d13 = cosT1 * cosP1MinusP3 * sinT3 - cosT3 * sinT1;
    double g = 1 - s12 * s12 - s13 * s13 - s23 * s23 + 2.0 * s12 * s13 * s23;
    double denom1 = __builtin_sqrt((-1. + s12 * s12) * (-1. + s12 * s12) * g);
    double denom2 = __builtin_sqrt((-1. + s13 * s13) * (-1. + s13 * s13) * g);
    double denom3 = __builtin_sqrt((s12 * s12 - 1.) * (s13 * s13 - 1.) * (s12 * s12 - 1.) * (s13 * s13 - 1.) * g);
    double gradTheta = (d13 * (s12 * s12 - 1.0) + d12 * (s23 - s12 * s13)) / denom1
                     + (d12 * (s13 * s13 - 1.0) + d13 * (s23 - s12 * s13)) / denom2
                     - (d12 * (s13 * s13 - 1.) * (s13 - s12 * s23) + d13 * (s12 * s12 - 1.0) * (s12 - s13 * s23)) / denom3;
    gradTheta *= radius;
    double gradPhi = ((s12 * s13 - s23) * sinP1MinusP2 * sinT2 - (-1 + s12 * s12) * sinP1MinusP3 * sinT3) / (denom1)
                   - ((-1 + s13 * s13) * sinP1MinusP2 * sinT2 + (-(s12 * s13) + s23) * sinP1MinusP3 * sinT3) / (denom2)
                   + ((-1 + s13 * s13) * (s13 - s12 * s23) * sinP1MinusP2 * sinT2 + (-1 + s12 * s12) * (s12 - s13 * s23) *
                      sinP1MinusP3 * sinT3) / (denom3);
    gradPhi *= radius;
    double determinant = a.x * (b.y * c.z - b.z * c.y)
                       + a.y * (b.z * c.x - b.x * c.z)
                       + a.z * (b.x * c.y - b.y * c.x);
    d[0] = gradTheta * tH[0] + gradPhi * pH[0];
    d[1] = gradTheta * tH[1] + gradPhi * pH[1];
    d[2] = gradTheta * tH[2] + gradPhi * pH[2];
    if (determinant < 0) {
        d[0] = -1. * d[0];
        d[1] = -1. * d[1];
        d[2] = -1. * d[2];
    }
}

// Force contribution of one (vertex, cell) pair: the body of the reference's
// per-cell loop (sphericalVertexModel.cpp:386-413).  Returns false when the
// contribution is NaN and must be skipped.
SVM_HD bool pair_force(const PQ& cur, const PQ& last, const PQ& next, const PQ& cpos,
                       double areaDifference, double perimeterDifference, double Kr, double radius,
                       double* fSet) {
    double tH[3], pH[3], t[3];
    basis(cur, tH, pH);
    fSet[0] = 0.0; fSet[1] = 0.0; fSet[2] = 0.0;
    double cp = 2.0 * perimeterDifference;
    double ca = 2.0 * Kr * areaDifference;
    grad_geodesic(cur, last, tH, pH, t);
    fSet[0] -= cp * t[0]; fSet[1] -= cp * t[1]; fSet[2] -= cp * t[2];
    grad_geodesic(cur, next, tH, pH, t);
    fSet[0] -= cp * t[0]; fSet[1] -= cp * t[1]; fSet[2] -= cp * t[2];
    grad_triangle(cur, last, cpos, tH, pH, radius, t);
    fSet[0] -= ca * t[0]; fSet[1] -= ca * t[1]; fSet[2] -= ca * t[2];
    grad_triangle(cur, cpos, next, tH, pH, radius, t);
    fSet[0] -= ca * t[0]; fSet[1] -= ca * t[1]; fSet[2] -= ca * t[2];
    return !(fSet[0] != fSet[0]);
}

// sphericalDomain::move(p, disp) (sphericalDomain.h:203-209): project the
// displacement on the tangent plane, add, put back on the sphere.
SVM_HD void move_on_sphere(double* p, const double* disp, double radius) {
    double d0 = disp[0], d1 = disp[1], d2 = disp[2];
    double dvn = 0.0; dvn += d0 * p[0]; dvn += d1 * p[1]; dvn += d2 * p[2];
    double dnn = 0.0; dnn += p[0] * p[0]; dnn += p[1] * p[1]; dnn += p[2] * p[2];
    double s = dvn / dnn;
    d0 = d0 - s * p[0]; d1 = d1 - s * p[1]; d2 = d2 - s * p[2];
    double q0 = p[0] + d0, q1 = p[1] + d1, q2 = p[2] + d2;
    double nn = 0.0; nn += q0 * q0; nn += q1 * q1; nn += q2 * q2;
    double inv = 1.0 / __builtin_sqrt(nn);
    p[0] = radius * (inv * q0);
    p[1] = radius * (inv * q1);
    p[2] = radius * (inv * q2);
}

// projectToTangentPlaneAndNormalize(v, n) (sphericalDomain.h:173-177).
SVM_HD void project_normalize(double* v, const double* n) {
    double dvn = 0.0; dvn += v[0] * n[0]; dvn += v[1] * n[1]; dvn += v[2] * n[2];
    double dnn = 0.0; dnn += n[0] * n[0]; dnn += n[1] * n[1]; dnn += n[2] * n[2];
    double s = dvn / dnn;
    double a0 = v[0] - s * n[0], a1 = v[1] - s * n[1], a2 = v[2] - s * n[2];
    double nn = 0.0; nn += a0 * a0; nn += a1 * a1; nn += a2 * a2;
    double inv = 1.0 / __builtin_sqrt(nn);
    v[0] = inv * a0; v[1] = inv * a1; v[2] = inv * a2;
}

// rodriguesRotation(v, k, theta) (inc/functions.h:175-181).
SVM_HD void rodrigues(double* v, const double* k, double theta) {
    double cp0 = k[1] * v[2] - k[2] * v[1];
    double cp1 = k[2] * v[0] - k[0] * v[2];
    double cp2 = k[0] * v[1] - k[1] * v[0];
    double c = svm::cos(theta), s = svm::sin(theta);
    double kv = 0.0; kv += k[0] * v[0]; kv += k[1] * v[1]; kv += k[2] * v[2];
    double w = (1.0 - c) * kv;
    double a0 = c * v[0] + s * cp0 + w * k[0];
    double a1 = c * v[1] + s * cp1 + w * k[1];
    double a2 = c * v[2] + s * cp2 + w * k[2];
    v[0] = a0; v[1] = a1; v[2] = a2;
}

// geodesicDistance (sphericalDomain.h:94-112) with the point norms supplied:
// n1 = sqrt(x1*x1+y1*y1+z1*z1), n2 likewise.
SVM_HD double geodesic(const double* p1, const double* p2, double n1, double n2, double radius) {
    double numerator = p1[0] * p2[0] + p1[1] * p2[1] + p1[2] * p2[2];
    double denominator = n1 * n2;
    return radius * svm::acos(numerator / denominator);
}

// The T1 edge test `geodesic(p1, p2, n1, n2, R) < thr`
// (src/model/sphericalVertexModel.cpp:478-480) with a pre-filter that leaves
// the decision unchanged.  acos is decreasing with slope magnitude >= 1, so a
// cosine x <= cos(thr/R) - 1e-9 gives acos(x) >= thr/R + 1e-9, i.e. an arc at
// least 1e-9 R above thr -- many orders beyond the <= 1 ulp errors of the
// exact expression.  Only the remaining edges evaluate acos.
SVM_HD double t1_cos_cut(double thr, double radius) { return svm::cos(thr / radius) - 1e-9; }
SVM_HD bool short_edge(const double* p1, const double* p2, double n1, double n2, double radius, double thr,
                       double cos_cut) {
    double numerator = p1[0] * p2[0] + p1[1] * p2[1] + p1[2] * p2[2];
    double denominator = n1 * n2;
    double x = numerator / denominator;
    if (x <= cos_cut) return false;
    return radius * svm::acos(x) < thr;
}

SVM_HD double pnorm(const double* p) { return __builtin_sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]); }

// sphericalTriangleArea (sphericalDomain.h:135-166).
SVM_HD double triangle_area(const double* p1, const double* p2, const double* p3, double radius) {
    double a[3], b[3], c[3];
    double dn;
    dn = 0.0; dn += p1[0] * p1[0]; dn += p1[1] * p1[1]; dn += p1[2] * p1[2];
    double ia = 1.0 / __builtin_sqrt(dn);
    dn = 0.0; dn += p2[0] * p2[0]; dn += p2[1] * p2[1]; dn += p2[2] * p2[2];
    double ib = 1.0 / __builtin_sqrt(dn);
    dn = 0.0; dn += p3[0] * p3[0]; dn += p3[1] * p3[1]; dn += p3[2] * p3[2];
    double ic = 1.0 / __builtin_sqrt(dn);
    for (int k = 0; k < 3; ++k) { a[k] = ia * p1[k]; b[k] = ib * p2[k]; c[k] = ic * p3[k]; }
    double p1Dotp2 = 0.0; p1Dotp2 += a[0] * b[0]; p1Dotp2 += a[1] * b[1]; p1Dotp2 += a[2] * b[2];
    double p1Dotp3 = 0.0; p1Dotp3 += a[0] * c[0]; p1Dotp3 += a[1] * c[1]; p1Dotp3 += a[2] * c[2];
    double p2Dotp3 = 0.0; p2Dotp3 += b[0] * c[0]; p2Dotp3 += b[1] * c[1]; p2Dotp3 += b[2] * c[2];
    double area = -SVM_PI;
    area += svm::acos((p2Dotp3 - p1Dotp2 * p1Dotp3) / (__builtin_sqrt(1 - p1Dotp2 * p1Dotp2) * __builtin_sqrt(1 - p1Dotp3 * p1Dotp3)));
    area += svm::acos((p1Dotp2 - p1Dotp3 * p2Dotp3) / (__builtin_sqrt(1 - p1Dotp3 * p1Dotp3) * __builtin_sqrt(1 - p2Dotp3 * p2Dotp3)));
    area += svm::acos((p1Dotp3 - p1Dotp2 * p2Dotp3) / (__builtin_sqrt(1 - p1Dotp2 * p1Dotp2) * __builtin_sqrt(1 - p2Dotp3 * p2Dotp3)));
    area *= (radius * radius);
    return area;
}

}  // namespace svm
