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

// Scalar-type-generic square root: the FP64 path (parity) and the FP32 force
// variant share the expression trees below.
SVM_HD double vsqrt(double x) { return __builtin_sqrt(x); }
SVM_HD float vsqrt(float x) { return __builtin_sqrtf(x); }

// How the gradient expressions below take their square roots and quotients.
// ExactOps: the language's correctly rounded operators, for every argument.
// A quotient is written div(a, b, rb) with rb = rcp(b) prepared once per
// divisor; here rcp does nothing and div is a / b.
struct ExactOps {
    template <class T> SVM_HD T sqrt(T x) const { return vsqrt(x); }
    template <class T> SVM_HD T rcp(T b) const { return b; }
    template <class T> SVM_HD T div(T a, T b, T) const { return a / b; }
};

// FastOps (FP64, device): the same correctly rounded results from shorter
// sequences that are exact only for ordinary arguments.  Where an argument
// or a result leaves the range in which they are exact, the result is
// replaced by a NaN, which every later operation of the force expressions
// carries into the force contribution, so the contribution comes out NaN; a
// caller that sees any NaN contribution evaluates the vertex again with
// ExactOps (which also decides the reference's NaN skips).
//  * sqrt: the Newton refinement of v_rsq_f64 that the compiler's root wraps
//    (sqrt_in_range); exact for x in [2^-767, DBL_MAX], accepted here on
//    [2^-767, 2^1000), so that every root lies in [2^-384, 2^500).
//  * rcp + div: the refined reciprocal y of the divisor (v_rcp_f64 and two
//    Newton steps), shared by the quotients with that divisor, then
//    q0 = a y, r = fma(-b, q0, a) (exact), q = fma(r, y, q0): the quotient
//    sequence of the compiler's division without its operand scaling
//    (div_scale), scaled fma (div_fmas) and special-case fixup (div_fixup).
//    Divisors here are roots accepted by sqrt, so |b| is in [2^-384, 2^500);
//    the quotient is accepted on [2^-400, 2^400), which keeps a = q b, a y,
//    the remainder and the reciprocal's steps normal and finite.  There a
//    scaling by a power of two, if the operator applies one, commutes with
//    every rounding, and the fixup passes ordinary values through, so the
//    result is the division operator's, bit for bit: the same seed and the
//    same refinement.  Zero, infinite and NaN quotients are replaced.
// A check's result is consumed at once by the select that replaces the
// value, so no flag stays live across the expressions.
struct FastOps {
    __device__ static uint32_t hi(double x) { return (uint32_t)(__builtin_bit_cast(uint64_t, x) >> 32); }
    // x where ok, else a quiet NaN (only the high word is selected)
    __device__ static double keep(double x, bool ok) {
        const uint64_t b = __builtin_bit_cast(uint64_t, x);
        const uint32_t h = ok ? (uint32_t)(b >> 32) : 0x7ff80000u;
        return __builtin_bit_cast(double, ((uint64_t)h << 32) | (b & 0xffffffffull));
    }
    __device__ double sqrt(double x) const {
        return keep(sqrt_in_range(x), (hi(x) - 0x10000000u) < (0x7E700000u - 0x10000000u));
    }
    __device__ double rcp(double b) const {
#if defined(__HIP_DEVICE_COMPILE__)
        const double y0 = __builtin_amdgcn_rcp(b);
        const double e0 = __builtin_fma(-b, y0, 1.0);
        const double y1 = __builtin_fma(y0, e0, y0);
        const double e1 = __builtin_fma(-b, y1, 1.0);
        return __builtin_fma(y1, e1, y1);
#else
        return 1.0 / b;   // the host pass parses device code; it never runs it
#endif
    }
    __device__ double div(double a, double b, double y) const {
        const double q0 = a * y;
        const double r = __builtin_fma(-b, q0, a);
        const double q = __builtin_fma(r, y, q0);
        return keep(q, ((hi(q) & 0x7fffffffu) - 0x26F00000u) < (0x58F00000u - 0x26F00000u));
    }
};

// Spherical coordinates of a point, exactly as each reference gradient
// recomputes them (sphericalDomain.h:250-261, 350-365, 218-225).
template <class T>
struct PQT {
    T x, y, z;
    T cosT, sinT, cosP, sinP;
};
using PQ = PQT<double>;

template <class T, class Ops = ExactOps>
SVM_HD PQT<T> make_pq(T x, T y, T z, const Ops& ops = Ops()) {
    PQT<T> q;
    q.x = x; q.y = y; q.z = z;
    T r = ops.sqrt(x * x + y * y + z * z);
    q.cosT = ops.div(z, r, ops.rcp(r));
    q.sinT = ops.sqrt(T(1.0) - q.cosT * q.cosT);
    T s = ops.sqrt(x * x + y * y);
    const T rs = ops.rcp(s);
    q.cosP = ops.div(x, s, rs);
    q.sinP = ops.div(y, s, rs);
    return q;
}

// thetaHat, phiHat at q (sphericalDomain.h:226-231).
template <class T>
SVM_HD void basis(const PQT<T>& q, T* tH, T* pH) {
    tH[0] = q.cosT * q.cosP;
    tH[1] = q.cosT * q.sinP;
    tH[2] = -q.sinT;
    pH[0] = -q.sinP;
    pH[1] = q.cosP;
    pH[2] = T(0);
}

// d(arc length)/d v1 (sphericalDomain.h:248-273), v1 = a, v2 = b.
template <class T>
SVM_HD void grad_geodesic(const PQT<T>& a, const PQT<T>& b, const T* tH, const T* pH, T* d) {
    T sinP1MinusP2 = a.sinP * b.cosP - a.cosP * b.sinP;
    T cosP1MinusP2 = a.cosP * b.cosP + a.sinP * b.sinP;
    T denomPart = a.cosT * b.cosT + cosP1MinusP2 * a.sinT * b.sinT;
    T denomInverse = T(1.) / vsqrt(T(1) - denomPart * denomPart);
    T gradTheta = T(1.0) * (b.cosT * a.sinT - a.cosT * cosP1MinusP2 * b.sinT) * denomInverse;
    T gradPhi = b.sinT * sinP1MinusP2 * denomInverse;
    d[0] = gradTheta * tH[0] + gradPhi * pH[0];
    d[1] = gradTheta * tH[1] + gradPhi * pH[1];
    d[2] = gradTheta * tH[2] + gradPhi * pH[2];
}

// d(spherical triangle area)/d v1 (sphericalDomain.h:347-411), v1=a, v2=b, v3=c.
template <class T>
SVM_HD void grad_triangle(const PQT<T>& a, const PQT<T>& b, const PQT<T>& c, const T* tH, const T* pH, T radius,
                          T* d) {
    T cosT1 = a.cosT, cosT2 = b.cosT, cosT3 = c.cosT;
    T sinT1 = a.sinT, sinT2 = b.sinT, sinT3 = c.sinT;
    T cosP1 = a.cosP, cosP2 = b.cosP, cosP3 = c.cosP;
    T sinP1 = a.sinP, sinP2 = b.sinP, sinP3 = c.sinP;
    T sinP1MinusP2 = sinP1 * cosP2 - cosP1 * sinP2;
    T sinP1MinusP3 = sinP1 * cosP3 - cosP1 * sinP3;
    T cosP1MinusP2 = cosP1 * cosP2 + sinP1 * sinP2;
    T cosP1MinusP3 = cosP1 * cosP3 + sinP1 * sinP3;
    T cosP2MinusP3 = cosP2 * cosP3 + sinP2 * sinP3;
    T s12 = cosT1 * cosT2 + cosP1MinusP2 * sinT1 * sinT2;
    T s13 = cosT1 * cosT3 + cosP1MinusP3 * sinT1 * sinT3;
    T s23 = cosT2 * cosT3 + cosP2MinusP3 * sinT2 * sinT3;
    T d12 = cosT1 * cosP1MinusP2 * sinT2 - cosT2 * sinT1;
    T d13 = cosT1 * cosP1MinusP3 * sinT3 - cosT3 * sinT1;
    T g = T(1) - s12 * s12 - s13 * s13 - s23 * s23 + T(2.0) * s12 * s13 * s23;
    T denom1 = vsqrt((T(-1.) + s12 * s12) * (T(-1.) + s12 * s12) * g);
    T denom2 = vsqrt((T(-1.) + s13 * s13) * (T(-1.) + s13 * s13) * g);
    T denom3 = vsqrt((s12 * s12 - T(1.)) * (s13 * s13 - T(1.)) * (s12 * s12 - T(1.)) * (s13 * s13 - T(1.)) * g);
    T gradTheta = (d13 * (s12 * s12 - T(1.0)) + d12 * (s23 - s12 * s13)) / denom1
                + (d12 * (s13 * s13 - T(1.0)) + d13 * (s23 - s12 * s13)) / denom2
                - (d12 * (s13 * s13 - T(1.)) * (s13 - s12 * s23) + d13 * (s12 * s12 - T(1.0)) * (s12 - s13 * s23)) /
                      denom3;
    gradTheta *= radius;
    T gradPhi = ((s12 * s13 - s23) * sinP1MinusP2 * sinT2 - (T(-1) + s12 * s12) * sinP1MinusP3 * sinT3) / (denom1)
              - ((T(-1) + s13 * s13) * sinP1MinusP2 * sinT2 + (-(s12 * s13) + s23) * sinP1MinusP3 * sinT3) / (denom2)
              + ((T(-1) + s13 * s13) * (s13 - s12 * s23) * sinP1MinusP2 * sinT2 + (T(-1) + s12 * s12) * (s12 - s13 * s23) *
                 sinP1MinusP3 * sinT3) / (denom3);
    gradPhi *= radius;
    T determinant = a.x * (b.y * c.z - b.z * c.y)
                  + a.y * (b.z * c.x - b.x * c.z)
                  + a.z * (b.x * c.y - b.y * c.x);
    d[0] = gradTheta * tH[0] + gradPhi * pH[0];
    d[1] = gradTheta * tH[1] + gradPhi * pH[1];
    d[2] = gradTheta * tH[2] + gradPhi * pH[2];
    if (determinant < 0) {
        d[0] = T(-1.) * d[0];
        d[1] = T(-1.) * d[1];
        d[2] = T(-1.) * d[2];
    }
}

// Terms of the pair (a, b) shared by the geodesic and triangle gradients at a
// (sphericalDomain.h:248-273, 347-411): sin and cos of phi_a - phi_b, the
// cosine s of the angle between a and b, and dd = cosT_a cos(phi_a - phi_b)
// sinT_b - cosT_b sinT_a, in the reference's operation order.
template <class T>
struct PairT {
    T sinPd, cosPd, s, dd;
};
template <class T>
SVM_HD PairT<T> pair_terms(const PQT<T>& a, const PQT<T>& b) {
    PairT<T> p;
    p.sinPd = a.sinP * b.cosP - a.cosP * b.sinP;
    p.cosPd = a.cosP * b.cosP + a.sinP * b.sinP;
    p.s = a.cosT * b.cosT + p.cosPd * a.sinT * b.sinT;
    p.dd = a.cosT * p.cosPd * b.sinT - b.cosT * a.sinT;
    return p;
}

// grad_geodesic with the pair terms supplied.
template <class T, class Ops = ExactOps>
SVM_HD void grad_geodesic_p(const PQT<T>& a, const PQT<T>& b, const PairT<T>& P, const T* tH, const T* pH, T* d,
                            const Ops& ops = Ops()) {
    const T den = ops.sqrt(T(1) - P.s * P.s);
    T denomInverse = ops.div(T(1.), den, ops.rcp(den));
    T gradTheta = T(1.0) * (b.cosT * a.sinT - a.cosT * P.cosPd * b.sinT) * denomInverse;
    T gradPhi = b.sinT * P.sinPd * denomInverse;
    d[0] = gradTheta * tH[0] + gradPhi * pH[0];
    d[1] = gradTheta * tH[1] + gradPhi * pH[1];
    d[2] = gradTheta * tH[2] + gradPhi * pH[2];
}

// grad_triangle with the (a, b) and (a, c) pair terms supplied (pair_terms):
// the same expression trees, each pair evaluated once per vertex.
template <class T, class Ops = ExactOps>
SVM_HD void grad_triangle_p(const PQT<T>& a, const PQT<T>& b, const PQT<T>& c, const PairT<T>& P12,
                            const PairT<T>& P13, const T* tH, const T* pH, T radius, T* d, const Ops& ops = Ops()) {
    T cosT1 = a.cosT, cosT2 = b.cosT, cosT3 = c.cosT;
    T sinT1 = a.sinT, sinT2 = b.sinT, sinT3 = c.sinT;
    T cosP1 = a.cosP, cosP2 = b.cosP, cosP3 = c.cosP;
    T sinP1 = a.sinP, sinP2 = b.sinP, sinP3 = c.sinP;
    T sinP1MinusP2 = P12.sinPd;
    T sinP1MinusP3 = P13.sinPd;
    T cosP2MinusP3 = cosP2 * cosP3 + sinP2 * sinP3;
    T s12 = P12.s;
    T s13 = P13.s;
    T s23 = cosT2 * cosT3 + cosP2MinusP3 * sinT2 * sinT3;
    T d12 = P12.dd;
    T d13 = P13.dd;
    T g = T(1) - s12 * s12 - s13 * s13 - s23 * s23 + T(2.0) * s12 * s13 * s23;
    T denom1 = ops.sqrt((T(-1.) + s12 * s12) * (T(-1.) + s12 * s12) * g);
    T denom2 = ops.sqrt((T(-1.) + s13 * s13) * (T(-1.) + s13 * s13) * g);
    T denom3 = ops.sqrt((s12 * s12 - T(1.)) * (s13 * s13 - T(1.)) * (s12 * s12 - T(1.)) * (s13 * s13 - T(1.)) * g);
    const T r1 = ops.rcp(denom1), r2 = ops.rcp(denom2), r3 = ops.rcp(denom3);
    T gradTheta = ops.div(d13 * (s12 * s12 - T(1.0)) + d12 * (s23 - s12 * s13), denom1, r1)
                + ops.div(d12 * (s13 * s13 - T(1.0)) + d13 * (s23 - s12 * s13), denom2, r2)
                - ops.div(d12 * (s13 * s13 - T(1.)) * (s13 - s12 * s23) + d13 * (s12 * s12 - T(1.0)) * (s12 - s13 * s23),
                          denom3, r3);
    gradTheta *= radius;
    T gradPhi = ops.div((s12 * s13 - s23) * sinP1MinusP2 * sinT2 - (T(-1) + s12 * s12) * sinP1MinusP3 * sinT3, denom1, r1)
              - ops.div((T(-1) + s13 * s13) * sinP1MinusP2 * sinT2 + (-(s12 * s13) + s23) * sinP1MinusP3 * sinT3,
                        denom2, r2)
              + ops.div((T(-1) + s13 * s13) * (s13 - s12 * s23) * sinP1MinusP2 * sinT2 +
                            (T(-1) + s12 * s12) * (s12 - s13 * s23) * sinP1MinusP3 * sinT3,
                        denom3, r3);
    gradPhi *= radius;
    T determinant = a.x * (b.y * c.z - b.z * c.y)
                  + a.y * (b.z * c.x - b.x * c.z)
                  + a.z * (b.x * c.y - b.y * c.x);
    d[0] = gradTheta * tH[0] + gradPhi * pH[0];
    d[1] = gradTheta * tH[1] + gradPhi * pH[1];
    d[2] = gradTheta * tH[2] + gradPhi * pH[2];
    if (determinant < 0) {
        d[0] = T(-1.) * d[0];
        d[1] = T(-1.) * d[1];
        d[2] = T(-1.) * d[2];
    }
}

// Force contribution of one (vertex, cell) pair: the body of the reference's
// per-cell loop (sphericalVertexModel.cpp:386-413).  Returns false when the
// contribution is NaN and must be skipped.
template <class T>
SVM_HD bool pair_force(const PQT<T>& cur, const PQT<T>& last, const PQT<T>& next, const PQT<T>& cpos,
                       T areaDifference, T perimeterDifference, T Kr, T radius, T* fSet) {
    T tH[3], pH[3], t[3];
    basis(cur, tH, pH);
    fSet[0] = T(0.0); fSet[1] = T(0.0); fSet[2] = T(0.0);
    T cp = T(2.0) * perimeterDifference;
    T ca = T(2.0) * Kr * areaDifference;
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

// ------------------------------------------------- vector-form gradients ----
// The same tangent-plane gradients as grad_geodesic / grad_triangle, written
// in a form that evaluates well in single precision (the FP32 force path)
// and, in double precision, stays accurate on large spheres where the
// reference's expressions lose digits.  The reference's trig-basis
// expressions cancel catastrophically in FP32 for the small angles between
// neighbouring points (the Gram determinant g of a small triangle is ~1e-5
// from O(1) terms).  Here every quantity is built from the unit vector a of
// the vertex and the scaled differences e = (u - v)/R of its neighbours,
// formed in FP64, so no O(1) terms cancel.
template <class T>
struct V3 {
    T x, y, z;
};
template <class T> SVM_HD V3<T> operator+(V3<T> a, V3<T> b) { return V3<T>{a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> SVM_HD V3<T> operator-(V3<T> a, V3<T> b) { return V3<T>{a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> SVM_HD V3<T> operator*(T s, V3<T> a) { return V3<T>{s * a.x, s * a.y, s * a.z}; }
template <class T> SVM_HD T dot3(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> SVM_HD V3<T> cross3(V3<T> a, V3<T> b) {
    return V3<T>{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
using f3 = V3<float>;

// d(arc length |v u|)/d v on the sphere: -(e - (a.e) a) / |a x e|
// (the angle alpha between a and a + e has sin(alpha) = |a x e|).
template <class T>
SVM_HD V3<T> grad_arc_v(V3<T> a, V3<T> e) {
    V3<T> t = e - dot3(a, e) * a;
    V3<T> x = cross3(a, e);
    return (T(-1) / vsqrt(dot3(x, x))) * t;
}

// d(signed area of the spherical triangle v, b, c)/d v for b = a + eb,
// c = a + ec: tan(E/2) = N/D with N = a.(b x c) = a.(eb x ec) and
// D = 1 + a.b + b.c + c.a = 4 - (|eb|^2 + |ec|^2 + |eb - ec|^2)/2, so
// grad_a E = 2 (D (b x c) - N (b + c)) / (N^2 + D^2), projected on the
// tangent plane; area = R^2 E and d/dv = (1/R) d/da.  The fan triangles
// (v, last, C) and (v, C, next) of a clockwise ring have N > 0; one that
// folds over (a non-convex cell) has N < 0 and counts negatively, so the fan
// areas keep summing to the cell's angle-sum area -- as in the reference's
// chain rule (checked against finite differences of the energy).
template <class T>
SVM_HD V3<T> grad_tri_v(V3<T> a, V3<T> eb, V3<T> ec, T radius) {
    V3<T> bxc = cross3(a, ec - eb) + cross3(eb, ec);
    T N = dot3(a, cross3(eb, ec));
    V3<T> dbc = eb - ec;
    T D = T(4) - T(0.5) * (dot3(eb, eb) + dot3(ec, ec) + dot3(dbc, dbc));
    V3<T> tb = bxc - N * a;                       // tangent part of b x c (a.(b x c) = N)
    V3<T> s = eb + ec;
    V3<T> ts = s - dot3(a, s) * a;                // tangent part of b + c
    T k = T(2) * radius / (N * N + D * D);
    return k * (D * tb - N * ts);
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

// rodriguesRotation(v, k, theta) (inc/functions.h:175-181), with c = cos(theta)
// and s = sin(theta) given.
SVM_HD void rodrigues_cs(double* v, const double* k, double c, double s) {
    double cp0 = k[1] * v[2] - k[2] * v[1];
    double cp1 = k[2] * v[0] - k[0] * v[2];
    double cp2 = k[0] * v[1] - k[1] * v[0];
    double kv = 0.0; kv += k[0] * v[0]; kv += k[1] * v[1]; kv += k[2] * v[2];
    double w = (1.0 - c) * kv;
    double a0 = c * v[0] + s * cp0 + w * k[0];
    double a1 = c * v[1] + s * cp1 + w * k[1];
    double a2 = c * v[2] + s * cp2 + w * k[2];
    v[0] = a0; v[1] = a1; v[2] = a2;
}
SVM_HD void rodrigues(double* v, const double* k, double theta) {
    rodrigues_cs(v, k, svm::cos(theta), svm::sin(theta));
}

// geodesicDistance (sphericalDomain.h:94-112) with the point norms supplied:
// n1 = sqrt(x1*x1+y1*y1+z1*z1), n2 likewise.
SVM_HD double geodesic(const double* p1, const double* p2, double n1, double n2, double radius) {
    double numerator = p1[0] * p2[0] + p1[1] * p2[1] + p1[2] * p2[2];
    double denominator = n1 * n2;
    return radius * svm::acos_hi(numerator / denominator);
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
    return radius * svm::acos_hi(x) < thr;
}

SVM_HD double pnorm(const double* p) { return svm::sqrt_rn(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]); }

// sphericalTriangleArea (sphericalDomain.h:135-166).
SVM_HD double triangle_area(const double* p1, const double* p2, const double* p3, double radius) {
    double a[3], b[3], c[3];
    double dn;
    dn = 0.0; dn += p1[0] * p1[0]; dn += p1[1] * p1[1]; dn += p1[2] * p1[2];
    double ia = 1.0 / svm::sqrt_rn(dn);
    dn = 0.0; dn += p2[0] * p2[0]; dn += p2[1] * p2[1]; dn += p2[2] * p2[2];
    double ib = 1.0 / svm::sqrt_rn(dn);
    dn = 0.0; dn += p3[0] * p3[0]; dn += p3[1] * p3[1]; dn += p3[2] * p3[2];
    double ic = 1.0 / svm::sqrt_rn(dn);
    for (int k = 0; k < 3; ++k) { a[k] = ia * p1[k]; b[k] = ib * p2[k]; c[k] = ic * p3[k]; }
    double p1Dotp2 = 0.0; p1Dotp2 += a[0] * b[0]; p1Dotp2 += a[1] * b[1]; p1Dotp2 += a[2] * b[2];
    double p1Dotp3 = 0.0; p1Dotp3 += a[0] * c[0]; p1Dotp3 += a[1] * c[1]; p1Dotp3 += a[2] * c[2];
    double p2Dotp3 = 0.0; p2Dotp3 += b[0] * c[0]; p2Dotp3 += b[1] * c[1]; p2Dotp3 += b[2] * c[2];
    double area = -SVM_PI;
    area += svm::acos((p2Dotp3 - p1Dotp2 * p1Dotp3) / (svm::sqrt_rn(1 - p1Dotp2 * p1Dotp2) * svm::sqrt_rn(1 - p1Dotp3 * p1Dotp3)));
    area += svm::acos((p1Dotp2 - p1Dotp3 * p2Dotp3) / (svm::sqrt_rn(1 - p1Dotp3 * p1Dotp3) * svm::sqrt_rn(1 - p2Dotp3 * p2Dotp3)));
    area += svm::acos((p1Dotp3 - p1Dotp2 * p2Dotp3) / (svm::sqrt_rn(1 - p1Dotp2 * p1Dotp2) * svm::sqrt_rn(1 - p2Dotp3 * p2Dotp3)));
    area *= (radius * radius);
    return area;
}

}  // namespace svm
