// svm_fp64.h -- FP64 elementary functions and counter-based random numbers
// shared bit-for-bit by the HIP kernels (gfx950) and the host code.
//
// Why this file exists: the vertex model's trajectories are chaotic, so one
// ulp of difference between the host libm and the device math library in
// acos/sin/cos eventually flips a T1 decision.  Every transcendental that
// touches the state is therefore computed by the functions below, which use
// only IEEE-754 basic operations (+ - * / sqrt, all correctly rounded on both
// x86-64 and gfx950) and integer bit manipulation.  Compiled with
// -ffp-contract=off on both sides they give identical bits everywhere.
//
// Methods (all published):
//   acos      : fdlibm e_acos.c (Sun Microsystems, 1993) -- rational
//               approximation of asin on [0,0.5], half-angle reduction above.
//   sin, cos  : fdlibm k_sin.c / k_cos.c kernel polynomials on [-pi/4,pi/4]
//               with a three-part Cody-Waite reduction by pi/2 (valid for
//               |x| < 2^19 * pi/2, far beyond the |x| <= 2*pi used here).
//   log       : fdlibm e_log.c -- reduction to [sqrt(2)/2, sqrt(2)], then the
//               odd series of log((1+s)/(1-s)).
//   Philox4x32-10 : Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
//               easy as 1, 2, 3", SC'11.
// Accuracy against glibc is checked by the oracle tests (<= 1 ulp for acos,
// sin, cos, log on the ranges used).
//
// The acos, sin, cos and log code below follows fdlibm's e_acos.c, k_sin.c,
// k_cos.c and e_log.c (their coefficients and evaluation steps), which carry
// this notice:
//
//   Copyright (C) 1993 by Sun Microsystems, Inc. All rights reserved.
//
//   Developed at SunSoft, a Sun Microsystems, Inc. business.
//   Permission to use, copy, modify, and distribute this
//   software is freely granted, provided that this notice
//   is preserved.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define SVM_HD __host__ __device__ inline __attribute__((always_inline))
#define SVM_HD_OUTLINE __host__ __device__ inline __attribute__((noinline))
#else
#define SVM_HD static inline __attribute__((always_inline))
#define SVM_HD_OUTLINE static inline __attribute__((noinline))
#endif

namespace svm {

// double-precision value of pi used by the model (the nearest double to pi).
#define SVM_PI 3.141592653589793115997963468544185161590576171875

SVM_HD uint64_t dbits(double x) { uint64_t u; __builtin_memcpy(&u, &x, 8); return u; }
SVM_HD double bitsd(uint64_t u) { double x; __builtin_memcpy(&x, &u, 8); return x; }
SVM_HD int32_t hiword(double x) { return (int32_t)(dbits(x) >> 32); }
SVM_HD uint32_t loword(double x) { return (uint32_t)dbits(x); }
SVM_HD double withwords(int32_t hi, uint32_t lo) {
    return bitsd(((uint64_t)(uint32_t)hi << 32) | (uint64_t)lo);
}

// ---------------------------------------------------------------- sqrt ----
// Correctly rounded square root, the same bits as __builtin_sqrt.  On the
// device, __builtin_sqrt expands to a Newton refinement of v_rsq_f64 wrapped
// in an exponent scaling (for arguments below 2^-767) and a class check
// (+-0 and +inf return themselves).  For arguments in [2^-767, DBL_MAX] the
// scaling is by 2^0 and the class check passes the refined value through, so
// there the refinement alone is the same computation.
// sqrt_in_range: that refinement alone; the caller guarantees the argument is
// in [2^-767, DBL_MAX].
SVM_HD double sqrt_in_range(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = y * 0.5;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    return g;
#else
    return __builtin_sqrt(x);
#endif
}
// sqrt_rn: the refinement where it is exact; every other argument (tiny,
// zero, negative, infinite, NaN) takes __builtin_sqrt itself.  This saves the
// wrapper's eight instructions where they do nothing, at the cost of a branch.
SVM_HD double sqrt_rn(double x) {
    if (__builtin_expect(!(x >= 0x1p-767 && x <= 0x1.fffffffffffffp+1023), 0)) return __builtin_sqrt(x);
    return sqrt_in_range(x);
}

// ---------------------------------------------------------------- acos ----
// The three argument ranges of the method share one rational approximation
// R(z) = p(z)/q(z); it is evaluated once for the range's z and the three
// range-specific combinations are formed and selected, so a wavefront whose
// lanes fall in different ranges does not execute three divergent paths.
// Each lane's result is exactly the arithmetic of its own range.
SVM_HD double acos(double x) {
    const double pi = 3.14159265358979311600e+00;
    const double pio2_hi = 1.57079632679489655800e+00;
    const double pio2_lo = 6.12323399573676603587e-17;
    const double pS0 = 1.66666666666666657415e-01;
    const double pS1 = -3.25565818622400915405e-01;
    const double pS2 = 2.01212532134862925881e-01;
    const double pS3 = -4.00555345006794114027e-02;
    const double pS4 = 7.91534994289814532176e-04;
    const double pS5 = 3.47933107596021167570e-05;
    const double qS1 = -2.40339491173441421878e+00;
    const double qS2 = 2.02094576023350569471e+00;
    const double qS3 = -6.88283971605453293030e-01;
    const double qS4 = 7.70381505559019352791e-02;
    int32_t hx = hiword(x);
    int32_t ix = hx & 0x7fffffff;
    bool small = ix < 0x3fe00000;                           // |x| < 0.5
    double z = small ? x * x : (hx < 0 ? (1.0 + x) * 0.5 : (1.0 - x) * 0.5);
    double p = z * (pS0 + z * (pS1 + z * (pS2 + z * (pS3 + z * (pS4 + z * pS5)))));
    double q = 1.0 + z * (qS1 + z * (qS2 + z * (qS3 + z * qS4)));
    double r = p / q;
    double s = sqrt_rn(z);
    double df = bitsd(dbits(s) & 0xffffffff00000000ull);
    double c = (z - df * df) / (s + df);
    double res_small = pio2_hi - (x - (pio2_lo - x * r));   // |x| < 0.5
    double res_neg = pi - 2.0 * (s + (r * s - pio2_lo));     // x <= -0.5
    double res_pos = 2.0 * (df + (r * s + c));               // x >= 0.5
    double res = small ? res_small : (hx < 0 ? res_neg : res_pos);
    if (ix <= 0x3c600000) res = pio2_hi + pio2_lo;          // |x| < 2^-57
    if (ix >= 0x3ff00000) {                                 // |x| >= 1
        if (((ix - 0x3ff00000) | (int32_t)loword(x)) == 0) res = hx > 0 ? 0.0 : pi + 2.0 * pio2_lo;
        else res = (x - x) / (x - x);                       // |x| > 1 or NaN
    }
    return res;
}

// acos for arguments that are mostly in [0.5, 1): the cosines of short arcs
// (cell perimeters, the T1 edge test).  There acos takes its x >= 0.5 range
// and nothing else, so evaluating that range alone is the same arithmetic on
// the same operands; any other argument goes to acos itself.
SVM_HD double acos_hi(double x) {
    if (__builtin_expect(!(x >= 0.5 && x < 1.0), 0)) return svm::acos(x);
    const double pS0 = 1.66666666666666657415e-01;
    const double pS1 = -3.25565818622400915405e-01;
    const double pS2 = 2.01212532134862925881e-01;
    const double pS3 = -4.00555345006794114027e-02;
    const double pS4 = 7.91534994289814532176e-04;
    const double pS5 = 3.47933107596021167570e-05;
    const double qS1 = -2.40339491173441421878e+00;
    const double qS2 = 2.02094576023350569471e+00;
    const double qS3 = -6.88283971605453293030e-01;
    const double qS4 = 7.70381505559019352791e-02;
    double z = (1.0 - x) * 0.5;
    double p = z * (pS0 + z * (pS1 + z * (pS2 + z * (pS3 + z * (pS4 + z * pS5)))));
    double q = 1.0 + z * (qS1 + z * (qS2 + z * (qS3 + z * qS4)));
    double r = p / q;
    double s = sqrt_in_range(z);     // z = (1 - x) / 2 is in [2^-54, 0.25]
    double df = bitsd(dbits(s) & 0xffffffff00000000ull);
    double c = (z - df * df) / (s + df);
    return 2.0 * (df + (r * s + c));
}

// acos for arguments that are mostly in (-1, 0.5): the cosines of a cell
// polygon's interior angles (about 120 degrees).  There acos takes its
// |x| < 0.5 and x <= -0.5 ranges; these two are evaluated together and
// selected, without the third range's division.  Any other argument goes to
// acos itself.
SVM_HD double acos_lo(double x) {
    if (__builtin_expect(!(x < 0.5 && x > -1.0), 0)) return svm::acos(x);
    const double pi = 3.14159265358979311600e+00;
    const double pio2_hi = 1.57079632679489655800e+00;
    const double pio2_lo = 6.12323399573676603587e-17;
    const double pS0 = 1.66666666666666657415e-01;
    const double pS1 = -3.25565818622400915405e-01;
    const double pS2 = 2.01212532134862925881e-01;
    const double pS3 = -4.00555345006794114027e-02;
    const double pS4 = 7.91534994289814532176e-04;
    const double pS5 = 3.47933107596021167570e-05;
    const double qS1 = -2.40339491173441421878e+00;
    const double qS2 = 2.02094576023350569471e+00;
    const double qS3 = -6.88283971605453293030e-01;
    const double qS4 = 7.70381505559019352791e-02;
    const int32_t ix = hiword(x) & 0x7fffffff;
    const bool small = ix < 0x3fe00000;                     // |x| < 0.5; otherwise x <= -0.5
    const double z = small ? x * x : (1.0 + x) * 0.5;
    const double p = z * (pS0 + z * (pS1 + z * (pS2 + z * (pS3 + z * (pS4 + z * pS5)))));
    const double q = 1.0 + z * (qS1 + z * (qS2 + z * (qS3 + z * qS4)));
    const double r = p / q;
    const double s = sqrt_in_range(small ? 0.25 : z);       // used by the x <= -0.5 range, where z >= 2^-54
    double res = small ? pio2_hi - (x - (pio2_lo - x * r)) : pi - 2.0 * (s + (r * s - pio2_lo));
    if (ix <= 0x3c600000) res = pio2_hi + pio2_lo;          // |x| < 2^-57
    return res;
}

// ------------------------------------------------------ sin/cos kernels ----
// x in [-pi/4, pi/4], y the tail of the reduced argument.
SVM_HD double ksin(double x, double y, int iy) {
    const double S1 = -1.66666666666666324348e-01;
    const double S2 = 8.33333333332248946124e-03;
    const double S3 = -1.98412698298579493134e-04;
    const double S4 = 2.75573137070700676789e-06;
    const double S5 = -2.50507602534068634195e-08;
    const double S6 = 1.58969099521155010221e-10;
    int32_t ix = hiword(x) & 0x7fffffff;
    if (ix < 0x3e400000) {                      // |x| < 2^-27
        if ((int)x == 0) return x;
    }
    double z = x * x;
    double v = z * x;
    double r = S2 + z * (S3 + z * (S4 + z * (S5 + z * S6)));
    if (iy == 0) return x + v * (S1 + z * r);
    return x - ((z * (0.5 * y - v * r) - y) - v * S1);
}

SVM_HD double kcos(double x, double y) {
    const double C1 = 4.16666666666666019037e-02;
    const double C2 = -1.38888888888741095749e-03;
    const double C3 = 2.48015872894767294178e-05;
    const double C4 = -2.75573143513906633035e-07;
    const double C5 = 2.08757232129817482790e-09;
    const double C6 = -1.13596475577881948265e-11;
    int32_t ix = hiword(x) & 0x7fffffff;
    if (ix < 0x3e400000) {                      // |x| < 2^-27
        if ((int)x == 0) return 1.0;
    }
    double z = x * x;
    double r = z * (C1 + z * (C2 + z * (C3 + z * (C4 + z * (C5 + z * C6)))));
    if (ix < 0x3fd33333) return 1.0 - (0.5 * z - (z * r - x * y));
    double qx = (ix > 0x3fe90000) ? 0.28125 : withwords(ix - 0x00200000, 0u);
    double hz = 0.5 * z - qx;
    double a = 1.0 - qx;
    return a - (hz - (z * r - x * y));
}

// Reduce x by n*pi/2 with a three-part split of pi/2 (33+33+53 bits).
// Returns n; y0 + y1 is the reduced argument.  Valid for |x| < 2^19*pi/2.
SVM_HD int rem_pio2(double x, double* y0, double* y1) {
    const double invpio2 = 6.36619772367581382433e-01;
    const double pio2_1 = 1.57079632673412561417e+00;
    const double pio2_1t = 6.07710050650619224932e-11;
    const double pio2_2 = 6.07710050630396597660e-11;
    const double pio2_2t = 2.02226624879595063154e-21;
    const double pio2_3 = 2.02226624871116645580e-21;
    const double pio2_3t = 8.47842766036889956997e-32;
    double t = x < 0 ? -x : x;
    int n = (int)(t * invpio2 + 0.5);
    double fn = (double)n;
    double r = t - fn * pio2_1;
    double w = fn * pio2_1t;
    double y = r - w;
    int32_t j = hiword(t) >> 20;
    int32_t i = j - ((hiword(y) >> 20) & 0x7ff);
    if (i > 16) {                               // second iteration
        double tt = r;
        w = fn * pio2_2;
        r = tt - w;
        w = fn * pio2_2t - ((tt - r) - w);
        y = r - w;
        i = j - ((hiword(y) >> 20) & 0x7ff);
        if (i > 49) {                           // third iteration
            tt = r;
            w = fn * pio2_3;
            r = tt - w;
            w = fn * pio2_3t - ((tt - r) - w);
            y = r - w;
        }
    }
    double yl = (r - y) - w;
    if (x < 0) { *y0 = -y; *y1 = -yl; return -n; }
    *y0 = y; *y1 = yl;
    return n;
}

SVM_HD double sin(double x) {
    int32_t ix = hiword(x) & 0x7fffffff;
    if (ix <= 0x3fe921fb) return ksin(x, 0.0, 0);   // |x| <= pi/4
    if (ix >= 0x7ff00000) return x - x;             // inf or NaN
    double y0, y1;
    int n = rem_pio2(x, &y0, &y1);
    switch (n & 3) {
        case 0: return ksin(y0, y1, 1);
        case 1: return kcos(y0, y1);
        case 2: return -ksin(y0, y1, 1);
        default: return -kcos(y0, y1);
    }
}

SVM_HD double cos(double x) {
    int32_t ix = hiword(x) & 0x7fffffff;
    if (ix <= 0x3fe921fb) return kcos(x, 0.0);
    if (ix >= 0x7ff00000) return x - x;
    double y0, y1;
    int n = rem_pio2(x, &y0, &y1);
    switch (n & 3) {
        case 0: return kcos(y0, y1);
        case 1: return -ksin(y0, y1, 1);
        case 2: return -kcos(y0, y1);
        default: return ksin(y0, y1, 1);
    }
}

// cos(+-pi/2) as cos() above returns it for the double nearest pi/2 (the
// residual 6.123e-17 of that double); sin() returns exactly +-1 there.  A
// quarter turn (the T1 flip) uses these instead of evaluating both functions.
constexpr double kCosQuarterTurn = 0x1.1a62633145c07p-54;

// ----------------------------------------------------------------- log ----
SVM_HD double log(double x) {
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    const double two54 = 1.80143985094819840000e+16;
    const double Lg1 = 6.666666666666735130e-01;
    const double Lg2 = 3.999999999940941908e-01;
    const double Lg3 = 2.857142874366239149e-01;
    const double Lg4 = 2.222219843214978396e-01;
    const double Lg5 = 1.818357216161805012e-01;
    const double Lg6 = 1.531383769920937332e-01;
    const double Lg7 = 1.479819860511658591e-01;
    int32_t hx = hiword(x);
    uint32_t lx = loword(x);
    int32_t k = 0;
    if (hx < 0x00100000) {                      // x < 2^-1022
        if (((hx & 0x7fffffff) | (int32_t)lx) == 0) return -two54 / 0.0;  // log(+-0) = -inf
        if (hx < 0) return (x - x) / 0.0;       // log(-#) = NaN
        k -= 54;
        x *= two54;                             // scale up subnormal
        hx = hiword(x);
    }
    if (hx >= 0x7ff00000) return x + x;
    k += (hx >> 20) - 1023;
    hx &= 0x000fffff;
    int32_t i = (hx + 0x95f64) & 0x100000;
    x = withwords(hx | (i ^ 0x3ff00000), loword(x));   // normalize x or x/2
    k += (i >> 20);
    double f = x - 1.0;
    double dk;
    if ((0x000fffff & (2 + hx)) < 3) {          // |f| < 2^-20
        if (f == 0.0) {
            if (k == 0) return 0.0;
            dk = (double)k;
            return dk * ln2_hi + dk * ln2_lo;
        }
        double R = f * f * (0.5 - 0.33333333333333333 * f);
        if (k == 0) return f - R;
        dk = (double)k;
        return dk * ln2_hi - ((R - dk * ln2_lo) - f);
    }
    double s = f / (2.0 + f);
    dk = (double)k;
    double z = s * s;
    i = hx - 0x6147a;
    double w = z * z;
    int32_t j = 0x6b851 - hx;
    double t1 = w * (Lg2 + w * (Lg4 + w * Lg6));
    double t2 = z * (Lg1 + w * (Lg3 + w * (Lg5 + w * Lg7)));
    i |= j;
    double R = t2 + t1;
    if (i > 0) {
        double hfsq = 0.5 * f * f;
        if (k == 0) return f - (hfsq - s * (hfsq + R));
        return dk * ln2_hi - ((hfsq - (s * (hfsq + R) + dk * ln2_lo)) - f);
    }
    if (k == 0) return f - s * (f - R);
    return dk * ln2_hi - ((s * (f - R) - dk * ln2_lo) - f);
}

// -------------------------------------------------------- Philox4x32-10 ----
struct u32x4 { uint32_t v[4]; };

SVM_HD void mulhilo32(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}

SVM_HD u32x4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                           uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo32(M0, c0, &hi0, &lo0);
        mulhilo32(M1, c2, &hi1, &lo1);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += W0; k1 += W1;
    }
    u32x4 out;
    out.v[0] = c0; out.v[1] = c1; out.v[2] = c2; out.v[3] = c3;
    return out;
}

// 53-bit uniform on [0,1) from two 32-bit words (a gives the top 27 bits).
SVM_HD double u01(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

// Stream identifiers: the fourth Philox counter word.
enum : uint32_t {
    STREAM_CELL_POINTS = 0x50545331u,   // initial cell centres
    STREAM_DIRECTORS   = 0x44495231u,   // initial vertex directors
    STREAM_BROWNIAN    = 0x42524e31u,   // Brownian displacement noise
    STREAM_ACTIVE      = 0x41435431u,   // active director rotation angle
    STREAM_BATCH       = 0x42415431u,   // batch parameter draws
    STREAM_MLP         = 0x4d4c5031u,   // network initialisation / sampling
    STREAM_VICSEK      = 0x56494331u    // Vicsek random tangent direction
};

// Two uniforms on [0,1) for draw `idx` of `stream` at `step` under `seed`.
SVM_HD void uniform2(uint64_t seed, uint32_t stream, uint64_t step, uint32_t idx,
                     double* u0, double* u1) {
    u32x4 r = philox4x32_10(idx, (uint32_t)step, (uint32_t)(step >> 32), stream,
                            (uint32_t)seed, (uint32_t)(seed >> 32));
    *u0 = u01(r.v[0], r.v[1]);
    *u1 = u01(r.v[2], r.v[3]);
}

// Two standard normals by the Box-Muller transform of one uniform pair.
SVM_HD void normal2(uint64_t seed, uint32_t stream, uint64_t step, uint32_t idx,
                    double* z0, double* z1) {
    double a, b;
    uniform2(seed, stream, step, idx, &a, &b);
    double rad = sqrt_rn(-2.0 * svm::log(1.0 - a));
    double th = (2.0 * SVM_PI) * b;
    *z0 = rad * svm::cos(th);
    *z1 = rad * svm::sin(th);
}

// Three standard normals for vertex `v` at `step` (components x, y, z).
SVM_HD void normal3(uint64_t seed, uint64_t step, uint32_t v, double* n) {
    double z0, z1, z2, z3;
    normal2(seed, STREAM_BROWNIAN, step, 2u * v, &z0, &z1);
    normal2(seed, STREAM_BROWNIAN, step, 2u * v + 1u, &z2, &z3);
    n[0] = z0; n[1] = z1; n[2] = z2;
    (void)z3;
}

}  // namespace svm
