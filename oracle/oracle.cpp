// oracle.cpp -- CPU restatement of the reference's CPU path for the spherical
// vertex model.  TEST INFRASTRUCTURE ONLY: it is the checker the HIP library is
// compared against, and is never linked into or called by the product.
//
// Every function follows the reference operation for operation (same
// expression trees, same loop orders, same in-place T1 rewiring), citing the
// reference file:line it restates.  Differences from the reference, all
// deliberate and documented:
//   * acos/sin/cos/log come from svm_fp64.h (own fdlibm-method code, <= 1 ulp
//     from glibc) unless built with -DORACLE_LIBM, so that host and device
//     agree bit for bit.
//   * random draws come from Philox4x32-10 streams (svm_fp64.h) instead of
//     mt19937/curand: the reference's draws are not reproducible anyway
//     (src/utilities/noiseSource.h:21-31).
//   * FIRE reductions can use either the reference's sequential sum or the
//     fixed-shape pairwise tree the device uses (reduce_mode).
// Build: oracle/Makefile (g++ -O3 -ffp-contract=off, no fast-math).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>
#include <algorithm>

#include "../curvedspacevertexmodel_amd/csrc/svm_fp64.h"

#ifdef ORACLE_LIBM
#define O_ACOS std::acos
#define O_SIN std::sin
#define O_COS std::cos
#else
#define O_ACOS svm::acos
#define O_SIN svm::sin
#define O_COS svm::cos
#endif

namespace {

const double PI = SVM_PI;  // inc/std_include.h:47

// dVec restated: inc/dDimensionalVectorTypes.h:19-135 (component loops).
struct dVec {
    double x[3];
    dVec() {}
    explicit dVec(double v) { x[0] = v; x[1] = v; x[2] = v; }
    double& operator[](int i) { return x[i]; }
    double operator[](int i) const { return x[i]; }
    dVec& operator-=(const dVec& o) { for (int d = 0; d < 3; ++d) x[d] -= o.x[d]; return *this; }
    dVec& operator+=(const dVec& o) { for (int d = 0; d < 3; ++d) x[d] += o.x[d]; return *this; }
};
inline dVec operator+(const dVec& a, const dVec& b) { dVec r; for (int d = 0; d < 3; ++d) r.x[d] = a.x[d] + b.x[d]; return r; }
inline dVec operator-(const dVec& a, const dVec& b) { dVec r; for (int d = 0; d < 3; ++d) r.x[d] = a.x[d] - b.x[d]; return r; }
inline dVec operator*(double a, const dVec& b) { dVec r; for (int d = 0; d < 3; ++d) r.x[d] = a * b.x[d]; return r; }
inline dVec operator*(const dVec& b, double a) { dVec r; for (int d = 0; d < 3; ++d) r.x[d] = a * b.x[d]; return r; }
inline bool operator==(const dVec& a, const dVec& b) { for (int d = 0; d < 3; ++d) if (a.x[d] != b.x[d]) return false; return true; }

// inc/functions.h:131-181
inline double dot(const dVec& a, const dVec& b) { double s = 0.0; for (int d = 0; d < 3; ++d) s += a.x[d] * b.x[d]; return s; }
inline dVec cross(const dVec& a, const dVec& b) {
    dVec r;
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
    return r;
}
inline double norm(const dVec& p) { return std::sqrt(dot(p, p)); }
inline void rodriguesRotation(dVec& v, dVec& k, double theta) {
    dVec cp = cross(k, v);
    dVec ans = O_COS(theta) * v + O_SIN(theta) * cp + (1.0 - O_COS(theta)) * dot(k, v) * k;
    v = ans;
}

// sphericalDomain restated: inc/sphericalDomain.h:13-420 (only the members
// on the vertex-model path).
struct Sphere {
    double radius = 1.0, inverseRadius = 1.0;
    explicit Sphere(double r = 1.0) { radius = r; inverseRadius = 1.0 / r; }

    void putInBoxVirtual(dVec& p) const { p = p * (1.0 / norm(p)); }            // :179-182
    void putInBoxReal(dVec& p) const { dVec pt = p * (1.0 / norm(p)) * radius; p = pt; }  // :183-187
    void projectToTangentPlane(dVec& v, const dVec& n) const {                    // :169-172
        v = v - (dot(v, n) / dot(n, n)) * n;
    }
    void projectToTangentPlaneAndNormalize(dVec& v, const dVec& n) const {        // :173-177
        v = v - (dot(v, n) / dot(n, n)) * n;
        v = v * (1.0 / norm(v));
    }
    void move(dVec& p1, const dVec& velocityDirection) const {                   // :203-209
        dVec disp = velocityDirection;
        projectToTangentPlane(disp, p1);
        p1 = p1 + disp;
        putInBoxReal(p1);
    }
    void geodesicDistance(const dVec& p1, const dVec& p2, double& dist) const {  // :94-112
        double numerator = p1[0] * p2[0] + p1[1] * p2[1] + p1[2] * p2[2];
        double denominator = std::sqrt(p1[0] * p1[0] + p1[1] * p1[1] + p1[2] * p1[2]) *
                             std::sqrt(p2[0] * p2[0] + p2[1] * p2[1] + p2[2] * p2[2]);
        dist = radius * O_ACOS(numerator / denominator);
    }
    void includedAngle(const dVec& p1, const dVec& p2, const dVec& p3, double& angle) const {  // :114-133
        dVec pt1 = p1, pt2 = p2, pt3 = p3;
        dVec crossI = cross(pt2, pt3);
        dVec crossIm1 = cross(pt1, pt2);
        double determinant = pt1[0] * (pt2[1] * pt3[2] - pt2[2] * pt3[1])
                           + pt1[1] * (pt2[2] * pt3[0] - pt2[0] * pt3[2])
                           + pt1[2] * (pt2[0] * pt3[1] - pt2[1] * pt3[0]);
        angle = O_ACOS(-dot(crossI, crossIm1) / (norm(crossI) * norm(crossIm1)));
        if (determinant > 0) angle *= -1;
        if (angle < 0) angle += 2. * PI;
    }
    void sphericalTriangleArea(const dVec& p1, const dVec& p2, const dVec& p3, double& area) const {  // :135-166
        dVec pt1 = p1, pt2 = p2, pt3 = p3;
        putInBoxVirtual(pt1);
        putInBoxVirtual(pt2);
        putInBoxVirtual(pt3);
        double p1Dotp2 = dot(pt1, pt2);
        double p1Dotp3 = dot(pt1, pt3);
        double p2Dotp3 = dot(pt2, pt3);
        area = -PI;
        area += O_ACOS((p2Dotp3 - p1Dotp2 * p1Dotp3) / (std::sqrt(1 - p1Dotp2 * p1Dotp2) * std::sqrt(1 - p1Dotp3 * p1Dotp3)));
        area += O_ACOS((p1Dotp2 - p1Dotp3 * p2Dotp3) / (std::sqrt(1 - p1Dotp3 * p1Dotp3) * std::sqrt(1 - p2Dotp3 * p2Dotp3)));
        area += O_ACOS((p1Dotp3 - p1Dotp2 * p2Dotp3) / (std::sqrt(1 - p1Dotp2 * p1Dotp2) * std::sqrt(1 - p2Dotp3 * p2Dotp3)));
        area *= (radius * radius);
    }
    void cartesianSphericalBasisChange(const dVec& c, dVec& thetaHat, dVec& phiHat) const {  // :218-232
        double r = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        double cosT = c[2] / r;
        double sinT = std::sqrt(1. - cosT * cosT);
        double denom = std::sqrt(c[0] * c[0] + c[1] * c[1]);
        double cosP = c[0] / denom;
        double sinP = c[1] / denom;
        thetaHat[0] = cosT * cosP;
        thetaHat[1] = cosT * sinP;
        thetaHat[2] = -sinT;
        phiHat[0] = -sinP;
        phiHat[1] = cosP;
        phiHat[2] = 0;
    }
    void gradientGeodesicDistance(const dVec& v1, const dVec& v2, dVec& derivative,
                                  const dVec& thetaHat, const dVec& phiHat) const {  // :248-273
        double r1 = std::sqrt(dot(v1, v1));
        double r2 = std::sqrt(dot(v2, v2));
        double cosT1 = v1[2] / r1;
        double cosT2 = v2[2] / r2;
        double sinT1 = std::sqrt(1 - cosT1 * cosT1);
        double sinT2 = std::sqrt(1 - cosT2 * cosT2);
        double cosP1 = v1[0] / std::sqrt(v1[0] * v1[0] + v1[1] * v1[1]);
        double cosP2 = v2[0] / std::sqrt(v2[0] * v2[0] + v2[1] * v2[1]);
        double sinP1 = v1[1] / std::sqrt(v1[0] * v1[0] + v1[1] * v1[1]);
        double sinP2 = v2[1] / std::sqrt(v2[0] * v2[0] + v2[1] * v2[1]);
        double sinP1MinusP2 = sinP1 * cosP2 - cosP1 * sinP2;
        double cosP1MinusP2 = cosP1 * cosP2 + sinP1 * sinP2;
        double denomPart = cosT1 * cosT2 + cosP1MinusP2 * sinT1 * sinT2;
        double denomInverse = 1. / std::sqrt(1 - denomPart * denomPart);
        double gradTheta = 1.0 * (cosT2 * sinT1 - cosT1 * cosP1MinusP2 * sinT2) * denomInverse;
        double gradPhi = sinT2 * sinP1MinusP2 * denomInverse;
        derivative = gradTheta * thetaHat + gradPhi * phiHat;
    }
    void gradientTriangleArea(const dVec& v1, const dVec& v2, const dVec& v3, dVec& derivative,
                              const dVec& thetaHat, const dVec& phiHat) const {  // :347-411
        double r1 = std::sqrt(dot(v1, v1));
        double r2 = std::sqrt(dot(v2, v2));
        double r3 = std::sqrt(dot(v3, v3));
        double cosT1 = v1[2] / r1;
        double cosT2 = v2[2] / r2;
        double cosT3 = v3[2] / r3;
        double sinT1 = std::sqrt(1 - cosT1 * cosT1);
        double sinT2 = std::sqrt(1 - cosT2 * cosT2);
        double sinT3 = std::sqrt(1 - cosT3 * cosT3);
        double cosP1 = v1[0] / std::sqrt(v1[0] * v1[0] + v1[1] * v1[1]);
        double cosP2 = v2[0] / std::sqrt(v2[0] * v2[0] + v2[1] * v2[1]);
        double cosP3 = v3[0] / std::sqrt(v3[0] * v3[0] + v3[1] * v3[1]);
        double sinP1 = v1[1] / std::sqrt(v1[0] * v1[0] + v1[1] * v1[1]);
        double sinP2 = v2[1] / std::sqrt(v2[0] * v2[0] + v2[1] * v2[1]);
        double sinP3 = v3[1] / std::sqrt(v3[0] * v3[0] + v3[1] * v3[1]);
        double sinP1MinusP2 = sinP1 * cosP2 - cosP1 * sinP2;
        double sinP1MinusP3 = sinP1 * cosP3 - cosP1 * sinP3;
        double cosP1MinusP2 = cosP1 * cosP2 + sinP1 * sinP2;
        double cosP1MinusP3 = cosP1 * cosP3 + sinP1 * sinP3;
        double cosP2MinusP3 = cosP2 * cosP3 + sinP2 * sinP3;
        double s12, s13, s23, d12, d13, denom1, denom2, denom3;
        s12 = cosT1 * cosT2 + cosP1MinusP2 * sinT1 * sinT2;
        s13 = cosT1 * cosT3 + cosP1MinusP3 * sinT1 * sinT3;
        s23 = cosT2 * cosT3 + cosP2MinusP3 * sinT2 * sinT3;
        d12 = cosT1 * cosP1MinusP2 * sinT2 - cosT2 * sinT1;
        d13 = cosT1 * cosP1MinusP3 * sinT3 - cosT3 * sinT1;
        denom1 = std::sqrt((-1. + s12 * s12) * (-1. + s12 * s12) * (1 - s12 * s12 - s13 * s13 - s23 * s23 + 2.0 * s12 * s13 * s23));
        denom2 = std::sqrt((-1. + s13 * s13) * (-1. + s13 * s13) * (1 - s12 * s12 - s13 * s13 - s23 * s23 + 2.0 * s12 * s13 * s23));
        denom3 = std::sqrt((s12 * s12 - 1.) * (s13 * s13 - 1.) * (s12 * s12 - 1.) * (s13 * s13 - 1.) * (1 - s12 * s12 - s13 * s13 - s23 * s23 + 2.0 * s12 * s13 * s23));
        double gradTheta = (d13 * (s12 * s12 - 1.0) + d12 * (s23 - s12 * s13)) / denom1
                         + (d12 * (s13 * s13 - 1.0) + d13 * (s23 - s12 * s13)) / denom2
                         - (d12 * (s13 * s13 - 1.) * (s13 - s12 * s23) + d13 * (s12 * s12 - 1.0) * (s12 - s13 * s23)) / denom3;
        gradTheta *= radius;
        double gradPhi = ((s12 * s13 - s23) * sinP1MinusP2 * sinT2 - (-1 + s12 * s12) * sinP1MinusP3 * sinT3) / (denom1)
                       - ((-1 + s13 * s13) * sinP1MinusP2 * sinT2 + (-(s12 * s13) + s23) * sinP1MinusP3 * sinT3) / (denom2)
                       + ((-1 + s13 * s13) * (s13 - s12 * s23) * sinP1MinusP2 * sinT2 + (-1 + s12 * s12) * (s12 - s13 * s23) *
                          sinP1MinusP3 * sinT3) / (denom3);
        gradPhi *= radius;
        double determinant = v1[0] * (v2[1] * v3[2] - v2[2] * v3[1])
                           + v1[1] * (v2[2] * v3[0] - v2[0] * v3[2])
                           + v1[2] * (v2[0] * v3[1] - v2[1] * v3[0]);
        derivative = gradTheta * thetaHat + gradPhi * phiHat;
        if (determinant < 0) derivative = -1. * derivative;
    }
};

// Fixed-shape pairwise reduction shared with the device: 256-element chunks
// folded by halving strides (128..1), then the chunk partials folded the same
// way over the next power of two.  Missing elements are 0.0.
double tree_sum(const std::vector<double>& x) {
    size_t n = x.size();
    size_t nc = (n + 255) / 256;
    if (nc == 0) return 0.0;
    size_t m = 1;
    while (m < nc) m <<= 1;
    std::vector<double> part(m, 0.0);
    double s[256];
    for (size_t c = 0; c < nc; ++c) {
        for (int i = 0; i < 256; ++i) {
            size_t k = c * 256 + i;
            s[i] = k < n ? x[k] : 0.0;
        }
        for (int h = 128; h >= 1; h >>= 1)
            for (int i = 0; i < h; ++i) s[i] = s[i] + s[i + h];
        part[c] = s[0];
    }
    for (size_t h = m / 2; h >= 1; h >>= 1)
        for (size_t i = 0; i < h; ++i) part[i] = part[i] + part[i + h];
    return part[0];
}

// The model: state arrays in the reference's layouts (AoS dVec; Index2D
// rings ring[c*maxDeg + k], inc/indexer.h:41-44; v-v and v-c lists stride 3).
struct Model {
    int nCells = 0, N = 0, maxDeg = 0;
    Sphere sphere;
    double Kr = 1.0, t1Threshold = 0.1;
    std::vector<dVec> positions, velocities, forces, directors, cellPositions;
    std::vector<double> masses;
    std::vector<int> neighbors, vertexCellNeighbors, cellNeighbors;
    std::vector<unsigned> cellNumberOfNeighbors, numberOfNeighbors;
    std::vector<double> ap, app;                        // (area, perimeter), (a0, p0)
    std::vector<dVec> curVert, lastVert, nextVert;      // per (vertex, slot) staging
    int nT1 = 0;                                        // flips performed (diagnostic)
    int nFanRecompute = 0;                              // fan fallbacks taken
    int nNan = 0;                                       // NaN force contributions skipped
    int nT1Skipped = 0;                                 // flips skipped: no cell l
    int reduce_mode = 0;                                // 0: sequential sums, 1: tree_sum

    int cni(int k, int c) const { return c * maxDeg + k; }

    // sphericalVertexModel::recomputeAreasCPU, src/model/sphericalVertexModel.cpp:162-207
    void recomputeAreasCPU() {
        for (int cc = 0; cc < nCells; ++cc) {
            int neighs = cellNumberOfNeighbors[cc];
            dVec cellPos = cellPositions[cc];
            int lastVertexIdx = cellNeighbors[cni(neighs - 2, cc)];
            int curVertexIdx = cellNeighbors[cni(neighs - 1, cc)];
            dVec lastVertexPos = positions[lastVertexIdx];
            dVec curVertexPos = positions[curVertexIdx];
            double area = 0, tempVal;
            for (int nn = 0; nn < neighs; ++nn) {
                int nextVertexIdx = cellNeighbors[cni(nn, cc)];
                dVec nextVertexPos = positions[nextVertexIdx];
                sphere.sphericalTriangleArea(cellPos, lastVertexPos, curVertexPos, tempVal);
                area += tempVal;
                lastVertexPos = curVertexPos;
                curVertexIdx = nextVertexIdx;
                curVertexPos = nextVertexPos;
            }
            ap[2 * cc] = area;
        }
    }

    // sphericalVertexModel::computeGeometryCPU, src/model/sphericalVertexModel.cpp:209-311
    void computeGeometry() {
        double totalArea = 0;
        std::vector<double> areas(reduce_mode ? nCells : 0);
        for (int cc = 0; cc < nCells; ++cc) {
            int neighs = cellNumberOfNeighbors[cc];
            dVec cellPos(0.0);
            for (int nn = 0; nn < neighs; ++nn) cellPos = cellPos + positions[cellNeighbors[cni(nn, cc)]];
            sphere.putInBoxReal(cellPos);
            cellPositions[cc] = cellPos;
            int lastVertexIdx = cellNeighbors[cni(neighs - 3, cc)];
            int curVertexIdx = cellNeighbors[cni(neighs - 2, cc)];
            int nextVertexIdx = cellNeighbors[cni(neighs - 1, cc)];
            (void)lastVertexIdx;
            dVec lastVertexPos = positions[lastVertexIdx];
            dVec curVertexPos = positions[curVertexIdx];
            dVec nextVertexPos = positions[nextVertexIdx];
            double perimeter = 0, area = 0, tempVal;
            for (int nn = 0; nn < neighs; ++nn) {
                int vNeighs = numberOfNeighbors[curVertexIdx];
                int forceSetIdx = -1;
                for (int vn = 0; vn < vNeighs; ++vn) {
                    int newIdx = 3 * curVertexIdx + vn;
                    if (vertexCellNeighbors[newIdx] == cc) forceSetIdx = newIdx;
                }
                int vIdxPlus2 = cellNeighbors[cni(nn, cc)];
                dVec positive2VertexPos = positions[vIdxPlus2];
                sphere.geodesicDistance(lastVertexPos, curVertexPos, tempVal);
                perimeter += tempVal;
                sphere.includedAngle(lastVertexPos, curVertexPos, nextVertexPos, tempVal);
                area += tempVal;
                curVert[forceSetIdx] = curVertexPos;
                lastVert[forceSetIdx] = lastVertexPos;
                nextVert[forceSetIdx] = nextVertexPos;
                lastVertexPos = curVertexPos;
                curVertexIdx = nextVertexIdx;
                curVertexPos = nextVertexPos;
                nextVertexIdx = vIdxPlus2;
                nextVertexPos = positive2VertexPos;
            }
            area = (area - (neighs - 2) * PI);
            int extraAngularArea = (int)std::floor(area / (1.0 * PI));
            if (extraAngularArea > 0) area -= extraAngularArea * PI;
            area = area * sphere.radius * sphere.radius;
            ap[2 * cc] = area;
            ap[2 * cc + 1] = perimeter;
            totalArea += area;
            if (reduce_mode) areas[cc] = area;
        }
        if (reduce_mode) totalArea = tree_sum(areas);
        double excessArea = totalArea - 4.0 * PI * sphere.radius * sphere.radius;
        if (std::fabs(excessArea) > 1e-6) {
            ++nFanRecompute;
            recomputeAreasCPU();
        }
    }

    // sphericalVertexModel::computeForceCPU, src/model/sphericalVertexModel.cpp:356-420
    void computeForces() {
        computeGeometry();
        for (int vertexIndex = 0; vertexIndex < N; ++vertexIndex) {
            dVec f(0.0);
            int vNeighs = numberOfNeighbors[vertexIndex];
            for (int cc = 0; cc < vNeighs; ++cc) {
                dVec fSet(0.0), tempVar;
                int c = 3 * vertexIndex + cc;
                int cellIndex = vertexCellNeighbors[c];
                dVec cPos = cellPositions[cellIndex];
                dVec vLast = lastVert[c], vCur = curVert[c], vNext = nextVert[c];
                double areaDifference = ap[2 * cellIndex] - app[2 * cellIndex];
                double perimeterDifference = ap[2 * cellIndex + 1] - app[2 * cellIndex + 1];
                dVec thetaHat, phiHat;
                sphere.cartesianSphericalBasisChange(vCur, thetaHat, phiHat);
                sphere.gradientGeodesicDistance(vCur, vLast, tempVar, thetaHat, phiHat);
                fSet -= 2.0 * perimeterDifference * tempVar;
                sphere.gradientGeodesicDistance(vCur, vNext, tempVar, thetaHat, phiHat);
                fSet -= 2.0 * perimeterDifference * tempVar;
                sphere.gradientTriangleArea(vCur, vLast, cPos, tempVar, thetaHat, phiHat);
                fSet -= 2.0 * Kr * areaDifference * tempVar;
                sphere.gradientTriangleArea(vCur, cPos, vNext, tempVar, thetaHat, phiHat);
                fSet -= 2.0 * Kr * areaDifference * tempVar;
                if (!std::isnan(fSet[0])) f += fSet;
                else ++nNan;
            }
            forces[vertexIndex] = f;
        }
    }

    // sphericalVertexModel::computeEnergy, src/model/sphericalVertexModel.cpp:150-160
    double computeEnergy(int reduce_mode) const {
        if (reduce_mode == 0) {
            double energy = 0.0;
            for (int i = 0; i < nCells; ++i)
                energy += Kr * (ap[2 * i] - app[2 * i]) * (ap[2 * i] - app[2 * i]) +
                          (ap[2 * i + 1] - app[2 * i + 1]) * (ap[2 * i + 1] - app[2 * i + 1]);
            return energy;
        }
        std::vector<double> e(nCells);
        for (int i = 0; i < nCells; ++i)
            e[i] = Kr * (ap[2 * i] - app[2 * i]) * (ap[2 * i] - app[2 * i]) +
                   (ap[2 * i + 1] - app[2 * i + 1]) * (ap[2 * i + 1] - app[2 * i + 1]);
        return tree_sum(e);
    }

    // simpleModel::computeKineticEnergy, src/model/simpleModel.cpp:63-73
    double computeKineticEnergy(int reduce_mode) const {
        if (reduce_mode == 0) {
            double en = 0.0;
            for (int ii = 0; ii < N; ++ii) en += 0.5 * masses[ii] * dot(velocities[ii], velocities[ii]);
            return en;
        }
        std::vector<double> e(N);
        for (int ii = 0; ii < N; ++ii) e[ii] = 0.5 * masses[ii] * dot(velocities[ii], velocities[ii]);
        return tree_sum(e);
    }

    // sphericalVertexModel::growCellVerticesList, src/model/sphericalVertexModel.cpp:733-758
    void growCellVerticesList(int newVertexMax) {
        int oldMax = maxDeg;
        int newMax = newVertexMax + 1;
        std::vector<int> nc(newMax * nCells, 0);
        for (int cell = 0; cell < nCells; ++cell)
            for (unsigned n = 0; n < cellNumberOfNeighbors[cell]; ++n)
                nc[cell * newMax + n] = cellNeighbors[cell * oldMax + n];
        maxDeg = newMax;
        cellNeighbors.swap(nc);
    }

    // sphericalVertexModel::getCellVertexSetForT1, src/model/sphericalVertexModel.cpp:611-731.
    // Returns false when the four cells cannot all be determined: vertex1 and
    // vertex2 then share all three cells (a cell touching both ends of the
    // edge), there is no cell l, and the reference would read its int4
    // cellSet uninitialised (left over from an earlier flip); such a flip is
    // skipped here, as the triangle rule skips a flip.
    bool getCellVertexSetForT1(int vertex1, int vertex2, int cs[4], int vs[4], bool& growList) {
        cs[0] = cs[1] = cs[2] = cs[3] = -1;
        int cell1 = vertexCellNeighbors[3 * vertex1];
        int cell2 = vertexCellNeighbors[3 * vertex1 + 1];
        int cell3 = vertexCellNeighbors[3 * vertex1 + 2];
        for (int ff = 0; ff < 3; ++ff) {
            int ctest = vertexCellNeighbors[3 * vertex2 + ff];
            if (ctest != cell1 && ctest != cell2 && ctest != cell3) cs[3] = ctest;
        }
        // (the reference's "find c and d" walk at :631-640 has no effect)
        int vlast, vcur, vnext = -1, cneigh;
        int cells[3] = {cell1, cell2, cell3};
        for (int t = 0; t < 3; ++t) {
            int cell = cells[t];
            cneigh = cellNumberOfNeighbors[cell];
            vlast = cellNeighbors[cni(cneigh - 2, cell)];
            vcur = cellNeighbors[cni(cneigh - 1, cell)];
            for (int cn = 0; cn < cneigh; ++cn) {
                vnext = cellNeighbors[cni(cn, cell)];
                if (vcur == vertex1) break;
                vlast = vcur;
                vcur = vnext;
            }
            if (vlast == vertex2) cs[0] = cell;
            else if (vnext == vertex2) cs[2] = cell;
            else cs[1] = cell;
        }
        if (cs[0] < 0 || cs[1] < 0 || cs[2] < 0 || cs[3] < 0) return false;
        cneigh = cellNumberOfNeighbors[cs[1]];
        vlast = cellNeighbors[cni(cneigh - 2, cs[1])];
        vcur = cellNeighbors[cni(cneigh - 1, cs[1])];
        for (int cn = 0; cn < cneigh; ++cn) {
            vnext = cellNeighbors[cni(cn, cs[1])];
            if (vcur == vertex1) break;
            vlast = vcur;
            vcur = vnext;
        }
        vs[0] = vlast;
        vs[1] = vnext;
        cneigh = cellNumberOfNeighbors[cs[3]];
        vlast = cellNeighbors[cni(cneigh - 2, cs[3])];
        vcur = cellNeighbors[cni(cneigh - 1, cs[3])];
        for (int cn = 0; cn < cneigh; ++cn) {
            vnext = cellNeighbors[cni(cn, cs[3])];
            if (vcur == vertex2) break;
            vlast = vcur;
            vcur = vnext;
        }
        vs[3] = vlast;
        vs[2] = vnext;
        if ((int)cellNumberOfNeighbors[cs[0]] == maxDeg || (int)cellNumberOfNeighbors[cs[1]] == maxDeg ||
            (int)cellNumberOfNeighbors[cs[2]] == maxDeg || (int)cellNumberOfNeighbors[cs[3]] == maxDeg)
            growList = true;
        return true;
    }

    // sphericalVertexModel::testAndPerformT1TransitionsCPU, src/model/sphericalVertexModel.cpp:439-604
    void testAndPerformT1Transitions() {
        int vMax = maxDeg;
        for (int vertex1 = 0; vertex1 < N; ++vertex1) {
            dVec v1 = positions[vertex1];
            for (int vv = 0; vv < 3; ++vv) {
                int vertex2 = neighbors[3 * vertex1 + vv];
                if (vertex1 < vertex2) {
                    dVec v2 = positions[vertex2];
                    double arcLength;
                    sphere.geodesicDistance(v1, v2, arcLength);
                    if (arcLength < t1Threshold) {
                        bool grow = false;
                        int cs[4], vs[4] = {0, 0, 0, 0};
                        if (!getCellVertexSetForT1(vertex1, vertex2, cs, vs, grow)) {
                            ++nT1Skipped;
                            continue;
                        }
                        if (cellNumberOfNeighbors[cs[0]] == 3 || cellNumberOfNeighbors[cs[2]] == 3) continue;
                        if (grow) {
                            vMax += 1;
                            growCellVerticesList(vMax);
                        }
                        dVec midpoint = 0.5 * (v1 + v2);
                        sphere.putInBoxVirtual(midpoint);
                        dVec vC = positions[vs[2]];
                        double determinant = vC[0] * (v1[1] * v2[2] - v1[2] * v2[1])
                                           + vC[1] * (v1[2] * v2[0] - v1[0] * v2[2])
                                           + vC[2] * (v1[0] * v2[1] - v1[1] * v2[0]);
                        determinant = determinant > 0 ? 1. : -1.;
                        rodriguesRotation(v1, midpoint, -0.5 * determinant * PI);
                        rodriguesRotation(v2, midpoint, -0.5 * determinant * PI);
                        dVec diff = 0.5 * (v1 - v2);
                        v1 = v1 + diff;
                        v2 = v2 - diff;
                        sphere.putInBoxReal(v1);
                        sphere.putInBoxReal(v2);
                        positions[vertex1] = v1;
                        positions[vertex2] = v2;
                        for (int vert = 0; vert < 3; ++vert) {
                            if (vertexCellNeighbors[3 * vertex1 + vert] == cs[2]) vertexCellNeighbors[3 * vertex1 + vert] = cs[3];
                            if (vertexCellNeighbors[3 * vertex2 + vert] == cs[0]) vertexCellNeighbors[3 * vertex2 + vert] = cs[1];
                            if (neighbors[3 * vs[1] + vert] == vertex1) neighbors[3 * vs[1] + vert] = vertex2;
                            if (neighbors[3 * vs[2] + vert] == vertex2) neighbors[3 * vs[2] + vert] = vertex1;
                            if (neighbors[3 * vertex1 + vert] == vs[1]) neighbors[3 * vertex1 + vert] = vs[2];
                            if (neighbors[3 * vertex2 + vert] == vs[2]) neighbors[3 * vertex2 + vert] = vs[1];
                        }
                        // cell i loses vertex2
                        int cneigh = cellNumberOfNeighbors[cs[0]];
                        int cidx = 0;
                        for (int cc = 0; cc < cneigh - 1; ++cc) {
                            if (cellNeighbors[cni(cc, cs[0])] == vertex2) cidx += 1;
                            cellNeighbors[cni(cc, cs[0])] = cellNeighbors[cni(cidx, cs[0])];
                            cidx += 1;
                        }
                        cellNumberOfNeighbors[cs[0]] -= 1;
                        // cell j gains vertex2 after vertex1
                        cneigh = cellNumberOfNeighbors[cs[1]];
                        std::vector<int> cvcopy1(cneigh + 1);
                        cidx = 0;
                        for (int cc = 0; cc < cneigh; ++cc) {
                            int cellIndex = cellNeighbors[cni(cc, cs[1])];
                            cvcopy1[cidx++] = cellIndex;
                            if (cellIndex == vertex1) cvcopy1[cidx++] = vertex2;
                        }
                        for (int cc = 0; cc < cneigh + 1; ++cc) cellNeighbors[cni(cc, cs[1])] = cvcopy1[cc];
                        cellNumberOfNeighbors[cs[1]] += 1;
                        // cell k loses vertex1
                        cneigh = cellNumberOfNeighbors[cs[2]];
                        cidx = 0;
                        for (int cc = 0; cc < cneigh - 1; ++cc) {
                            if (cellNeighbors[cni(cc, cs[2])] == vertex1) cidx += 1;
                            cellNeighbors[cni(cc, cs[2])] = cellNeighbors[cni(cidx, cs[2])];
                            cidx += 1;
                        }
                        cellNumberOfNeighbors[cs[2]] -= 1;
                        // cell l gains vertex1 after vertex2
                        cneigh = cellNumberOfNeighbors[cs[3]];
                        std::vector<int> cvcopy2(cneigh + 1);
                        cidx = 0;
                        for (int cc = 0; cc < cneigh; ++cc) {
                            int cellIndex = cellNeighbors[cni(cc, cs[3])];
                            cvcopy2[cidx++] = cellIndex;
                            if (cellIndex == vertex2) cvcopy2[cidx++] = vertex1;
                        }
                        for (int cc = 0; cc < cneigh + 1; ++cc) cellNeighbors[cni(cc, cs[3])] = cvcopy2[cc];
                        cellNumberOfNeighbors[cs[3]] = cneigh + 1;
                        ++nT1;
                    }
                }
            }
        }
    }

    // sphericalVertexModel::moveParticlesCPU, src/model/sphericalVertexModel.cpp:111-135
    void moveParticles(const std::vector<dVec>& disp, double scale) {
        if (scale == 1.) {
            for (int ii = 0; ii < N; ++ii) sphere.move(positions[ii], disp[ii]);
        } else {
            for (int ii = 0; ii < N; ++ii) sphere.move(positions[ii], scale * disp[ii]);
        }
        testAndPerformT1Transitions();
    }
};

}  // namespace

// ------------------------------------------------------------- C API ------
extern "C" {

// sphericalModel::setParticlePositionsRandomly (src/model/sphericalModel.cpp:37-51)
// followed by setRadius(sqrt(N/4pi)) (src/model/sphericalModel.cpp:11-22,
// src/model/sphericalVertexModel.cpp:12).  Draw i is Philox stream
// CELL_POINTS, counter i.
void or_generate_cell_points(int n, uint64_t seed, double* out) {
    Sphere unit(1.0);
    Sphere big(std::sqrt(n / (4.0 * PI)));
    for (int ii = 0; ii < n; ++ii) {
        double u, v;
        svm::uniform2(seed, svm::STREAM_CELL_POINTS, 0, (uint32_t)ii, &u, &v);
        double phi = 2.0 * PI * u;
        double theta = O_ACOS(2.0 * v - 1);
        dVec p;
        p[0] = 1.0 * O_SIN(theta) * O_COS(phi);
        p[1] = 1.0 * O_SIN(theta) * O_SIN(phi);
        p[2] = 1.0 * O_COS(theta);
        unit.putInBoxReal(p);
        big.putInBoxReal(p);
        for (int d = 0; d < 3; ++d) out[3 * ii + d] = p[d];
    }
}

// Director initialisation of the vertex-model constructor
// (src/model/sphericalVertexModel.cpp:58-73).  Draw ii is Philox stream
// DIRECTORS, counter ii.
void or_generate_directors(int nv, const double* pos, uint64_t seed, double* out) {
    Sphere sph(1.0);
    for (int ii = 0; ii < nv; ++ii) {
        double u2, v2;
        svm::uniform2(seed, svm::STREAM_DIRECTORS, 0, (uint32_t)ii, &u2, &v2);
        double phi = 2.0 * PI * u2;
        double theta = O_ACOS(2.0 * v2 - 1);
        dVec n, p;
        n[0] = 1.0 * O_SIN(theta) * O_COS(phi);
        n[1] = 1.0 * O_SIN(theta) * O_SIN(phi);
        n[2] = 1.0 * O_COS(theta);
        for (int d = 0; d < 3; ++d) p[d] = pos[3 * ii + d];
        sph.projectToTangentPlaneAndNormalize(n, p);
        for (int d = 0; d < 3; ++d) out[3 * ii + d] = n[d];
    }
}

// neighborList::computeCPU (src/utilities/neighborList.cpp:102-166) without
// the cell list: every pair, the difference by
// periodicBoundaryConditions::minDist in a cube of side 2 * box_half
// (inc/periodicBoundaryConditions.h:124-133: whole box lengths added or
// removed until each component is within +-box_half), its norm
// sqrt(dot(d, d)) (inc/functions.h:131-138, 170-173), kept when
// !(dist >= range) (neighborList.cpp:139).  Neighbours in increasing index
// order.  counts[n]; with M >= every count, idx[n*M], vec[3*n*M], dist[n*M]
// (unused entries zero).  Returns the largest count.
int or_neighbor_list(int n, const double* pts, double range, double box_half, int M, int* counts, int* idx,
                     double* vec, double* dist) {
    const double box = 2.0 * box_half;
    int maxc = 0;
    for (int i = 0; i < n; ++i) {
        int c = 0;
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            double d[3];
            for (int k = 0; k < 3; ++k) {
                d[k] = pts[3 * i + k] - pts[3 * j + k];
                while (d[k] < -box_half) d[k] += box;
                while (d[k] > box_half) d[k] -= box;
            }
            double dd = 0.0;
            for (int k = 0; k < 3; ++k) dd += d[k] * d[k];
            const double r = std::sqrt(dd);
            if (r >= range) continue;
            if (c < M) {
                const size_t e = (size_t)i * M + c;
                idx[e] = j;
                for (int k = 0; k < 3; ++k) vec[3 * e + k] = d[k];
                dist[e] = r;
            }
            ++c;
        }
        counts[i] = c;
        maxc = std::max(maxc, c);
    }
    return maxc;
}

// Topology of the vertex model from a spherical Delaunay triangulation,
// following convexHullCGALInterface::sphericalConvexHullForVertexModel
// (src/model/convexHullCGAL.cpp:82-186) and getFacetCenter (:67-80).
// tris[3t..3t+2] is triangle t, counter-clockwise seen from outside; its
// three entries play the role of the facet circulator (sum order for the
// centroid, slot order for the v-c and v-v lists: slot k holds cell tris[k]
// and the triangle across the edge (tris[k-1], tris[k])).  Cell rings are the
// triangles around each point, clockwise seen from outside (the CGAL
// halfedge-around-vertex order), starting at the lowest triangle index.
// Outputs: vpos[3V], vn[3V], vcn[3V], ring[n*maxDeg], cnn[n]; returns maxDeg.
int or_topology_from_triangles(int n, const double* points, int ntri, const int* tris,
                               double* vpos, int* vn, int* vcn, int* ring, int* cnn, int ring_cap) {
    double radius = std::sqrt(points[0] * points[0] + points[1] * points[1] + points[2] * points[2]);
    // edge -> triangle map: key (a,b) directed edge a->b inside triangle t
    std::vector<std::vector<std::pair<int, int>>> out_edges(n);  // a -> (b, t)
    for (int t = 0; t < ntri; ++t)
        for (int k = 0; k < 3; ++k) {
            int a = tris[3 * t + k], b = tris[3 * t + (k + 1) % 3];
            out_edges[a].push_back({b, t});
        }
    auto tri_of_edge = [&](int a, int b) -> int {
        for (auto& e : out_edges[a]) if (e.first == b) return e.second;
        return -1;
    };
    for (int t = 0; t < ntri; ++t) {
        dVec center(0.0);
        for (int k = 0; k < 3; ++k) {
            int c = tris[3 * t + k];
            center[0] += points[3 * c + 0];
            center[1] += points[3 * c + 1];
            center[2] += points[3 * c + 2];
        }
        center = (1.0 / 3) * center;
        double nrm = std::sqrt(center[0] * center[0] + center[1] * center[1] + center[2] * center[2]);
        center = (radius / nrm) * center;
        for (int d = 0; d < 3; ++d) vpos[3 * t + d] = center[d];
        for (int k = 0; k < 3; ++k) {
            int prev = tris[3 * t + (k + 2) % 3], cur = tris[3 * t + k];
            vcn[3 * t + k] = cur;
            // the facet on the other side of halfedge prev->cur holds cur->prev
            vn[3 * t + k] = tri_of_edge(cur, prev);
        }
    }
    int maxDeg = 0;
    for (int c = 0; c < n; ++c) maxDeg = std::max<int>(maxDeg, (int)out_edges[c].size());
    if (maxDeg > ring_cap) return -maxDeg;
    for (int c = 0; c < n; ++c) {
        // triangles around c, clockwise from outside: from triangle (c, b, x)
        // step to the triangle holding the edge c->x... walk via edge c->b:
        // the triangle containing directed edge c->b is (c,b,*); the next one
        // clockwise shares the edge (c, prev) where prev is its third vertex.
        int deg = (int)out_edges[c].size();
        int start = -1, startb = -1;
        for (auto& e : out_edges[c]) if (start < 0 || e.second < start) { start = e.second; startb = e.first; }
        int t = start, b = startb;
        for (int k = 0; k < deg; ++k) {
            ring[c * ring_cap + k] = t;
            // triangle t = (c, b, x) in ccw order; clockwise neighbour around c
            // is the triangle across edge (c, b), which holds directed edge b->c
            // and is (c, y, b): its out-edge from c goes to y.
            int tn = tri_of_edge(b, c);
            int y = -1;
            for (int kk = 0; kk < 3; ++kk)
                if (tris[3 * tn + kk] == c) y = tris[3 * tn + (kk + 1) % 3];
            t = tn; b = y;
        }
        cnn[c] = deg;
    }
    return maxDeg;
}

void* or_create(int nCells, int nVertices, double R, const double* pos, const int* vn, const int* vcn,
                const int* ring, int maxDeg, const int* cnn) {
    Model* m = new Model();
    m->nCells = nCells;
    m->N = nVertices;
    m->maxDeg = maxDeg;
    m->sphere = Sphere(R);
    m->positions.resize(nVertices);
    m->velocities.assign(nVertices, dVec(0.0));
    m->forces.assign(nVertices, dVec(0.0));
    m->directors.assign(nVertices, dVec(0.0));
    m->masses.assign(nVertices, 1.0);
    m->cellPositions.assign(nCells, dVec(0.0));
    m->curVert.assign(3 * nVertices, dVec(0.0));
    m->lastVert.assign(3 * nVertices, dVec(0.0));
    m->nextVert.assign(3 * nVertices, dVec(0.0));
    for (int i = 0; i < nVertices; ++i)
        for (int d = 0; d < 3; ++d) m->positions[i][d] = pos[3 * i + d];
    m->neighbors.assign(vn, vn + 3 * nVertices);
    m->vertexCellNeighbors.assign(vcn, vcn + 3 * nVertices);
    m->numberOfNeighbors.assign(nVertices, 3u);
    m->cellNeighbors.assign(ring, ring + (size_t)nCells * maxDeg);
    m->cellNumberOfNeighbors.assign(cnn, cnn + nCells);
    m->ap.assign(2 * nCells, 0.0);
    m->app.assign(2 * nCells, 0.0);
    for (int c = 0; c < nCells; ++c) { m->app[2 * c] = 1.0; m->app[2 * c + 1] = 3.8; }
    return m;
}

void or_destroy(void* h) { delete (Model*)h; }

void or_set_preferences(void* h, const double* a0p0) {
    Model* m = (Model*)h;
    for (int c = 0; c < 2 * m->nCells; ++c) m->app[c] = a0p0[c];
}
void or_set_reduce_mode(void* h, int mode) { ((Model*)h)->reduce_mode = mode; }
void or_set_kr(void* h, double kr) { ((Model*)h)->Kr = kr; }
void or_set_t1_threshold(void* h, double t) { ((Model*)h)->t1Threshold = t; }
void or_set_directors(void* h, const double* n) {
    Model* m = (Model*)h;
    for (int i = 0; i < m->N; ++i) for (int d = 0; d < 3; ++d) m->directors[i][d] = n[3 * i + d];
}
void or_set_velocities(void* h, const double* v) {
    Model* m = (Model*)h;
    for (int i = 0; i < m->N; ++i) for (int d = 0; d < 3; ++d) m->velocities[i][d] = v[3 * i + d];
}
void or_set_positions(void* h, const double* p) {
    Model* m = (Model*)h;
    for (int i = 0; i < m->N; ++i) for (int d = 0; d < 3; ++d) m->positions[i][d] = p[3 * i + d];
}

void or_compute_geometry(void* h) { ((Model*)h)->computeGeometry(); }
void or_compute_forces(void* h) { ((Model*)h)->computeForces(); }
double or_energy(void* h, int reduce_mode) { return ((Model*)h)->computeEnergy(reduce_mode); }
double or_kinetic_energy(void* h, int reduce_mode) { return ((Model*)h)->computeKineticEnergy(reduce_mode); }
void or_enforce_topology(void* h) { ((Model*)h)->testAndPerformT1Transitions(); }

void or_move(void* h, const double* disp, double scale) {
    Model* m = (Model*)h;
    std::vector<dVec> d(m->N);
    for (int i = 0; i < m->N; ++i) for (int k = 0; k < 3; ++k) d[i][k] = disp[3 * i + k];
    m->moveParticles(d, scale);
}

// brownianDynamics::integrateEOMCPU, src/updaters/brownianDynamics.cpp:29-48.
// Noise for component dd of vertex ii at timestep `step`: normal3(seed, step, ii).
void or_brownian_steps(void* h, int nsteps, double deltaT, double mu, double temperature,
                       uint64_t seed, uint64_t step0) {
    Model* m = (Model*)h;
    std::vector<dVec> disp(m->N);
    for (int s = 0; s < nsteps; ++s) {
        uint64_t step = step0 + (uint64_t)s;
        m->computeForces();
        double forcePrefactor = deltaT * mu;
        double noisePrefactor = std::sqrt(2.0 * forcePrefactor * temperature);
        for (int ii = 0; ii < m->N; ++ii) {
            double xi[3] = {0.0, 0.0, 0.0};
            if (temperature != 0.0) svm::normal3(seed, step, (uint32_t)ii, xi);
            for (int dd = 0; dd < 3; ++dd)
                disp[ii][dd] = xi[dd] * noisePrefactor + m->forces[ii][dd] * forcePrefactor;
        }
        m->moveParticles(disp, 1.0);
    }
}

// sphericalSelfPropelledParticle::integrateEOMCPU,
// src/updaters/sphericalSelfPropelledParticle.cpp:16-47.  The rotation angle
// of vertex ii at timestep `step` uses uniform draw (STREAM_ACTIVE, ii).
void or_active_steps(void* h, int nsteps, double deltaT, double mu, double v0, double tau,
                     uint64_t seed, uint64_t step0) {
    Model* m = (Model*)h;
    std::vector<dVec> disp(m->N);
    for (int s = 0; s < nsteps; ++s) {
        uint64_t step = step0 + (uint64_t)s;
        m->computeForces();
        for (int ii = 0; ii < m->N; ++ii)
            disp[ii] = deltaT * (v0 * m->directors[ii] + mu * m->forces[ii]);
        m->moveParticles(disp, 1.0);
        for (int ii = 0; ii < m->N; ++ii) {
            double u, unused;
            svm::uniform2(seed, svm::STREAM_ACTIVE, step, (uint32_t)ii, &u, &unused);
            double randomNumber = u * (.5 - (-.5)) + (-.5);
            double theta = randomNumber * std::sqrt(2.0 * deltaT * tau);
            m->sphere.projectToTangentPlaneAndNormalize(m->directors[ii], m->positions[ii]);
            dVec nDisp = m->directors[ii];
            rodriguesRotation(nDisp, m->positions[ii], theta);
            m->directors[ii] = nDisp;
        }
    }
}

// sphericalVectorialVicsek::integrateEOMCPU
// (src/updaters/sphericalVectorialVicsek.cpp:5-66) on the vertex model, whose
// degrees of freedom are the vertices and their neighbours the three vertex
// neighbours.  The random direction of vertex ii in the step that takes the
// step count from k to k + 1 uses uniform draws (STREAM_VICSEK, ii) at k + 1.
void or_vicsek_steps(void* h, int nsteps, double deltaT, double mu, double v0, double tau, double Eta,
                     uint64_t seed, uint64_t step0) {
    Model* m = (Model*)h;
    std::vector<dVec> disp(m->N), nDisp(m->N);
    for (int s = 0; s < nsteps; ++s) {
        uint64_t step = step0 + (uint64_t)s + 1;
        m->computeForces();
        for (int ii = 0; ii < m->N; ++ii)
            disp[ii] = deltaT * (v0 * m->directors[ii] + mu * m->forces[ii]);
        m->moveParticles(disp, 1.0);
        for (int ii = 0; ii < m->N; ++ii) {
            double u, w;
            svm::uniform2(seed, svm::STREAM_VICSEK, step, (uint32_t)ii, &u, &w);
            double phi = 2.0 * PI * u;
            double theta = O_ACOS(2.0 * w - 1);
            dVec spherePoint;
            spherePoint[0] = 1.0 * O_SIN(theta) * O_COS(phi);
            spherePoint[1] = 1.0 * O_SIN(theta) * O_SIN(phi);
            spherePoint[2] = 1.0 * O_COS(theta);
            m->sphere.projectToTangentPlane(spherePoint, m->positions[ii]);
            spherePoint = spherePoint * (1.0 / norm(spherePoint));
            int mm = 3;
            nDisp[ii] = m->directors[ii];
            for (int jj = 0; jj < mm; ++jj) nDisp[ii] += m->directors[m->neighbors[3 * ii + jj]];
            mm += 1;
            double mi = 1.0 / mm;
            nDisp[ii] = nDisp[ii] * mi + spherePoint * Eta;
            m->sphere.projectToTangentPlaneAndNormalize(nDisp[ii], m->positions[ii]);
        }
        for (int ii = 0; ii < m->N; ++ii) {
            m->directors[ii] += (deltaT / tau) * (nDisp[ii] - m->directors[ii]);
            m->sphere.projectToTangentPlane(m->directors[ii], m->positions[ii]);
            m->directors[ii] = m->directors[ii] * (1.0 / norm(m->directors[ii]));
        }
    }
}

// energyMinimizerFIRE::minimize + velocityVerlet::integrateEOMCPU + fireStepCPU
// (src/updaters/energyMinimizerFIRE.cpp:181-259, src/updaters/velocityVerlet.cpp:6-38).
// params: [deltaT, alphaStart, deltaTMax, deltaTInc, deltaTDec, alphaDec,
//          NMin, forceCutoff, alphaMin, maxIterations]
// state (in/out): [iterations, deltaT, alpha, NSinceNegativePower, forceMax, Power]
// stop_rms: 0 = reference stop test forceMax = sqrt(sum F^2)/Ndof;
//           1 = rms test sqrt(sum F^2 / Ndof).
void or_fire_minimize(void* h, const double* params, double* state, int reduce_mode, int stop_rms) {
    Model* m = (Model*)h;
    double alphaStart = params[1], deltaTMax = params[2], deltaTInc = params[3];
    double deltaTDec = params[4], alphaDec = params[5];
    int NMin = (int)params[6];
    double forceCutoff = params[7], alphaMin = params[8];
    int maxIterations = (int)params[9];
    double deltaTMin = params[0] * .01;          // energyMinimizerFIRE.h:41
    int iterations = (int)state[0];
    double deltaT = state[1], alpha = state[2];
    int NSinceNegativePower = (int)state[3];
    double Power = state[5];
    int Ndof = m->N;
    std::vector<dVec> disp(Ndof);
    m->computeForces();
    double forceMax = 110.0;
    std::vector<double> tf(Ndof), tp(Ndof), tv(Ndof);
    while ((iterations < maxIterations) && (forceMax > forceCutoff)) {
        iterations += 1;
        // velocity Verlet
        for (int i = 0; i < Ndof; ++i) {
            disp[i] = deltaT * m->velocities[i] + (0.5 * deltaT * deltaT) * m->forces[i];
            m->velocities[i] += (0.5 / m->masses[i]) * deltaT * m->forces[i];
        }
        m->moveParticles(disp, 1.0);
        m->computeForces();
        for (int i = 0; i < Ndof; ++i) m->velocities[i] += (0.5 / m->masses[i]) * deltaT * m->forces[i];
        // FIRE step
        Power = 0.0;
        double forceNorm = 0.0, velocityNorm = 0.0;
        if (reduce_mode == 0) {
            for (int i = 0; i < Ndof; ++i) {
                Power += dot(m->forces[i], m->velocities[i]);
                double fdot = dot(m->forces[i], m->forces[i]);
                forceNorm += fdot;
                velocityNorm += dot(m->velocities[i], m->velocities[i]);
            }
        } else {
            for (int i = 0; i < Ndof; ++i) {
                tp[i] = dot(m->forces[i], m->velocities[i]);
                tf[i] = dot(m->forces[i], m->forces[i]);
                tv[i] = dot(m->velocities[i], m->velocities[i]);
            }
            Power = tree_sum(tp);
            forceNorm = tree_sum(tf);
            velocityNorm = tree_sum(tv);
        }
        if (stop_rms) forceMax = std::sqrt(forceNorm / (double)Ndof);
        else forceMax = std::sqrt(forceNorm) / (double)Ndof;
        double scaling = 0.0;
        if (forceNorm > 0.) scaling = std::sqrt(velocityNorm / forceNorm);
        for (int i = 0; i < Ndof; ++i)
            m->velocities[i] = (1.0 - alpha) * m->velocities[i] + alpha * scaling * m->forces[i];
        if (Power > 0) {
            if (NSinceNegativePower > NMin) {
                deltaT = std::min(deltaT * deltaTInc, deltaTMax);
                alpha = alpha * alphaDec;
                alpha = std::max(alpha, alphaMin);
            }
            NSinceNegativePower += 1;
        } else {
            NSinceNegativePower = 0;
            deltaT = deltaT * deltaTDec;
            deltaT = std::max(deltaT, deltaTMin);
            alpha = alphaStart;
            for (int i = 0; i < Ndof; ++i) m->velocities[i] = dVec(0.0);
        }
    }
    state[0] = iterations;
    state[1] = deltaT;
    state[2] = alpha;
    state[3] = NSinceNegativePower;
    state[4] = forceMax;
    state[5] = Power;
}

// energyMinimizerAdam::minimize + adamStepCPU (src/updaters/energyMinimizerAdam.cpp:55-103).
// params: [beta1, beta2, epsilon, deltaT, forceCutoff, maxIterations]
// state (in/out): [iterations, beta1t, beta2t, forceMax]
// mom, mom2 (in/out): the biased first and second moment estimates, [N][3].
// Every iteration computes the forces, updates the moments, moves by
// -deltaT mc / sqrt(vc) per component (epsilon where sqrt(vc) is 0) and
// enforces the topology; forceMax = sqrt(sum F^2) / N of that iteration's
// forces.
void or_adam_minimize(void* h, const double* params, double* state, double* mom, double* mom2, int reduce_mode) {
    Model* m = (Model*)h;
    double beta1 = params[0], beta2 = params[1], epsilon = params[2], deltaT = params[3];
    double forceCutoff = params[4];
    int maxIterations = (int)params[5];
    int iterations = (int)state[0];
    double beta1t = state[1], beta2t = state[2];
    int Ndof = m->N;
    std::vector<dVec> disp(Ndof);
    std::vector<double> tf(Ndof);
    double forceMax = 110.0;
    while ((iterations < maxIterations) && (forceMax > forceCutoff)) {
        iterations += 1;
        m->computeForces();
        double forceNorm = 0.0;
        for (int nn = 0; nn < Ndof; ++nn) {
            const dVec& g = m->forces[nn];
            double fdot = dot(g, g);
            if (reduce_mode == 0) forceNorm += fdot;
            else tf[nn] = fdot;
            for (int dd = 0; dd < 3; ++dd) {
                double mm = beta1 * mom[3 * nn + dd] + (beta1 - 1) * g[dd];
                double vv = beta2 * mom2[3 * nn + dd] + (1 - beta2) * (g[dd] * g[dd]);
                mom[3 * nn + dd] = mm;
                mom2[3 * nn + dd] = vv;
                double mc = (1.0 / (1.0 - beta1t)) * mm;
                double vc = (1.0 / (1.0 - beta2t)) * vv;
                double rootvc = std::sqrt(vc);
                if (rootvc == 0) rootvc = epsilon;
                disp[nn][dd] = -deltaT * mc / (rootvc);
            }
        }
        if (reduce_mode) forceNorm = tree_sum(tf);
        m->moveParticles(disp, 1.0);
        forceMax = std::sqrt(forceNorm) / Ndof;
        beta1t *= beta1;
        beta2t *= beta2;
    }
    state[0] = iterations;
    state[1] = beta1t;
    state[2] = beta2t;
    state[3] = forceMax;
}

// dynamicalFeatures::computeMSD and computeOverlapFunction
// (src/analysis/dynamicalFeatures.cpp:17-53) on the cell positions, against
// ref [nCells][3]: the first floor(nCells * fraction) cells (all when
// fraction >= 1); a NaN distance adds nothing to the MSD but still counts in
// N; the overlap counts distances below the cutoff.  reduce_mode 1 sums with
// tree_sum (the device's order).  out: [msd, overlap].
void or_msd_overlap(void* h, const double* ref, double fraction, double cutoff, int reduce_mode, double* out) {
    Model* m = (Model*)h;
    int N = m->nCells;
    if (fraction < 1) N = (int)std::floor(N * fraction);
    std::vector<double> sq(N), ov(N);
    double msd = 0.0, overlap = 0.0;
    for (int ii = 0; ii < N; ++ii) {
        dVec init(0.0);
        for (int d = 0; d < 3; ++d) init[d] = ref[3 * ii + d];
        dVec cur = m->cellPositions[ii];
        double disp;
        m->sphere.geodesicDistance(init, cur, disp);
        sq[ii] = std::isnan(disp) ? 0.0 : disp * disp;
        ov[ii] = disp < cutoff ? 1.0 : 0.0;
        msd += sq[ii];
        overlap += ov[ii];
    }
    if (reduce_mode) {
        msd = tree_sum(sq);
        overlap = tree_sum(ov);
    }
    out[0] = msd / N;
    out[1] = overlap / N;
}

// ---- accessors ----
int or_num_vertices(void* h) { return ((Model*)h)->N; }
int or_max_deg(void* h) { return ((Model*)h)->maxDeg; }
int or_num_t1(void* h) { return ((Model*)h)->nT1; }
int or_num_t1_skipped(void* h) { return ((Model*)h)->nT1Skipped; }
int or_num_fan(void* h) { return ((Model*)h)->nFanRecompute; }
int or_num_nan(void* h) { return ((Model*)h)->nNan; }
void or_get_positions(void* h, double* out) {
    Model* m = (Model*)h;
    for (int i = 0; i < m->N; ++i) for (int d = 0; d < 3; ++d) out[3 * i + d] = m->positions[i][d];
}
void or_get_forces(void* h, double* out) {
    Model* m = (Model*)h;
    for (int i = 0; i < m->N; ++i) for (int d = 0; d < 3; ++d) out[3 * i + d] = m->forces[i][d];
}
void or_get_velocities(void* h, double* out) {
    Model* m = (Model*)h;
    for (int i = 0; i < m->N; ++i) for (int d = 0; d < 3; ++d) out[3 * i + d] = m->velocities[i][d];
}
void or_get_directors(void* h, double* out) {
    Model* m = (Model*)h;
    for (int i = 0; i < m->N; ++i) for (int d = 0; d < 3; ++d) out[3 * i + d] = m->directors[i][d];
}
void or_get_cell_positions(void* h, double* out) {
    Model* m = (Model*)h;
    for (int i = 0; i < m->nCells; ++i) for (int d = 0; d < 3; ++d) out[3 * i + d] = m->cellPositions[i][d];
}
void or_get_area_perimeter(void* h, double* out) {
    Model* m = (Model*)h;
    for (int i = 0; i < 2 * m->nCells; ++i) out[i] = m->ap[i];
}
// topology in the reference layout; ring is written with stride `cap`.
int or_get_topology(void* h, int* vn, int* vcn, int* ring, int cap, int* cnn) {
    Model* m = (Model*)h;
    if (cap < m->maxDeg) {
        int need = 0;
        for (int c = 0; c < m->nCells; ++c) need = std::max<int>(need, m->cellNumberOfNeighbors[c]);
        if (cap < need) return -need;
    }
    for (int i = 0; i < 3 * m->N; ++i) { vn[i] = m->neighbors[i]; vcn[i] = m->vertexCellNeighbors[i]; }
    for (int c = 0; c < m->nCells; ++c) {
        cnn[c] = m->cellNumberOfNeighbors[c];
        for (int k = 0; k < cap; ++k)
            ring[c * cap + k] = k < (int)m->cellNumberOfNeighbors[c] ? m->cellNeighbors[m->cni(k, c)] : -1;
    }
    return 0;
}
double or_tree_sum(const double* x, int n) { return tree_sum(std::vector<double>(x, x + n)); }

// math entry points for the tests
double or_acos(double x) { return O_ACOS(x); }
// the library's short-arc form of acos, compared with svm::acos by the tests
double or_acos_hi(double x) { return svm::acos_hi(x); }
double or_acos_lo(double x) { return svm::acos_lo(x); }
double or_own_acos(double x) { return svm::acos(x); }
// the library's constant for cos(+-pi/2) in the T1 quarter turn
double or_cos_quarter_turn() { return svm::kCosQuarterTurn; }
double or_sin(double x) { return O_SIN(x); }
double or_cos(double x) { return O_COS(x); }
double or_log(double x) { return svm::log(x); }
void or_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    svm::u32x4 r = svm::philox4x32_10(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1]);
    for (int i = 0; i < 4; ++i) out[i] = r.v[i];
}
// the draws the updaters consume, for the tests that pin the updaters
void or_uniform2(uint64_t seed, uint32_t stream, uint64_t step, uint32_t idx, double* out) {
    svm::uniform2(seed, stream, step, idx, &out[0], &out[1]);
}
void or_normal3(uint64_t seed, uint64_t step, uint32_t v, double* out) { svm::normal3(seed, step, v, out); }

}  // extern "C"
