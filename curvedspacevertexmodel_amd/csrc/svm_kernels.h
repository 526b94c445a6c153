// svm_kernels.h -- HIP kernels of the per-timestep hot path (gfx950, wave64).
//
// One dynamics step of a batch of spheres is four launches over the whole
// batch (every sphere's tiles in one grid):
//   k_geometry  (thread per cell)   cell centre, area, perimeter; per-sphere
//                                   total-area check with the fan fallback
//   k_vertex    (thread per vertex) chain-rule force over the 3 cells of the
//                                   vertex, fused with the updater epilogue
//                                   (Brownian / active displacement and the
//                                   move on the sphere, or the FIRE
//                                   half-kick, power/norm reductions and the
//                                   FIRE controller)
//   k_t1_detect (thread per vertex) short-edge test -> candidate bit mask;
//                                   active-director update
//   k_t1_resolve (block per sphere) applies T1 flips in exactly the order of
//                                   the reference's sequential CPU loop
// Per-sphere reductions are deterministic: 256-row tiles folded by a fixed
// halving tree, then the tile partials folded the same way by the last block
// of that sphere to arrive.
#pragma once
#include <hip/hip_runtime.h>
#include "svm_device.h"
#include "svm_geom.h"

namespace svm {

enum VertexMode { VM_FORCE = 0, VM_BROWNIAN = 1, VM_ACTIVE = 2, VM_FIRE = 3, VM_FORCE_F32 = 4 };

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

// True in every thread of the last block (of n) to reach this point for one
// sphere.  The caller's partial-sum stores must precede the call.
__device__ inline bool arrive_last(int32_t* counter, int n) {
    __shared__ int s_last;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        int prev = atomicAdd(counter, 1);
        s_last = (prev == n - 1);
        if (s_last) atomicExch(counter, 0);
    }
    __syncthreads();
    if (s_last) __threadfence();
    return s_last != 0;
}

// In-place halving tree over part[i*stride], i < n, padded with zeros to the
// next power of two.  Whole block; the result is valid in every thread.
__device__ inline double fold_partials(double* part, int stride, int n) {
    int m = 1;
    while (m < n) m <<= 1;
    for (int h = m >> 1; h >= 1; h >>= 1) {
        for (int i = threadIdx.x; i < h; i += blockDim.x) {
            if (i < n) {
                double a = __hip_atomic_load(part + (size_t)i * stride, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                double b = (i + h < n) ? __hip_atomic_load(part + (size_t)(i + h) * stride, __ATOMIC_RELAXED,
                                                           __HIP_MEMORY_SCOPE_AGENT)
                                       : 0.0;
                __hip_atomic_store(part + (size_t)i * stride, a + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        __threadfence_block();
        __syncthreads();
    }
    double r = __hip_atomic_load(part, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    return r;
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
// and recomputeAreasCPU (:162-207): one thread per cell.
__global__ __launch_bounds__(SVM_TILE) void k_geometry(Dev d, int pin, int skip_frozen) {
    __shared__ double red[SVM_TILE + 1];
    const int tile = blockIdx.x;
    const int s = d.ctile_sys[tile];
    SysParams* S = d.sys + s;
    if (skip_frozen && S->f_frozen) return;
    const int c = tile * SVM_TILE + threadIdx.x;
    const int local = c - S->co;
    const double R = S->R;
    const double* X = d.px[pin];
    const double* Y = d.py[pin];
    const double* Z = d.pz[pin];
    const int Cs = d.Cs;
    double area = 0.0;
    if (local < S->nc) {
        const int n = d.cnn[c];
        const int32_t* rc = d.ring + c;
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (int k = 0; k < n; ++k) {
            int v = rc[(size_t)k * Cs];
            sx = sx + X[v];
            sy = sy + Y[v];
            sz = sz + Z[v];
        }
        double cs[3] = {sx, sy, sz};
        double inv = 1.0 / __builtin_sqrt(dot0(cs, cs));
        double cxv = R * (inv * sx), cyv = R * (inv * sy), czv = R * (inv * sz);
        d.cx[c] = cxv;
        d.cy[c] = cyv;
        d.cz[c] = czv;

        int il = rc[(size_t)(n - 3) * Cs], ic = rc[(size_t)(n - 2) * Cs], in = rc[(size_t)(n - 1) * Cs];
        double pl[3] = {X[il], Y[il], Z[il]};
        double pc[3] = {X[ic], Y[ic], Z[ic]};
        double pn[3] = {X[in], Y[in], Z[in]};
        double nl = pnorm(pl), nc = pnorm(pc), nnx = pnorm(pn);
        double cm[3] = {pl[1] * pc[2] - pl[2] * pc[1], pl[2] * pc[0] - pl[0] * pc[2], pl[0] * pc[1] - pl[1] * pc[0]};
        double ncm = __builtin_sqrt(dot0(cm, cm));
        double perimeter = 0.0, angles = 0.0;
        for (int k = 0; k < n; ++k) {
            int ip = rc[(size_t)k * Cs];
            double pp[3] = {X[ip], Y[ip], Z[ip]};
            perimeter += geodesic(pl, pc, nl, nc, R);
            double ci[3] = {pc[1] * pn[2] - pc[2] * pn[1], pc[2] * pn[0] - pc[0] * pn[2], pc[0] * pn[1] - pc[1] * pn[0]};
            double nci = __builtin_sqrt(dot0(ci, ci));
            double determinant = pl[0] * ci[0] + pl[1] * ci[1] + pl[2] * ci[2];
            double angle = svm::acos(-dot0(ci, cm) / (nci * ncm));
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
    }
    double part = block_tree256(area, red);
    if (threadIdx.x == 0) d.cpart[tile] = part;
    if (arrive_last(&S->counter_geo, S->nctiles)) {
        double total = fold_partials(d.cpart + S->ctile0, 1, S->nctiles);
        double excessArea = total - 4.0 * SVM_PI * R * R;
        if (__builtin_fabs(excessArea) > 1e-6) {
            // triangle-fan areas for every cell of this sphere
            for (int cc = S->co + threadIdx.x; cc < S->co + S->nc; cc += blockDim.x) {
                const int n = d.cnn[cc];
                const int32_t* rc = d.ring + cc;
                double cp[3] = {d.cx[cc], d.cy[cc], d.cz[cc]};
                int il = rc[(size_t)(n - 2) * Cs], ic = rc[(size_t)(n - 1) * Cs];
                double pl[3] = {X[il], Y[il], Z[il]};
                double pc[3] = {X[ic], Y[ic], Z[ic]};
                double a = 0.0;
                for (int k = 0; k < n; ++k) {
                    int ip = rc[(size_t)k * Cs];
                    double pn[3] = {X[ip], Y[ip], Z[ip]};
                    a += triangle_area(cp, pl, pc, R);
                    pl[0] = pc[0]; pl[1] = pc[1]; pl[2] = pc[2];
                    pc[0] = pn[0]; pc[1] = pn[1]; pc[2] = pn[2];
                }
                d.A[cc] = a;
            }
            if (threadIdx.x == 0) S->n_fan += 1;
        }
    }
}

// --------------------------------------------------------------- vertices --
__device__ inline PQ pq_at(const double* X, const double* Y, const double* Z, int v) {
    return make_pq(X[v], Y[v], Z[v]);
}

__device__ inline const PQ& pick3(int k, const PQ& a, const PQ& b, const PQ& c) {
    return k == 0 ? a : (k == 1 ? b : c);
}

// Force on vertex v (sphericalVertexModel::computeForceCPU,
// src/model/sphericalVertexModel.cpp:380-416).  Returns the number of skipped
// NaN contributions.
__device__ inline int vertex_force(const Dev& d, const SysParams* S, const double* X, const double* Y,
                                   const double* Z, int v, double* f) {
    const double R = S->R, Kr = S->kr;
    PQ qs = pq_at(X, Y, Z, v);
    int4 nb = d.vn[v];
    int4 cl = d.vcn[v];
    int ln = d.vln[v];
    PQ q0 = pq_at(X, Y, Z, nb.x);
    PQ q1 = pq_at(X, Y, Z, nb.y);
    PQ q2 = pq_at(X, Y, Z, nb.z);
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    int nan = 0;
#pragma unroll
    for (int sl = 0; sl < 3; ++sl) {
        int c = sl == 0 ? cl.x : (sl == 1 ? cl.y : cl.z);
        int ls = (ln >> (4 * sl)) & 3;
        int ns = (ln >> (4 * sl + 2)) & 3;
        double areaDifference = d.A[c] - d.a0[c];
        double perimeterDifference = d.P[c] - d.p0[c];
        PQ qc = make_pq(d.cx[c], d.cy[c], d.cz[c]);
        double fs[3];
        if (pair_force(qs, pick3(ls, q0, q1, q2), pick3(ns, q0, q1, q2), qc, areaDifference, perimeterDifference,
                       Kr, R, fs)) {
            f0 = f0 + fs[0];
            f1 = f1 + fs[1];
            f2 = f2 + fs[2];
        } else {
            ++nan;
        }
    }
    f[0] = f0; f[1] = f1; f[2] = f2;
    return nan;
}

template <int MODE>
__global__ __launch_bounds__(SVM_TILE) void k_vertex(Dev d, int pin, int pout, int nin, int store_forces) {
    __shared__ double red[SVM_TILE + 1];
    const int tile = blockIdx.x;
    const int s = d.vtile_sys[tile];
    SysParams* S = d.sys + s;
    if (MODE == VM_FIRE && S->f_frozen) return;
    const int v = tile * SVM_TILE + threadIdx.x;
    const int local = v - S->vo;
    const bool valid = local < S->nv;
    const double* X = d.px[pin];
    const double* Y = d.py[pin];
    const double* Z = d.pz[pin];
    double f[3] = {0.0, 0.0, 0.0};
    double tp = 0.0, tf = 0.0, tv = 0.0;
    int nan = 0;
    if (valid) {
        nan = vertex_force(d, S, X, Y, Z, v, f);
        if (MODE == VM_FORCE || store_forces) {
            d.fx[v] = f[0]; d.fy[v] = f[1]; d.fz[v] = f[2];
        }
        if (MODE == VM_BROWNIAN || MODE == VM_ACTIVE) {
            double disp[3];
            if (MODE == VM_BROWNIAN) {
                double forcePrefactor = S->dt * S->mu;
                double noisePrefactor = __builtin_sqrt(2.0 * forcePrefactor * S->T);
                double xi[3] = {0.0, 0.0, 0.0};
                if (S->T != 0.0) normal3(S->seed, S->step, (uint32_t)local, xi);
                for (int k = 0; k < 3; ++k) disp[k] = xi[k] * noisePrefactor + f[k] * forcePrefactor;
            } else {
                double n0 = d.nx[nin][v], n1 = d.ny[nin][v], n2 = d.nz[nin][v];
                double dt = S->dt, v0 = S->v0, mu = S->mu;
                disp[0] = dt * (v0 * n0 + mu * f[0]);
                disp[1] = dt * (v0 * n1 + mu * f[1]);
                disp[2] = dt * (v0 * n2 + mu * f[2]);
            }
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
    if (MODE == VM_FIRE) {
        double sp = block_tree256(tp, red);
        double sf = block_tree256(tf, red);
        double sv = block_tree256(tv, red);
        if (threadIdx.x == 0) {
            d.vpart[3 * tile + 0] = sp;
            d.vpart[3 * tile + 1] = sf;
            d.vpart[3 * tile + 2] = sv;
        }
        if (arrive_last(&S->counter_vtx, S->nvtiles)) {
            double Power = fold_partials(d.vpart + 3 * S->vtile0 + 0, 3, S->nvtiles);
            double forceNorm = fold_partials(d.vpart + 3 * S->vtile0 + 1, 3, S->nvtiles);
            double velocityNorm = fold_partials(d.vpart + 3 * S->vtile0 + 2, 3, S->nvtiles);
            if (threadIdx.x == 0) {
                // energyMinimizerFIRE::fireStepCPU (src/updaters/energyMinimizerFIRE.cpp:199-232)
                double Ndof = (double)S->nv;
                double forceMax = S->f_stop_rms ? __builtin_sqrt(forceNorm / Ndof) : __builtin_sqrt(forceNorm) / Ndof;
                double scaling = 0.0;
                if (forceNorm > 0.) scaling = __builtin_sqrt(velocityNorm / forceNorm);
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
    const int tile = blockIdx.x;
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
// src/model/sphericalVertexModel.cpp:111-131) for sphere `only` (-1: all).
__global__ __launch_bounds__(SVM_TILE) void k_move_ext(Dev d, int pin, int pout, int only, double scale) {
    const int tile = blockIdx.x;
    const int s = d.vtile_sys[tile];
    const SysParams* S = d.sys + s;
    const int v = tile * SVM_TILE + threadIdx.x;
    if (v - S->vo >= S->nv) return;
    double p[3] = {d.px[pin][v], d.py[pin][v], d.pz[pin][v]};
    if (only < 0 || only == s) {
        double disp[3] = {d.disp[3 * (size_t)v], d.disp[3 * (size_t)v + 1], d.disp[3 * (size_t)v + 2]};
        if (scale != 1.) { disp[0] = scale * disp[0]; disp[1] = scale * disp[1]; disp[2] = scale * disp[2]; }
        move_on_sphere(p, disp, S->R);
    }
    d.px[pout][v] = p[0];
    d.py[pout][v] = p[1];
    d.pz[pout][v] = p[2];
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

// ---------------------------------------------------------------- T1 ------
// Short-edge test of testAndPerformT1TransitionsCPU
// (src/model/sphericalVertexModel.cpp:469-481): bit v of t1mask is set when
// vertex v has an edge (v, w), v < w, shorter than the threshold.
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
    if ((threadIdx.x & 63) == 0) {
        d.t1mask[v >> 6] = m;
        if (m) atomicOr(&S->t1_flag, 1);
    }
}

__device__ inline int& ringref(const Dev& d, int c, int k) { return d.ring[(size_t)k * d.Cs + c]; }
__device__ inline int& vcnref(const Dev& d, int v, int k) { return ((int*)&d.vcn[v])[k]; }
__device__ inline int& vnref(const Dev& d, int v, int k) { return ((int*)&d.vn[v])[k]; }

// getCellVertexSetForT1 (src/model/sphericalVertexModel.cpp:611-731).
// Returns false when the four cells cannot all be determined (vertex1 and
// vertex2 share all three cells, so there is no cell l); the flip is then
// skipped, where the reference would read an uninitialised cell set.
__device__ inline bool cell_vertex_set(const Dev& d, int vertex1, int vertex2, int* cs, int* vs) {
    cs[0] = cs[1] = cs[2] = cs[3] = -1;
    int cell1 = vcnref(d, vertex1, 0), cell2 = vcnref(d, vertex1, 1), cell3 = vcnref(d, vertex1, 2);
    for (int ff = 0; ff < 3; ++ff) {
        int ctest = vcnref(d, vertex2, ff);
        if (ctest != cell1 && ctest != cell2 && ctest != cell3) cs[3] = ctest;
    }
    int cells[3] = {cell1, cell2, cell3};
    int vlast, vcur, vnext = -1, cneigh;
    for (int t = 0; t < 3; ++t) {
        int cell = cells[t];
        cneigh = d.cnn[cell];
        vlast = ringref(d, cell, cneigh - 2);
        vcur = ringref(d, cell, cneigh - 1);
        for (int cn = 0; cn < cneigh; ++cn) {
            vnext = ringref(d, cell, cn);
            if (vcur == vertex1) break;
            vlast = vcur;
            vcur = vnext;
        }
        if (vlast == vertex2) cs[0] = cell;
        else if (vnext == vertex2) cs[2] = cell;
        else cs[1] = cell;
    }
    if (cs[0] < 0 || cs[1] < 0 || cs[2] < 0 || cs[3] < 0) return false;
    cneigh = d.cnn[cs[1]];
    vlast = ringref(d, cs[1], cneigh - 2);
    vcur = ringref(d, cs[1], cneigh - 1);
    for (int cn = 0; cn < cneigh; ++cn) {
        vnext = ringref(d, cs[1], cn);
        if (vcur == vertex1) break;
        vlast = vcur;
        vcur = vnext;
    }
    vs[0] = vlast;
    vs[1] = vnext;
    cneigh = d.cnn[cs[3]];
    vlast = ringref(d, cs[3], cneigh - 2);
    vcur = ringref(d, cs[3], cneigh - 1);
    for (int cn = 0; cn < cneigh; ++cn) {
        vnext = ringref(d, cs[3], cn);
        if (vcur == vertex2) break;
        vlast = vcur;
        vcur = vnext;
    }
    vs[3] = vlast;
    vs[2] = vnext;
    return true;
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

// One T1 flip of edge (vertex1, vertex2): the body of
// testAndPerformT1TransitionsCPU (src/model/sphericalVertexModel.cpp:496-598).
// v1 is the caller's copy of vertex1's position.  Returns 0, or 4 when a
// cell would outgrow the ring capacity (nothing is modified then).
__device__ inline int t1_flip(const Dev& d, int pcur, int vertex1, int vertex2, const int* cs, const int* vs,
                              double* v1, double R, int* tmp) {
    if (d.cnn[cs[1]] + 1 > d.cap || d.cnn[cs[3]] + 1 > d.cap) return 4;
    double* X = d.px[pcur];
    double* Y = d.py[pcur];
    double* Z = d.pz[pcur];
    double v2[3] = {X[vertex2], Y[vertex2], Z[vertex2]};
    double mid[3] = {0.5 * (v1[0] + v2[0]), 0.5 * (v1[1] + v2[1]), 0.5 * (v1[2] + v2[2])};
    double im = 1.0 / __builtin_sqrt(dot0(mid, mid));
    mid[0] = im * mid[0]; mid[1] = im * mid[1]; mid[2] = im * mid[2];
    double vC[3] = {X[vs[2]], Y[vs[2]], Z[vs[2]]};
    double determinant = vC[0] * (v1[1] * v2[2] - v1[2] * v2[1])
                       + vC[1] * (v1[2] * v2[0] - v1[0] * v2[2])
                       + vC[2] * (v1[0] * v2[1] - v1[1] * v2[0]);
    determinant = determinant > 0 ? 1. : -1.;
    rodrigues(v1, mid, -0.5 * determinant * SVM_PI);
    rodrigues(v2, mid, -0.5 * determinant * SVM_PI);
    double diff[3] = {0.5 * (v1[0] - v2[0]), 0.5 * (v1[1] - v2[1]), 0.5 * (v1[2] - v2[2])};
    for (int k = 0; k < 3; ++k) { v1[k] = v1[k] + diff[k]; v2[k] = v2[k] - diff[k]; }
    double i1 = 1.0 / __builtin_sqrt(dot0(v1, v1));
    double i2 = 1.0 / __builtin_sqrt(dot0(v2, v2));
    for (int k = 0; k < 3; ++k) { v1[k] = R * (i1 * v1[k]); v2[k] = R * (i2 * v2[k]); }
    X[vertex1] = v1[0]; Y[vertex1] = v1[1]; Z[vertex1] = v1[2];
    X[vertex2] = v2[0]; Y[vertex2] = v2[1]; Z[vertex2] = v2[2];
    for (int vert = 0; vert < 3; ++vert) {
        if (vcnref(d, vertex1, vert) == cs[2]) vcnref(d, vertex1, vert) = cs[3];
        if (vcnref(d, vertex2, vert) == cs[0]) vcnref(d, vertex2, vert) = cs[1];
        if (vnref(d, vs[1], vert) == vertex1) vnref(d, vs[1], vert) = vertex2;
        if (vnref(d, vs[2], vert) == vertex2) vnref(d, vs[2], vert) = vertex1;
        if (vnref(d, vertex1, vert) == vs[1]) vnref(d, vertex1, vert) = vs[2];
        if (vnref(d, vertex2, vert) == vs[2]) vnref(d, vertex2, vert) = vs[1];
    }
    // cell i loses vertex2
    int cneigh = d.cnn[cs[0]];
    int cidx = 0;
    for (int cc = 0; cc < cneigh - 1; ++cc) {
        if (ringref(d, cs[0], cc) == vertex2) cidx += 1;
        ringref(d, cs[0], cc) = ringref(d, cs[0], cidx);
        cidx += 1;
    }
    d.cnn[cs[0]] -= 1;
    // cell j gains vertex2 between vertex1 and b
    cneigh = d.cnn[cs[1]];
    cidx = 0;
    for (int cc = 0; cc < cneigh; ++cc) {
        int e = ringref(d, cs[1], cc);
        tmp[cidx++] = e;
        if (e == vertex1) tmp[cidx++] = vertex2;
    }
    for (int cc = 0; cc < cneigh + 1; ++cc) ringref(d, cs[1], cc) = tmp[cc];
    d.cnn[cs[1]] += 1;
    // cell k loses vertex1
    cneigh = d.cnn[cs[2]];
    cidx = 0;
    for (int cc = 0; cc < cneigh - 1; ++cc) {
        if (ringref(d, cs[2], cc) == vertex1) cidx += 1;
        ringref(d, cs[2], cc) = ringref(d, cs[2], cidx);
        cidx += 1;
    }
    d.cnn[cs[2]] -= 1;
    // cell l gains vertex1 between vertex2 and a
    cneigh = d.cnn[cs[3]];
    cidx = 0;
    for (int cc = 0; cc < cneigh; ++cc) {
        int e = ringref(d, cs[3], cc);
        tmp[cidx++] = e;
        if (e == vertex2) tmp[cidx++] = vertex1;
    }
    for (int cc = 0; cc < cneigh + 1; ++cc) ringref(d, cs[3], cc) = tmp[cc];
    d.cnn[cs[3]] = cneigh + 1;
    return 0;
}

// Applies the T1 flips of one sphere in the order of the reference's
// sequential loop (vertex1 ascending, slots 0..2, later tests seeing earlier
// flips).  Only vertices that can do anything are visited: those flagged by
// k_t1_detect, plus every vertex > the current one whose position, edges or
// neighbours' positions a flip changed.  Any other vertex's test reads data
// unchanged since detection, so skipping it is exact.  If the dirty-vertex
// heap fills up, the remaining vertices are all visited in order (still exact).
__global__ __launch_bounds__(SVM_TILE) void k_t1_resolve(Dev d, int pcur, int nin, int nout, int dir_mode,
                                                         int advance_step, int fire_mode) {
    __shared__ int heapbuf[SVM_HEAP_CAP];
    __shared__ int moved[SVM_LIST_CAP];
    __shared__ int touched[SVM_LIST_CAP];
    __shared__ int scan[SVM_TILE];
    __shared__ int tmp[72];
    __shared__ int sh_flag, sh_nmoved, sh_ntouched, sh_err;
    const int s = blockIdx.x;
    SysParams* S = d.sys + s;
    const int t = threadIdx.x;
    if (t == 0) {
        bool frozen = fire_mode && S->f_frozen;
        sh_flag = frozen ? 0 : S->t1_flag;
        if (!frozen) S->t1_flag = 0;
        if (fire_mode && !frozen) S->mix_pending = 0;
        sh_nmoved = 0;
        sh_ntouched = 0;
        sh_err = 0;
    }
    __syncthreads();
    if (!sh_flag) {
        if (t == 0 && advance_step) S->step += 1;
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
        const double R = S->R, thr = S->thr;
        double* X = d.px[pcur];
        double* Y = d.py[pcur];
        double* Z = d.pz[pcur];
        Heap heap{heapbuf, 0};
        int ci = 0, last = -1, nm = 0, nt = 0, err = 0;
        bool full_scan = false;
        long long flips = 0, skipped = 0;
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
            const int gu = vo + u;
            double v1[3] = {X[gu], Y[gu], Z[gu]};
            for (int vv = 0; vv < 3; ++vv) {
                int vertex2 = vnref(d, gu, vv);
                if (gu < vertex2) {
                    double q[3] = {X[vertex2], Y[vertex2], Z[vertex2]};
                    double arcLength = geodesic(v1, q, pnorm(v1), pnorm(q), R);
                    if (arcLength < thr) {
                        int cs[4], vs[4] = {0, 0, 0, 0};
                        if (!cell_vertex_set(d, gu, vertex2, cs, vs)) {
                            ++skipped;
                            continue;
                        }
                        bool inside = true;
                        for (int k = 0; k < 4; ++k) {
                            inside &= cs[k] >= S->co && cs[k] < S->co + S->nc;
                            inside &= vs[k] >= vo && vs[k] < vo + nv;
                        }
                        if (!inside) { err = 5; break; }
                        if (d.cnn[cs[0]] == 3 || d.cnn[cs[2]] == 3) continue;
                        int r = t1_flip(d, pcur, gu, vertex2, cs, vs, v1, R, tmp);
                        if (r) { err = r; break; }
                        ++flips;
                        if (nm + 2 <= SVM_LIST_CAP) { moved[nm++] = gu; moved[nm++] = vertex2; }
                        else nm = SVM_LIST_CAP + 1;
                        for (int k = 0; k < 4; ++k) {
                            if (nt < SVM_LIST_CAP) touched[nt++] = cs[k];
                            else nt = SVM_LIST_CAP + 1;
                        }
                        if (!full_scan) {
                            int dirty[6] = {u, vertex2 - vo, vs[0] - vo, vs[1] - vo, vs[2] - vo, vs[3] - vo};
                            for (int k = 0; k < 6; ++k) {
                                if (dirty[k] > u && dirty[k] < nv) {
                                    if (heap.n < SVM_HEAP_CAP) heap.push(dirty[k]);
                                    else full_scan = true;
                                }
                            }
                        }
                    }
                }
            }
            if (err) break;
        }
        S->n_t1 += flips;
        S->n_t1_skipped += skipped;
        sh_nmoved = nm;
        sh_ntouched = nt;
        sh_err = err;
        if (err) atomicMax(&S->err, err);
    }
    __threadfence_block();
    __syncthreads();
    // vln for every vertex of every touched cell
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
    // directors of moved vertices (active dynamics)
    if (dir_mode) {
        const int nmv = sh_nmoved;
        double* X = d.px[pcur];
        double* Y = d.py[pcur];
        double* Z = d.pz[pcur];
        if (nmv > SVM_LIST_CAP) {
            for (int x = vo + t; x < vo + nv; x += SVM_TILE) {
                double p[3] = {X[x], Y[x], Z[x]};
                double nI[3] = {d.nx[nin][x], d.ny[nin][x], d.nz[nin][x]}, nO[3];
                director_update(S, x - vo, p, nI, nO);
                d.nx[nout][x] = nO[0]; d.ny[nout][x] = nO[1]; d.nz[nout][x] = nO[2];
            }
        } else {
            for (int i = t; i < nmv; i += SVM_TILE) {
                int x = moved[i];
                double p[3] = {X[x], Y[x], Z[x]};
                double nI[3] = {d.nx[nin][x], d.ny[nin][x], d.nz[nin][x]}, nO[3];
                director_update(S, x - vo, p, nI, nO);
                d.nx[nout][x] = nO[0]; d.ny[nout][x] = nO[1]; d.nz[nout][x] = nO[2];
            }
        }
    }
    __syncthreads();
    if (t == 0 && advance_step) S->step += 1;
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
    if (arrive_last(&S->counter_geo, S->nctiles)) {
        double total = fold_partials(d.cpart + S->ctile0, 1, S->nctiles);
        if (threadIdx.x == 0) S->energy_pot = total;
    }
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
    if (arrive_last(&S->counter_vtx, S->nvtiles)) {
        double total = fold_partials(d.vpart + 3 * S->vtile0, 3, S->nvtiles);
        if (threadIdx.x == 0) S->energy_kin = total;
    }
}

}  // namespace svm
