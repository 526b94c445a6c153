// svm_mlp.h -- the learned warm start for FIRE: a per-cell MLP that maps
// local shape descriptors to a tangent-plane displacement of the cell, with
// its forward pass on bf16 MFMA (FP32 accumulate) and its training (FP32
// backward, Adam) on the GPU.  No counterpart in the reference; FIRE itself
// (energyMinimizerFIRE) is unchanged and remains the fallback.
//
// Network: 12 descriptors (padded to 16) -> 64 -> 64 -> 64 -> 2, SiLU after
// each hidden layer.  Inputs are normalised by per-feature (mu, sigma) before
// the first layer; the output is scaled by out_scale.  Weights are stored as
// an FP32 master vector theta (layout below) and as bf16 copies used by the
// forward pass; activations feeding each layer are rounded to bf16, products
// accumulate in FP32.
//
// MFMA mapping (one wavefront per 16 cells): the weights are the A operand
// (16 output features x K inputs) and the activations the B operand (K
// inputs x 16 cells), with D[4(l/16)+i][l%16] in lane l.  The first layer
// (16 padded inputs) is one v_mfma_f32_16x16x16_bf16 per output block, with
// lane l holding A[l%16][4(l/16)+j] and B[4(l/16)+j][l%16].  The 64-wide
// layers use gfx950's v_mfma_f32_16x16x32_bf16 (K = 32, eight values per lane
// and operand): a lane supplies the four outputs it holds of each of two
// consecutive 16-row blocks of the previous layer, and the weights are loaded
// in that order (mlp_ldw32), so the layers chain in registers with no data
// exchange and half as many instructions as with K = 16.
#pragma once
#include "svm_device.h"

namespace svm {

#define SVM_MLP_IN 12
#define SVM_MLP_INP 16
#define SVM_MLP_H 64
// offsets in theta (floats)
#define SVM_MLP_W1 0                       // [64][16]
#define SVM_MLP_B1 1024                    // [64]
#define SVM_MLP_W2 1088                    // [64][64]
#define SVM_MLP_B2 5184
#define SVM_MLP_W3 5248                    // [64][64]
#define SVM_MLP_B3 9344
#define SVM_MLP_W4 9408                    // [16][64], rows 2..15 zero
#define SVM_MLP_B4 10432                   // [16], entries 2..15 zero
#define SVM_MLP_NP 10448
// per-cell training buffers
#define SVM_MLP_ZS 192                     // pre-activations of the 3 hidden layers
#define SVM_MLP_DS 194                     // deltas: 3 x 64 hidden + 2 output

typedef short mlp_v4s __attribute__((ext_vector_type(4)));
typedef float mlp_v4f __attribute__((ext_vector_type(4)));

struct MlpParams {
    const float* theta;              // FP32 master weights
    const unsigned short* w1;        // bf16 copies
    const unsigned short* w2;
    const unsigned short* w3;
    const unsigned short* w4;
    const float* mu;                 // [16] input normalisation (pads: 0, 1)
    const float* sig;                // [16]
    float out_scale;
};

// round to nearest even (the inputs here are finite)
__device__ __host__ inline unsigned short f2bf(float f) {
    unsigned u;
    __builtin_memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __host__ inline float bf2f(unsigned short b) {
    unsigned u = (unsigned)b << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}
// SiLU and its derivative with the hardware exponential and reciprocal
// (v_exp_f32, v_rcp_f32: about 1 ulp each); the activations are rounded to
// bf16 right after, so an IEEE-exact divide would buy nothing and cost about
// nine instructions per activation.
__device__ inline float sigmoid_fast(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ inline float silu(float z) { return z * sigmoid_fast(z); }
__device__ inline float dsilu(float z) {
    float s = sigmoid_fast(z);
    return s * (1.0f + z * (1.0f - s));
}

__device__ inline mlp_v4s mlp_ldw(const unsigned short* W, int in_dim, int ob, int kb, int lane) {
    const int row = ob * 16 + (lane & 15), k0 = kb * 16 + (lane >> 4) * 4;
    const unsigned short* p = W + (size_t)row * in_dim + k0;
    mlp_v4s r;
    r[0] = (short)p[0]; r[1] = (short)p[1]; r[2] = (short)p[2]; r[3] = (short)p[3];
    return r;
}
// Operands of gfx950's v_mfma_f32_16x16x32_bf16 for a 64-wide layer: K = 32
// inputs per instruction.  A lane of group q holds, from the previous layer,
// outputs q*4 .. q*4+3 of every 16-row block; it gives the instruction those
// of blocks 2*kb2 and 2*kb2+1 (eight values), and the weights are loaded in
// the same order, so the instruction's K slots are a fixed permutation of the
// inputs, the same for both operands.
typedef short mlp_v8s __attribute__((ext_vector_type(8)));
typedef __bf16 mlp_v8bf __attribute__((ext_vector_type(8)));
__device__ inline mlp_v8s mlp_ldw32(const unsigned short* W, int in_dim, int ob, int kb2, int lane) {
    const int row = ob * 16 + (lane & 15), q = lane >> 4;
    const unsigned short* p = W + (size_t)row * in_dim + kb2 * 32 + q * 4;
    mlp_v8s r;
    r[0] = (short)p[0]; r[1] = (short)p[1]; r[2] = (short)p[2]; r[3] = (short)p[3];
    r[4] = (short)p[16]; r[5] = (short)p[17]; r[6] = (short)p[18]; r[7] = (short)p[19];
    return r;
}
__device__ inline mlp_v4f mlp_mfma32(const mlp_v8s& w, const mlp_v4s& lo, const mlp_v4s& hi, mlp_v4f acc) {
    const mlp_v8s b = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(mlp_v8bf, w), __builtin_bit_cast(mlp_v8bf, b),
                                                   acc, 0, 0, 0);
}

// Four floats to the bf16 B operand of the MFMA, round to nearest even (as
// f2bf for finite values), two per v_cvt_pk_bf16_f32.
typedef float mlp_v2f __attribute__((ext_vector_type(2)));
typedef __bf16 mlp_v2bf __attribute__((ext_vector_type(2)));
__device__ inline mlp_v4s mlp_pack(float a, float b, float c, float d) {
    const mlp_v2bf lo = __builtin_convertvector((mlp_v2f){a, b}, mlp_v2bf);
    const mlp_v2bf hi = __builtin_convertvector((mlp_v2f){c, d}, mlp_v2bf);
    struct {
        mlp_v2bf lo, hi;
    } pair = {lo, hi};
    return __builtin_bit_cast(mlp_v4s, pair);
}

// Forward pass over n cells: X [n][16] raw descriptors -> Y [n][2].  With Z
// (and H) non-null, also the hidden pre-activations Z [n][192] and the bf16
// activations fed to layers 2..4 (as floats) H [n][192] for training.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_mlp_forward(MlpParams m, const float* X, int n, float* Y, float* Z,
                                                     float* H) {
    const int lane = threadIdx.x & 63, q = lane >> 4, col = lane & 15;
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6), nwaves = gridDim.x * 4;
    mlp_v4s w1[4];
    mlp_v8s w2[4][2], w3[4][2], w4[2];
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) {
        w1[ob] = mlp_ldw(m.w1, 16, ob, 0, lane);
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) {
            w2[ob][kb] = mlp_ldw32(m.w2, 64, ob, kb, lane);
            w3[ob][kb] = mlp_ldw32(m.w3, 64, ob, kb, lane);
        }
    }
    w4[0] = mlp_ldw32(m.w4, 64, 0, 0, lane);
    w4[1] = mlp_ldw32(m.w4, 64, 0, 1, lane);
    float b1[4][4], b2[4][4], b3[4][4], b4[4], mu[4], is[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int ob = 0; ob < 4; ++ob) {
            b1[ob][i] = m.theta[SVM_MLP_B1 + ob * 16 + q * 4 + i];
            b2[ob][i] = m.theta[SVM_MLP_B2 + ob * 16 + q * 4 + i];
            b3[ob][i] = m.theta[SVM_MLP_B3 + ob * 16 + q * 4 + i];
        }
        b4[i] = m.theta[SVM_MLP_B4 + q * 4 + i];
        mu[i] = m.mu[q * 4 + i];
        is[i] = 1.0f / m.sig[q * 4 + i];
    }
    for (int tile = wave; tile * 16 < n; tile += nwaves) {   // wave-uniform loop
        const int cell = tile * 16 + col;
        const bool valid = cell < n;
        float x[4] = {0.f, 0.f, 0.f, 0.f};
        if (valid) {
            const float* xr = X + (size_t)cell * 16 + q * 4;
#pragma unroll
            for (int j = 0; j < 4; ++j) x[j] = (xr[j] - mu[j]) * is[j];
        }
        const mlp_v4s h0 = mlp_pack(x[0], x[1], x[2], x[3]);
        mlp_v4s h1[4], h2[4], h3[4];
#pragma unroll
        for (int ob = 0; ob < 4; ++ob) {
            mlp_v4f acc = {0.f, 0.f, 0.f, 0.f};
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w1[ob], h0, acc, 0, 0, 0);
            float z[4], a[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) { z[i] = acc[i] + b1[ob][i]; a[i] = silu(z[i]); }
            h1[ob] = mlp_pack(a[0], a[1], a[2], a[3]);
            if (Z && valid)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    Z[(size_t)cell * SVM_MLP_ZS + ob * 16 + q * 4 + i] = z[i];
                    H[(size_t)cell * SVM_MLP_ZS + ob * 16 + q * 4 + i] = bf2f((unsigned short)h1[ob][i]);
                }
        }
#pragma unroll
        for (int ob = 0; ob < 4; ++ob) {
            mlp_v4f acc = {0.f, 0.f, 0.f, 0.f};
            acc = mlp_mfma32(w2[ob][0], h1[0], h1[1], acc);
            acc = mlp_mfma32(w2[ob][1], h1[2], h1[3], acc);
            float z[4], a[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) { z[i] = acc[i] + b2[ob][i]; a[i] = silu(z[i]); }
            h2[ob] = mlp_pack(a[0], a[1], a[2], a[3]);
            if (Z && valid)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    Z[(size_t)cell * SVM_MLP_ZS + 64 + ob * 16 + q * 4 + i] = z[i];
                    H[(size_t)cell * SVM_MLP_ZS + 64 + ob * 16 + q * 4 + i] = bf2f((unsigned short)h2[ob][i]);
                }
        }
#pragma unroll
        for (int ob = 0; ob < 4; ++ob) {
            mlp_v4f acc = {0.f, 0.f, 0.f, 0.f};
            acc = mlp_mfma32(w3[ob][0], h2[0], h2[1], acc);
            acc = mlp_mfma32(w3[ob][1], h2[2], h2[3], acc);
            float z[4], a[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) { z[i] = acc[i] + b3[ob][i]; a[i] = silu(z[i]); }
            h3[ob] = mlp_pack(a[0], a[1], a[2], a[3]);
            if (Z && valid)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    Z[(size_t)cell * SVM_MLP_ZS + 128 + ob * 16 + q * 4 + i] = z[i];
                    H[(size_t)cell * SVM_MLP_ZS + 128 + ob * 16 + q * 4 + i] = bf2f((unsigned short)h3[ob][i]);
                }
        }
        mlp_v4f acc = {0.f, 0.f, 0.f, 0.f};
        acc = mlp_mfma32(w4[0], h3[0], h3[1], acc);
        acc = mlp_mfma32(w4[1], h3[2], h3[3], acc);
        if (q == 0 && valid) {
            Y[(size_t)cell * 2] = m.out_scale * (acc[0] + b4[0]);
            Y[(size_t)cell * 2 + 1] = m.out_scale * (acc[1] + b4[1]);
        }
    }
}

// Backward, thread per cell: loss L = (1/n_total) sum |Y - T|^2 over the
// batch; deltas D [n][194]: dL/dz of hidden layers 1..3 and dL/d(layer-4
// output before scaling).  Per-block partial sums of |Y - T|^2 in lossp.
__global__ __launch_bounds__(256) void k_mlp_backward(MlpParams m, const float* Z, const float* Y, const float* T,
                                                      int n, float inv_total, float* D, float* lossp) {
    __shared__ float sw2[64][65], sw3[64][65], sw4[2][64];
    __shared__ float red[256];
    for (int e = threadIdx.x; e < 4096; e += 256) {
        sw2[e >> 6][e & 63] = bf2f(m.w2[e]);
        sw3[e >> 6][e & 63] = bf2f(m.w3[e]);
    }
    if (threadIdx.x < 128) sw4[threadIdx.x >> 6][threadIdx.x & 63] = bf2f(m.w4[threadIdx.x]);
    __syncthreads();
    const int c = blockIdx.x * 256 + threadIdx.x;
    float l2 = 0.0f;
    if (c < n) {
        const float* z = Z + (size_t)c * SVM_MLP_ZS;
        float* dc = D + (size_t)c * SVM_MLP_DS;
        float g[2];
        for (int o = 0; o < 2; ++o) {
            float e = Y[(size_t)c * 2 + o] - T[(size_t)c * 2 + o];
            l2 += e * e;
            g[o] = 2.0f * e * inv_total * m.out_scale;
            dc[192 + o] = g[o];
        }
        // deltas go to D as they are made; each matvec runs over the output
        // index with 64 accumulators in registers (LDS reads broadcast)
        for (int i = 0; i < 64; ++i) dc[128 + i] = (sw4[0][i] * g[0] + sw4[1][i] * g[1]) * dsilu(z[128 + i]);
        float acc[64];
#pragma unroll
        for (int i = 0; i < 64; ++i) acc[i] = 0.0f;
        for (int o = 0; o < 64; ++o) {
            const float dd = dc[128 + o];
#pragma unroll
            for (int i = 0; i < 64; ++i) acc[i] += sw3[o][i] * dd;
        }
#pragma unroll
        for (int i = 0; i < 64; ++i) {
            dc[64 + i] = acc[i] * dsilu(z[64 + i]);
            acc[i] = 0.0f;
        }
        for (int o = 0; o < 64; ++o) {
            const float dd = dc[64 + o];
#pragma unroll
            for (int i = 0; i < 64; ++i) acc[i] += sw2[o][i] * dd;
        }
#pragma unroll
        for (int i = 0; i < 64; ++i) dc[i] = acc[i] * dsilu(z[i]);
    }
    red[threadIdx.x] = l2;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) lossp[blockIdx.x] = red[0];
}

// Weight gradients: per chunk of 64 cells, every parameter's partial sum
// (deltas x the bf16 activations feeding that layer), into partial[chunk][NP].
__global__ __launch_bounds__(256) void k_mlp_wgrad(MlpParams m, const float* X, const float* H, const float* D,
                                                   int n, float* partial) {
    __shared__ float sd[64][SVM_MLP_DS + 1];
    __shared__ float sa[64][SVM_MLP_ZS + 16 + 1];   // [0,16): layer-1 inputs, [16,208): H
    const int c0 = blockIdx.x * 64;
    for (int e = threadIdx.x; e < 64 * SVM_MLP_DS; e += 256) {
        int r = e / SVM_MLP_DS, k = e % SVM_MLP_DS;
        sd[r][k] = c0 + r < n ? D[(size_t)(c0 + r) * SVM_MLP_DS + k] : 0.0f;
    }
    for (int e = threadIdx.x; e < 64 * SVM_MLP_ZS; e += 256) {
        int r = e / SVM_MLP_ZS, k = e % SVM_MLP_ZS;
        sa[r][16 + k] = c0 + r < n ? H[(size_t)(c0 + r) * SVM_MLP_ZS + k] : 0.0f;
    }
    for (int e = threadIdx.x; e < 64 * 16; e += 256) {
        int r = e >> 4, k = e & 15;
        float x = 0.0f;
        if (c0 + r < n) x = bf2f(f2bf((X[(size_t)(c0 + r) * 16 + k] - m.mu[k]) / m.sig[k]));
        sa[r][k] = x;
    }
    __syncthreads();
    for (int p = threadIdx.x; p < SVM_MLP_NP; p += 256) {
        int drow, acol;   // delta column, activation column (-1: bias)
        if (p < SVM_MLP_B1) { drow = p / 16; acol = p % 16; }
        else if (p < SVM_MLP_W2) { drow = p - SVM_MLP_B1; acol = -1; }
        else if (p < SVM_MLP_B2) { int q = p - SVM_MLP_W2; drow = 64 + q / 64; acol = 16 + q % 64; }
        else if (p < SVM_MLP_W3) { drow = 64 + p - SVM_MLP_B2; acol = -1; }
        else if (p < SVM_MLP_B3) { int q = p - SVM_MLP_W3; drow = 128 + q / 64; acol = 16 + 64 + q % 64; }
        else if (p < SVM_MLP_W4) { drow = 128 + p - SVM_MLP_B3; acol = -1; }
        else if (p < SVM_MLP_B4) {
            int q = p - SVM_MLP_W4;
            drow = q / 64 < 2 ? 192 + q / 64 : -1;
            acol = 16 + 128 + q % 64;
        } else {
            int q = p - SVM_MLP_B4;
            drow = q < 2 ? 192 + q : -1;
            acol = -1;
        }
        float s = 0.0f;
        if (drow >= 0)
            for (int r = 0; r < 64; ++r) s += sd[r][drow] * (acol < 0 ? 1.0f : sa[r][acol]);
        partial[(size_t)blockIdx.x * SVM_MLP_NP + p] = s;
    }
}

// grad[p] = sum over chunks of partial[chunk][p], in chunk order.
__global__ __launch_bounds__(256) void k_mlp_reduce(const float* partial, int nchunks, float* grad) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= SVM_MLP_NP) return;
    float s = 0.0f;
    for (int k = 0; k < nchunks; ++k) s += partial[(size_t)k * SVM_MLP_NP + p];
    grad[p] = s;
}

// Adam (Kingma & Ba 2015) on theta, then the bf16 copies.
__global__ __launch_bounds__(256) void k_mlp_adam(float* theta, const float* grad, float* mom, float* vel, float lr,
                                                  float beta1, float beta2, float eps, float c1, float c2) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= SVM_MLP_NP) return;
    const float g = grad[p];
    const float mt = beta1 * mom[p] + (1.0f - beta1) * g;
    const float vt = beta2 * vel[p] + (1.0f - beta2) * g * g;
    mom[p] = mt;
    vel[p] = vt;
    theta[p] -= lr * (mt / c1) / (__builtin_sqrtf(vt / c2) + eps);
}

__global__ __launch_bounds__(256) void k_mlp_to_bf16(const float* theta, unsigned short* w1, unsigned short* w2,
                                                     unsigned short* w3, unsigned short* w4) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p < 1024) w1[p] = f2bf(theta[SVM_MLP_W1 + p]);
    if (p < 4096) {
        w2[p] = f2bf(theta[SVM_MLP_W2 + p]);
        w3[p] = f2bf(theta[SVM_MLP_W3 + p]);
    }
    if (p < 1024) w4[p] = f2bf(theta[SVM_MLP_W4 + p]);
}

// ------------------------------------------------------------ descriptors --
// Per cell (after k_geometry and a force pass): the 12 descriptors and the
// cell's tangent frame (e1 along the tangent part of the cell force, the mean
// of its vertices' forces; e2 = n x e1).
//   0 A - A0            1 P - P0            2 P / sqrt(A)        3 n - 6
//   4 |F_cell|          5 mean |F_v|        6 <A_nb> - A         7 std(A_nb)
//   8, 9 sum (A_nb - A)(d_nb . e1, e2) / n  (d_nb: unit tangent toward C_nb)
//   10 <q_nb> - q       11 <P_nb - P0_nb>
__global__ __launch_bounds__(SVM_TILE) void k_cell_descriptors(Dev d, int pin, float* X, float* frame) {
    const int c = blockIdx.x * SVM_TILE + threadIdx.x;
    if (c >= d.Cs) return;
    const SysParams* S = d.sys + d.ctile_sys[c / SVM_TILE];
    float* x = X + (size_t)c * 16;
    float* fr = frame + (size_t)c * 6;
    if (c - S->co >= S->nc) {
        for (int k = 0; k < 16; ++k) x[k] = 0.0f;
        for (int k = 0; k < 6; ++k) fr[k] = 0.0f;
        return;
    }
    const int n = d.cnn[c];
    const double A = d.A[c], P = d.P[c];
    const double C[3] = {d.cx[c], d.cy[c], d.cz[c]};
    const double cn = svm::sqrt_rn(C[0] * C[0] + C[1] * C[1] + C[2] * C[2]);
    const double nh[3] = {C[0] / cn, C[1] / cn, C[2] / cn};
    double F[3] = {0.0, 0.0, 0.0}, fabs_sum = 0.0;
    for (int k = 0; k < n; ++k) {
        int v = d.ring[(size_t)k * d.Cs + c];
        double f0 = d.fx[v], f1 = d.fy[v], f2 = d.fz[v];
        F[0] += f0; F[1] += f1; F[2] += f2;
        fabs_sum += svm::sqrt_rn(f0 * f0 + f1 * f1 + f2 * f2);
    }
    for (int k = 0; k < 3; ++k) F[k] /= n;
    double fn = F[0] * nh[0] + F[1] * nh[1] + F[2] * nh[2];
    double Ft[3] = {F[0] - fn * nh[0], F[1] - fn * nh[1], F[2] - fn * nh[2]};
    double fmag = svm::sqrt_rn(Ft[0] * Ft[0] + Ft[1] * Ft[1] + Ft[2] * Ft[2]);
    double e1[3];
    if (fmag > 1e-300) {
        for (int k = 0; k < 3; ++k) e1[k] = Ft[k] / fmag;
    } else {   // no force: any tangent direction
        double a[3] = {__builtin_fabs(nh[0]) < 0.9 ? 1.0 : 0.0, __builtin_fabs(nh[0]) < 0.9 ? 0.0 : 1.0, 0.0};
        double t[3] = {a[1] * nh[2] - a[2] * nh[1], a[2] * nh[0] - a[0] * nh[2], a[0] * nh[1] - a[1] * nh[0]};
        double tn = svm::sqrt_rn(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
        for (int k = 0; k < 3; ++k) e1[k] = t[k] / tn;
    }
    const double e2[3] = {nh[1] * e1[2] - nh[2] * e1[1], nh[2] * e1[0] - nh[0] * e1[2], nh[0] * e1[1] - nh[1] * e1[0]};
    const double q = P / svm::sqrt_rn(A > 1e-12 ? A : 1e-12);
    double sA = 0.0, sA2 = 0.0, dip1 = 0.0, dip2 = 0.0, sq = 0.0, sdp = 0.0;
    for (int k = 0; k < n; ++k) {
        int v = d.ring[(size_t)k * d.Cs + c];
        int w = d.ring[(size_t)((k + 1) % n) * d.Cs + c];
        int4 cv = d.vcn[v], cw = d.vcn[w];
        int a3[3] = {cv.x, cv.y, cv.z}, b3[3] = {cw.x, cw.y, cw.z};
        int nb = c;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                if (a3[i] == b3[j] && a3[i] != c) nb = a3[i];
        const double An = d.A[nb], Pn = d.P[nb];
        sA += An;
        sA2 += An * An;
        sq += Pn / svm::sqrt_rn(An > 1e-12 ? An : 1e-12);
        sdp += Pn - d.p0[nb];
        double dv[3] = {d.cx[nb] - C[0], d.cy[nb] - C[1], d.cz[nb] - C[2]};
        double dn = dv[0] * nh[0] + dv[1] * nh[1] + dv[2] * nh[2];
        for (int k2 = 0; k2 < 3; ++k2) dv[k2] -= dn * nh[k2];
        double dl = svm::sqrt_rn(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
        if (dl > 0.0) {
            double u1 = (dv[0] * e1[0] + dv[1] * e1[1] + dv[2] * e1[2]) / dl;
            double u2 = (dv[0] * e2[0] + dv[1] * e2[1] + dv[2] * e2[2]) / dl;
            dip1 += (An - A) * u1;
            dip2 += (An - A) * u2;
        }
    }
    const double mA = sA / n;
    const double var = sA2 / n - mA * mA;
    x[0] = (float)(A - d.a0[c]);
    x[1] = (float)(P - d.p0[c]);
    x[2] = (float)q;
    x[3] = (float)(n - 6);
    x[4] = (float)fmag;
    x[5] = (float)(fabs_sum / n);
    x[6] = (float)(mA - A);
    x[7] = (float)svm::sqrt_rn(var > 0.0 ? var : 0.0);
    x[8] = (float)(dip1 / n);
    x[9] = (float)(dip2 / n);
    x[10] = (float)(sq / n - q);
    x[11] = (float)(sdp / n);
    for (int k = 12; k < 16; ++k) x[k] = 0.0f;
    for (int k = 0; k < 3; ++k) { fr[k] = (float)e1[k]; fr[3 + k] = (float)e2[k]; }
}

// Warm-start displacement of each vertex: the mean over its three cells of
// the predicted cell displacement y0 e1 + y1 e2, times `scale`, into d.disp;
// its length is capped at `cap` times the vertex's shortest edge (no cap for
// cap = 0); with a cap below 1/3 no move can fold a cell (vertices moving by
// less than a third of their edges keep every triangle's orientation).
__global__ __launch_bounds__(SVM_TILE) void k_warm_disp(Dev d, int pin, const float* Y, const float* frame,
                                                        double scale, double cap) {
    const int v = blockIdx.x * SVM_TILE + threadIdx.x;
    if (v >= d.Vs) return;
    const SysParams* S = d.sys + d.vtile_sys[v / SVM_TILE];
    double out[3] = {0.0, 0.0, 0.0};
    // a sphere whose minimiser has converged keeps its positions
    if (v - S->vo < S->nv && !minimiser_converged(S)) {
        int4 cl = d.vcn[v];
        int c3[3] = {cl.x, cl.y, cl.z};
        for (int k = 0; k < 3; ++k) {
            const float* fr = frame + (size_t)c3[k] * 6;
            const double y0 = Y[(size_t)c3[k] * 2], y1 = Y[(size_t)c3[k] * 2 + 1];
            for (int j = 0; j < 3; ++j) out[j] += y0 * fr[j] + y1 * fr[3 + j];
        }
        for (int j = 0; j < 3; ++j) out[j] *= scale / 3.0;
        const double* X = d.px[pin];
        const double* Yp = d.py[pin];
        const double* Zp = d.pz[pin];
        int4 nb = d.vn[v];
        int n3[3] = {nb.x, nb.y, nb.z};
        double emin = 1e300;
        for (int k = 0; k < 3; ++k) {
            const double dx = X[n3[k]] - X[v], dy = Yp[n3[k]] - Yp[v], dz = Zp[n3[k]] - Zp[v];
            emin = fmin(emin, svm::sqrt_rn(dx * dx + dy * dy + dz * dz));
        }
        const double len = svm::sqrt_rn(out[0] * out[0] + out[1] * out[1] + out[2] * out[2]);
        const double lim = cap * emin;             // cap 0: no limit
        if (cap > 0.0 && len > lim && len > 0.0)
            for (int j = 0; j < 3; ++j) out[j] *= lim / len;
    }
    for (int j = 0; j < 3; ++j) d.disp[3 * (size_t)v + j] = out[j];
}

}  // namespace svm
