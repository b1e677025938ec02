// Learned warm start for the FIRE relaxation: a small 3D-conv network that
// maps the total-force patch around a probe's start position to its relaxed
// displacement, trained on the GPU by the kernels below.  No counterpart in
// the reference (it is an addition in front of relaxTipStroke /
// relaxStrokesTilted); FIRE always finishes every position afterwards.
//
// Network (valid convolutions, SiLU after layers 1 and 2):
//   patch 5x5x3 nodes x 3 force components (RMS-normalised, bf16)
//   conv1 3 -> 32, kernel 3x3x3  -> 3x3x1 x 32   (GEMM K = 81, padded to 96)
//   conv2 32 -> 32, kernel 3x3x1 -> 1x1x1 x 32   (GEMM K = 288)
//   conv3 32 -> 3 (1x1x1)                        (VALU, FP32)
// Parameters (FP32 master, 11,971 values): W1[32][81] b1[32] W2[32][288]
// b2[32] W3[3][32] b3[3]; bf16 copies W1b[32][96], W2b[32][288],
// W2bT[288][32] feed the matrix cores.
//
// Matrix products run on v_mfma_f32_16x16x32_bf16 (FP32 accumulation).
// Fragment layout (checked on the device): lane l holds A[l%16][8*(l/16)+e]
// and B[8*(l/16)+e][l%16] (e = 0..7), and D[4*(l/16)+i][l%16] (i = 0..3).
//
// Training step on a minibatch of S samples (S a multiple of 64):
//   fwd1  Z1 = im2col(P) W1^T + b1, H1 = silu(Z1)         (per 64 samples x 9 positions)
//   fwd2  Z2 = H1 W2^T + b2, H2 = silu(Z2)
//   head  out = bf(H2) bf(W3)^T + b3, loss, dW3, db3, dZ2, db2
//   bwd2  dW2 += dZ2^T H1                                  (split over S, atomics)
//   bwd1a dZ1 = (dZ2 W2) * silu'(Z1), db1
//   bwd1b dW1 += dZ1^T A1                                  (split over 9S, atomics)
//   adam  update, refresh the bf16 copies, clear the gradient

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "pp_common.hpp"

namespace pp {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kN1 = 32, kK1 = 81, kK1p = 96, kN2 = 32, kK2 = 288, kN3 = 3;
constexpr int kPatch = 225, kPatchStride = 232;  // bf16 per patch row (16-byte aligned)
constexpr int kOffW1 = 0, kOffB1 = kN1 * kK1, kOffW2 = kOffB1 + kN1, kOffB2 = kOffW2 + kN2 * kK2;
constexpr int kOffW3 = kOffB2 + kN2, kOffB3 = kOffW3 + kN3 * kN2, kNParam = kOffB3 + kN3;
static_assert(kNParam == 11971, "parameter count");

__device__ __forceinline__ unsigned short f2bf(float x)
{
    unsigned u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float((unsigned)b << 16); }
__device__ __forceinline__ float bfr(float x) { return bf2f(f2bf(x)); }

__device__ __forceinline__ float silu(float z) { return z / (1.f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z)
{
    const float s = 1.f / (1.f + __expf(-z));
    return s * (1.f + z * (1.f - s));
}

__device__ __forceinline__ bf16x8 ld8(const unsigned short *p)
{
    return __builtin_bit_cast(bf16x8, *(const uint4 *)p);
}
__device__ __forceinline__ bf16x8 pack8(const unsigned short v[8])
{
    uint4 u;
    u.x = v[0] | ((unsigned)v[1] << 16);
    u.y = v[2] | ((unsigned)v[3] << 16);
    u.z = v[4] | ((unsigned)v[5] << 16);
    u.w = v[6] | ((unsigned)v[7] << 16);
    return __builtin_bit_cast(bf16x8, u);
}
__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// offset of conv1 tap k = ((dx*3 + dy)*3 + dz)*3 + c inside a patch row,
// relative to the window origin; -1 for the padding k >= 81
__constant__ short kTapOff[kK1p] = {
#define T3(dx, dy) \
    (short)(((dx)*5 + (dy)) * 9 + 0), (short)(((dx)*5 + (dy)) * 9 + 1), (short)(((dx)*5 + (dy)) * 9 + 2), \
    (short)(((dx)*5 + (dy)) * 9 + 3), (short)(((dx)*5 + (dy)) * 9 + 4), (short)(((dx)*5 + (dy)) * 9 + 5), \
    (short)(((dx)*5 + (dy)) * 9 + 6), (short)(((dx)*5 + (dy)) * 9 + 7), (short)(((dx)*5 + (dy)) * 9 + 8)
    T3(0, 0), T3(0, 1), T3(0, 2), T3(1, 0), T3(1, 1), T3(1, 2), T3(2, 0), T3(2, 1), T3(2, 2),
#undef T3
    -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};

struct Net {
    const float *theta;
    const unsigned short *W1b;   // [32][96]
    const unsigned short *W2b;   // [32][288]
    const unsigned short *W2bT;  // [288][32]
};

// ---------------------------------------------------------------- forward

// conv1 for one 64-sample tile and one output position p: Z1, H1 (and the
// transposed copies the backward pass reads).  One wave per (tile, p).
__global__ __launch_bounds__(64) void ws_fwd1_kernel(int S, const unsigned short *__restrict__ P, Net net,
                                                      float *__restrict__ Z1, unsigned short *__restrict__ H1,
                                                      unsigned short *__restrict__ H1T, unsigned short *__restrict__ A1T)
{
    const int l = threadIdx.x;
    const int tile = blockIdx.x, p = blockIdx.y;
    const int ox = p / 3, oy = p % 3;
    const int poff = (ox * 5 + oy) * 9;
    f32x4 acc[4][2];
#pragma unroll
    for (int r = 0; r < 4; r++)
        for (int t = 0; t < 2; t++) acc[r][t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const int k0 = 32 * q + 8 * (l >> 4);
        bf16x8 b[2];
#pragma unroll
        for (int t = 0; t < 2; t++) b[t] = ld8(net.W1b + (16 * t + (l & 15)) * kK1p + k0);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int s = 64 * tile + 16 * r + (l & 15);
            unsigned short v[8];
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int off = kTapOff[k0 + e];
                v[e] = off >= 0 ? P[(long)s * kPatchStride + poff + off] : (unsigned short)0;
                if (A1T) A1T[(long)(k0 + e) * (9L * S) + (long)s * 9 + p] = v[e];
            }
            const bf16x8 a = pack8(v);
#pragma unroll
            for (int t = 0; t < 2; t++) acc[r][t] = mfma(a, b[t], acc[r][t]);
        }
    }
#pragma unroll
    for (int t = 0; t < 2; t++) {
        const int j = 16 * t + (l & 15);
        const float bj = net.theta[kOffB1 + j];
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int s = 64 * tile + 16 * r + 4 * (l >> 4) + i;
                const float z = acc[r][t][i] + bj;
                const unsigned short h = f2bf(silu(z));
                const int k2 = p * 32 + j;
                if (Z1) Z1[(long)s * kK2 + k2] = z;
                H1[(long)s * kK2 + k2] = h;
                if (H1T) H1T[(long)k2 * S + s] = h;
            }
    }
}

// conv2 for one 64-sample tile: Z2, H2 (FP32).
__global__ __launch_bounds__(64) void ws_fwd2_kernel(int S, const unsigned short *__restrict__ H1, Net net,
                                                      float *__restrict__ Z2, float *__restrict__ H2)
{
    const int l = threadIdx.x;
    const int tile = blockIdx.x;
    f32x4 acc[4][2];
#pragma unroll
    for (int r = 0; r < 4; r++)
        for (int t = 0; t < 2; t++) acc[r][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int q = 0; q < kK2 / 32; q++) {
        const int k0 = 32 * q + 8 * (l >> 4);
        bf16x8 b[2];
#pragma unroll
        for (int t = 0; t < 2; t++) b[t] = ld8(net.W2b + (16 * t + (l & 15)) * kK2 + k0);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int s = 64 * tile + 16 * r + (l & 15);
            const bf16x8 a = ld8(H1 + (long)s * kK2 + k0);
#pragma unroll
            for (int t = 0; t < 2; t++) acc[r][t] = mfma(a, b[t], acc[r][t]);
        }
    }
#pragma unroll
    for (int t = 0; t < 2; t++) {
        const int j = 16 * t + (l & 15);
        const float bj = net.theta[kOffB2 + j];
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int s = 64 * tile + 16 * r + 4 * (l >> 4) + i;
                const float z = acc[r][t][i] + bj;
                Z2[(long)s * kN2 + j] = z;
                H2[(long)s * kN2 + j] = silu(z);
            }
    }
}

__device__ __forceinline__ float wave_sum(float v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Layer 3, loss and the head of the backward pass; one thread per sample.
// grad / loss are accumulated with one atomic per wave per value.
// Samples s >= nvalid are padding: they produce outputs but no gradient.
__global__ __launch_bounds__(256) void ws_head_kernel(int S, int nvalid, float inv_n, const float *__restrict__ Z2,
                                                      const float *__restrict__ H2, Net net,
                                                      const float *__restrict__ T, float *__restrict__ out,
                                                      float *__restrict__ grad, float *__restrict__ loss,
                                                      unsigned short *__restrict__ dZ2, unsigned short *__restrict__ dZ2T)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = s < S;
    const bool valid = s < nvalid;
    float h[kN2], o[kN3];
    const float *W3 = net.theta + kOffW3;
#pragma unroll
    for (int j = 0; j < kN2; j++) h[j] = live ? bfr(H2[(long)s * kN2 + j]) : 0.f;
#pragma unroll
    for (int m = 0; m < kN3; m++) {
        float a = net.theta[kOffB3 + m];
#pragma unroll
        for (int j = 0; j < kN2; j++) a = fmaf(h[j], bfr(W3[m * kN2 + j]), a);
        o[m] = a;
        if (out && live) out[(long)s * kN3 + m] = a;
    }
    if (!grad) return;
    float dy[kN3], lsum = 0.f;
#pragma unroll
    for (int m = 0; m < kN3; m++) {
        const float d = valid ? o[m] - T[(long)s * kN3 + m] : 0.f;
        lsum += d * d;
        dy[m] = 2.f * d * inv_n;
    }
    const int lane = threadIdx.x & 63;
    // dW3, db3
#pragma unroll
    for (int m = 0; m < kN3; m++) {
#pragma unroll
        for (int j = 0; j < kN2; j++) {
            const float g = wave_sum(dy[m] * h[j]);
            if (lane == 0) atomicAdd(&grad[kOffW3 + m * kN2 + j], g);
        }
        const float g = wave_sum(dy[m]);
        if (lane == 0) atomicAdd(&grad[kOffB3 + m], g);
    }
    const float ls = wave_sum(lsum * inv_n);
    if (lane == 0) atomicAdd(loss, ls);
    // dZ2 = (dY W3) * silu'(Z2), db2
#pragma unroll
    for (int j = 0; j < kN2; j++) {
        float dh = 0.f;
#pragma unroll
        for (int m = 0; m < kN3; m++) dh = fmaf(dy[m], bfr(W3[m * kN2 + j]), dh);
        const float dz = live ? dh * dsilu(Z2[(long)s * kN2 + j]) : 0.f;
        const float g = wave_sum(dz);
        if (lane == 0) atomicAdd(&grad[kOffB2 + j], g);
        if (live) {
            const unsigned short b = f2bf(dz);
            dZ2[(long)s * kN2 + j] = b;
            dZ2T[(long)j * S + s] = b;
        }
    }
}

// dW2[j][k2] += sum_s dZ2[s][j] H1[s][k2]: M = 32, N = 96 per block column,
// K = ks samples per block.
__global__ __launch_bounds__(64) void ws_bwd2_kernel(int S, int ks, const unsigned short *__restrict__ dZ2T,
                                                      const unsigned short *__restrict__ H1T, float *__restrict__ grad)
{
    const int l = threadIdx.x;
    const int nblk = blockIdx.x;  // 0..2: columns 96*nblk ..
    const int s0 = blockIdx.y * ks;
    f32x4 acc[2][6];
#pragma unroll
    for (int m = 0; m < 2; m++)
        for (int t = 0; t < 6; t++) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int q = 0; q < ks / 32; q++) {
        const int s = s0 + 32 * q + 8 * (l >> 4);
        bf16x8 a[2];
#pragma unroll
        for (int m = 0; m < 2; m++) a[m] = ld8(dZ2T + (long)(16 * m + (l & 15)) * S + s);
#pragma unroll
        for (int t = 0; t < 6; t++) {
            const int k2 = 96 * nblk + 16 * t + (l & 15);
            const bf16x8 b = ld8(H1T + (long)k2 * S + s);
#pragma unroll
            for (int m = 0; m < 2; m++) acc[m][t] = mfma(a[m], b, acc[m][t]);
        }
    }
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int t = 0; t < 6; t++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int j = 16 * m + 4 * (l >> 4) + i;
                const int k2 = 96 * nblk + 16 * t + (l & 15);
                atomicAdd(&grad[kOffW2 + j * kK2 + k2], acc[m][t][i]);
            }
}

// dA2 = dZ2 W2 (M = 64 samples, N = 288, K = 32), dZ1 = dA2 * silu'(Z1),
// written transposed as dZ1T[c][s*9 + p]; db1.
__global__ __launch_bounds__(64) void ws_bwd1a_kernel(int S, const unsigned short *__restrict__ dZ2, Net net,
                                                       const float *__restrict__ Z1, unsigned short *__restrict__ dZ1T,
                                                       float *__restrict__ grad)
{
    const int l = threadIdx.x;
    const int tile = blockIdx.x;
    bf16x8 a[4];
#pragma unroll
    for (int r = 0; r < 4; r++) a[r] = ld8(dZ2 + (long)(64 * tile + 16 * r + (l & 15)) * kN2 + 8 * (l >> 4));
    float db[2] = {0.f, 0.f};
    for (int t = 0; t < kK2 / 16; t++) {
        const int k2 = 16 * t + (l & 15);
        const bf16x8 b = ld8(net.W2bT + (long)k2 * kN2 + 8 * (l >> 4));
        const int p = k2 >> 5, c = k2 & 31;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            f32x4 d = mfma(a[r], b, f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int s = 64 * tile + 16 * r + 4 * (l >> 4) + i;
                const float dz = d[i] * dsilu(Z1[(long)s * kK2 + k2]);
                db[t & 1] += dz;
                dZ1T[(long)c * (9L * S) + (long)s * 9 + p] = f2bf(dz);
            }
        }
    }
    // lanes l, l^16, l^32, l^48 hold the same channel
#pragma unroll
    for (int h = 0; h < 2; h++) {
        float v = db[h];
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        if (l < 16) atomicAdd(&grad[kOffB1 + 16 * h + l], v);
    }
}

// dW1[j][k] += sum_{(s,p)} dZ1[(s,p)][j] A1[(s,p)][k]: M = 32, N = 96, K = kr
// rows of the (s,p) range per block.
__global__ __launch_bounds__(64) void ws_bwd1b_kernel(long K, int kr, const unsigned short *__restrict__ dZ1T,
                                                       const unsigned short *__restrict__ A1T, float *__restrict__ grad)
{
    const int l = threadIdx.x;
    const long r0 = (long)blockIdx.x * kr;
    f32x4 acc[2][6];
#pragma unroll
    for (int m = 0; m < 2; m++)
        for (int t = 0; t < 6; t++) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int q = 0; q < kr / 32; q++) {
        const long r = r0 + 32 * q + 8 * (l >> 4);
        bf16x8 a[2];
#pragma unroll
        for (int m = 0; m < 2; m++) a[m] = ld8(dZ1T + (long)(16 * m + (l & 15)) * K + r);
#pragma unroll
        for (int t = 0; t < 6; t++) {
            const bf16x8 b = ld8(A1T + (long)(16 * t + (l & 15)) * K + r);
#pragma unroll
            for (int m = 0; m < 2; m++) acc[m][t] = mfma(a[m], b, acc[m][t]);
        }
    }
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int t = 0; t < 6; t++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int j = 16 * m + 4 * (l >> 4) + i;
                const int k = 16 * t + (l & 15);
                if (k < kK1) atomicAdd(&grad[kOffW1 + j * kK1 + k], acc[m][t][i]);
            }
}

// Adam on the FP32 master parameters; refreshes the bf16 copies and clears
// the gradient.  step >= 1.
__global__ __launch_bounds__(256) void ws_adam_kernel(float *__restrict__ theta, float *__restrict__ grad,
                                                      float *__restrict__ m1, float *__restrict__ m2, float lr,
                                                      float beta1, float beta2, float eps, float c1, float c2,
                                                      unsigned short *__restrict__ W1b, unsigned short *__restrict__ W2b,
                                                      unsigned short *__restrict__ W2bT)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kNParam) return;
    float w = theta[i];
    if (lr > 0.f) {
        const float g = grad[i];
        const float a = beta1 * m1[i] + (1.f - beta1) * g;
        const float b = beta2 * m2[i] + (1.f - beta2) * g * g;
        m1[i] = a;
        m2[i] = b;
        w -= lr * (a * c1) / (sqrtf(b * c2) + eps);
        theta[i] = w;
    }
    grad[i] = 0.f;
    const unsigned short h = f2bf(w);
    if (i < kOffB1) {
        const int j = i / kK1, k = i % kK1;
        W1b[j * kK1p + k] = h;
    } else if (i >= kOffW2 && i < kOffB2) {
        const int j = (i - kOffW2) / kK2, k = (i - kOffW2) % kK2;
        W2b[j * kK2 + k] = h;
        W2bT[k * kN2 + j] = h;
    }
}

}  // namespace

// --------------------------------------------------------------- model object

struct WsModel {
    float *theta = nullptr, *grad = nullptr, *m1 = nullptr, *m2 = nullptr, *loss = nullptr;
    unsigned short *W1b = nullptr, *W2b = nullptr, *W2bT = nullptr;
    int cap = 0;  // samples of the scratch buffers (multiple of 64)
    float *Z1 = nullptr, *Z2 = nullptr, *H2 = nullptr;
    unsigned short *H1 = nullptr, *H1T = nullptr, *A1T = nullptr, *dZ2 = nullptr, *dZ2T = nullptr, *dZ1T = nullptr;
    unsigned short *Ppad = nullptr;  // padded copy of a partial last minibatch
    float *Tpad = nullptr;
    long step = 0;
    Net net() const { return Net{theta, W1b, W2b, W2bT}; }
};

int ws_alloc(WsModel &m, int cap)
{
    cap = (cap + 63) / 64 * 64;
    m.cap = cap;
    PP_TRY_HIP(hipMalloc((void **)&m.theta, kNParam * 4));
    PP_TRY_HIP(hipMalloc((void **)&m.grad, kNParam * 4));
    PP_TRY_HIP(hipMalloc((void **)&m.m1, kNParam * 4));
    PP_TRY_HIP(hipMalloc((void **)&m.m2, kNParam * 4));
    PP_TRY_HIP(hipMalloc((void **)&m.loss, 64));
    PP_TRY_HIP(hipMalloc((void **)&m.W1b, kN1 * kK1p * 2));
    PP_TRY_HIP(hipMalloc((void **)&m.W2b, kN2 * kK2 * 2));
    PP_TRY_HIP(hipMalloc((void **)&m.W2bT, kN2 * kK2 * 2));
    PP_TRY_HIP(hipMemset(m.grad, 0, kNParam * 4));
    PP_TRY_HIP(hipMemset(m.m1, 0, kNParam * 4));
    PP_TRY_HIP(hipMemset(m.m2, 0, kNParam * 4));
    PP_TRY_HIP(hipMemset(m.W1b, 0, kN1 * kK1p * 2));
    const long c = cap;
    PP_TRY_HIP(hipMalloc((void **)&m.Z1, c * kK2 * 4));
    PP_TRY_HIP(hipMalloc((void **)&m.Z2, c * kN2 * 4));
    PP_TRY_HIP(hipMalloc((void **)&m.H2, c * kN2 * 4));
    PP_TRY_HIP(hipMalloc((void **)&m.H1, c * kK2 * 2));
    PP_TRY_HIP(hipMalloc((void **)&m.H1T, c * kK2 * 2));
    PP_TRY_HIP(hipMalloc((void **)&m.A1T, c * 9 * kK1p * 2));
    PP_TRY_HIP(hipMalloc((void **)&m.dZ2, c * kN2 * 2));
    PP_TRY_HIP(hipMalloc((void **)&m.dZ2T, c * kN2 * 2));
    PP_TRY_HIP(hipMalloc((void **)&m.dZ1T, c * 9 * kN1 * 2));
    PP_TRY_HIP(hipMalloc((void **)&m.Ppad, c * kPatchStride * 2));
    PP_TRY_HIP(hipMalloc((void **)&m.Tpad, c * kN3 * 4));
    return 0;
}

void ws_free(WsModel &m)
{
    void *p[] = {m.theta, m.grad, m.m1, m.m2, m.loss, m.W1b, m.W2b, m.W2bT, m.Z1, m.Z2, m.H2,
                 m.H1, m.H1T, m.A1T, m.dZ2, m.dZ2T, m.dZ1T, m.Ppad, m.Tpad};
    for (void *q : p)
        if (q) (void)hipFree(q);
}

// bf16 copies from theta (and clear grad): Adam with lr = 0
int ws_refresh(WsModel &m, hipStream_t st)
{
    hipLaunchKernelGGL(ws_adam_kernel, dim3((kNParam + 255) / 256), dim3(256), 0, st, m.theta, m.grad, m.m1, m.m2, 0.f,
                       0.f, 0.f, 0.f, 0.f, 0.f, m.W1b, m.W2b, m.W2bT);
    return 0;
}

// Forward (and, with T, the gradient into m.grad and the loss into m.loss)
// over S rows of P (S a multiple of 64); nvalid rows carry targets.
int ws_pass(WsModel &m, int S, int nvalid, const unsigned short *P, const float *T, float *out, hipStream_t st)
{
    PP_REQUIRE(S % 64 == 0 && S <= m.cap && S > 0, "minibatch must be a positive multiple of 64 within capacity");
    const bool train = T != nullptr;
    const Net net = m.net();
    hipLaunchKernelGGL(ws_fwd1_kernel, dim3(S / 64, 9), dim3(64), 0, st, S, P, net, m.Z1, m.H1,
                       train ? m.H1T : nullptr, train ? m.A1T : nullptr);
    hipLaunchKernelGGL(ws_fwd2_kernel, dim3(S / 64), dim3(64), 0, st, S, m.H1, net, m.Z2, m.H2);
    const float inv_n = nvalid > 0 ? 1.f / (3.f * (float)nvalid) : 0.f;
    hipLaunchKernelGGL(ws_head_kernel, dim3((S + 255) / 256), dim3(256), 0, st, S, nvalid, inv_n, m.Z2, m.H2, net, T,
                       out, train ? m.grad : nullptr, m.loss, m.dZ2, m.dZ2T);
    if (!train) return 0;
    // dW2: K = S split into chunks of up to 512 samples
    int ks = 512;
    while (S % ks) ks >>= 1;
    hipLaunchKernelGGL(ws_bwd2_kernel, dim3(3, S / ks), dim3(64), 0, st, S, ks, m.dZ2T, m.H1T, m.grad);
    hipLaunchKernelGGL(ws_bwd1a_kernel, dim3(S / 64), dim3(64), 0, st, S, m.dZ2, net, m.Z1, m.dZ1T, m.grad);
    const long K = 9L * S;
    int kr = 4608;
    while (K % kr) kr >>= 1;
    hipLaunchKernelGGL(ws_bwd1b_kernel, dim3((unsigned)(K / kr)), dim3(64), 0, st, K, kr, m.dZ1T, m.A1T, m.grad);
    return 0;
}

}  // namespace pp

using namespace pp;

struct ppafm_ws_model {
    WsModel m;
};

extern "C" {

ppafm_ws_model *ppafm_ws_create(int max_batch)
{
    if (max_batch < 64) {
        set_error("max_batch must be >= 64");
        return nullptr;
    }
    ppafm_ws_model *w = new ppafm_ws_model;
    if (ws_alloc(w->m, max_batch)) {
        ws_free(w->m);
        delete w;
        return nullptr;
    }
    return w;
}

void ppafm_ws_destroy(ppafm_ws_model *w)
{
    if (!w) return;
    (void)hipDeviceSynchronize();
    ws_free(w->m);
    delete w;
}

int ppafm_ws_nparams(void) { return kNParam; }

int ppafm_ws_set_params(ppafm_ws_model *w, const float *theta, int n)
{
    PP_REQUIRE(w && theta, "model and theta must not be NULL");
    PP_REQUIRE(n == kNParam, "wrong parameter count");
    PP_TRY_HIP(hipMemcpy(w->m.theta, theta, kNParam * 4, hipMemcpyHostToDevice));
    PP_TRY_HIP(hipMemset(w->m.m1, 0, kNParam * 4));
    PP_TRY_HIP(hipMemset(w->m.m2, 0, kNParam * 4));
    w->m.step = 0;
    ws_refresh(w->m, nullptr);
    PP_TRY_HIP(hipDeviceSynchronize());
    return 0;
}

int ppafm_ws_get_params(const ppafm_ws_model *w, float *theta, int n)
{
    PP_REQUIRE(w && theta, "model and theta must not be NULL");
    PP_REQUIRE(n == kNParam, "wrong parameter count");
    PP_TRY_HIP(hipMemcpy(theta, w->m.theta, kNParam * 4, hipMemcpyDeviceToHost));
    return 0;
}

int ppafm_ws_forward(ppafm_ws_model *w, int n, const unsigned short *patches_dev, float *out_dev, void *stream)
{
    PP_REQUIRE(w && patches_dev && out_dev, "NULL argument");
    PP_REQUIRE(n >= 0, "n must be >= 0");
    hipStream_t st = as_stream(stream);
    WsModel &m = w->m;
    for (int s0 = 0; s0 < n; s0 += m.cap) {
        const int cnt = std::min(m.cap, n - s0);
        const int S = (cnt + 63) / 64 * 64;
        const unsigned short *P = patches_dev + (long)s0 * kPatchStride;
        if (S != cnt) {
            PP_TRY_HIP(hipMemsetAsync(m.Ppad, 0, (size_t)S * kPatchStride * 2, st));
            PP_TRY_HIP(hipMemcpyAsync(m.Ppad, P, (size_t)cnt * kPatchStride * 2, hipMemcpyDeviceToDevice, st));
            P = m.Ppad;
        }
        // padded rows: outputs land in Tpad, valid rows copied below
        float *o = S != cnt ? m.Tpad : out_dev + (long)s0 * kN3;
        int e = ws_pass(m, S, 0, P, nullptr, o, st);
        if (e) return e;
        if (S != cnt) PP_TRY_HIP(hipMemcpyAsync(out_dev + (long)s0 * kN3, m.Tpad, (size_t)cnt * kN3 * 4, hipMemcpyDeviceToDevice, st));
    }
    return 0;
}

// Loss and gradient of the mean loss over n samples (no update).
int ppafm_ws_loss_grad(ppafm_ws_model *w, int n, const unsigned short *patches_dev, const float *targets_dev,
                       float *grad_host, float *loss_host)
{
    PP_REQUIRE(w && patches_dev && targets_dev, "NULL argument");
    WsModel &m = w->m;
    PP_REQUIRE(n > 0 && n <= m.cap, "n must be within the model's minibatch capacity");
    const int S = (n + 63) / 64 * 64;
    const unsigned short *P = patches_dev;
    const float *T = targets_dev;
    if (S != n) {
        PP_TRY_HIP(hipMemset(m.Ppad, 0, (size_t)S * kPatchStride * 2));
        PP_TRY_HIP(hipMemcpy(m.Ppad, P, (size_t)n * kPatchStride * 2, hipMemcpyDeviceToDevice));
        PP_TRY_HIP(hipMemset(m.Tpad, 0, (size_t)S * kN3 * 4));
        PP_TRY_HIP(hipMemcpy(m.Tpad, T, (size_t)n * kN3 * 4, hipMemcpyDeviceToDevice));
        P = m.Ppad;
        T = m.Tpad;
    }
    PP_TRY_HIP(hipMemset(m.grad, 0, kNParam * 4));
    PP_TRY_HIP(hipMemset(m.loss, 0, 4));
    int e = ws_pass(m, S, n, P, T, nullptr, nullptr);
    if (e) return e;
    PP_TRY_HIP(hipDeviceSynchronize());
    if (grad_host) PP_TRY_HIP(hipMemcpy(grad_host, m.grad, kNParam * 4, hipMemcpyDeviceToHost));
    if (loss_host) PP_TRY_HIP(hipMemcpy(loss_host, m.loss, 4, hipMemcpyDeviceToHost));
    PP_TRY_HIP(hipMemset(m.grad, 0, kNParam * 4));
    return 0;
}

// Adam over n samples in minibatches of `batch` (a multiple of 64), in the
// given order (index array of n samples, device, may be NULL for 0..n-1 --
// then the data should already be shuffled).  lr follows a cosine schedule
// from lr0 to lr0*lr_end_frac over `steps` minibatches (steps <= 0: one pass).
// losses_host (steps entries, may be NULL) receives each minibatch loss.
int ppafm_ws_train(ppafm_ws_model *w, long n, const unsigned short *patches_dev, const float *targets_dev, int batch,
                   long steps, float lr0, float lr_end_frac, float *losses_host, void *stream)
{
    PP_REQUIRE(w && patches_dev && targets_dev, "NULL argument");
    WsModel &m = w->m;
    PP_REQUIRE(batch % 64 == 0 && batch > 0 && batch <= m.cap, "batch must be a multiple of 64 within capacity");
    PP_REQUIRE(n >= batch, "need at least one full minibatch");
    hipStream_t st = as_stream(stream);
    const long nb = n / batch;
    if (steps <= 0) steps = nb;
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    std::vector<float> losses;
    PP_TRY_HIP(hipMemsetAsync(m.grad, 0, kNParam * 4, st));
    for (long k = 0; k < steps; k++) {
        const long b = k % nb;
        const unsigned short *P = patches_dev + b * batch * (long)kPatchStride;
        const float *T = targets_dev + b * batch * (long)kN3;
        PP_TRY_HIP(hipMemsetAsync(m.loss + (k & 15), 0, 4, st));
        float *lsave = m.loss;
        m.loss = lsave + (k & 15);
        int e = ws_pass(m, batch, batch, P, T, nullptr, st);
        m.loss = lsave;
        if (e) return e;
        m.step++;
        const double frac = steps > 1 ? (double)k / (double)(steps - 1) : 1.0;
        const double lr = lr0 * (lr_end_frac + (1.0 - lr_end_frac) * 0.5 * (1.0 + cos(3.141592653589793 * frac)));
        const float c1 = (float)(1.0 / (1.0 - pow(beta1, (double)m.step)));
        const float c2 = (float)(1.0 / (1.0 - pow(beta2, (double)m.step)));
        hipLaunchKernelGGL(ws_adam_kernel, dim3((kNParam + 255) / 256), dim3(256), 0, st, m.theta, m.grad, m.m1, m.m2,
                           (float)lr, beta1, beta2, eps, c1, c2, m.W1b, m.W2b, m.W2bT);
        if (losses_host && ((k & 15) == 15 || k == steps - 1)) {
            float l16[16];
            PP_TRY_HIP(hipMemcpyAsync(l16, m.loss, 64, hipMemcpyDeviceToHost, st));
            PP_TRY_HIP(hipStreamSynchronize(st));
            const long base = k & ~15L;
            for (long q = base; q <= k; q++) losses_host[q] = l16[q & 15];
        }
    }
    PP_TRY_HIP(hipStreamSynchronize(st));
    return 0;
}

}  // extern "C"
