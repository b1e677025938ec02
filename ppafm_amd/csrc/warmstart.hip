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
#include "relax_dev.hpp"

namespace pp {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kN1 = 32, kK1 = 81, kK1p = 96, kN2 = 32, kK2 = 288, kN3 = 3;
constexpr int kPatch = 225, kPatchStride = 232;  // bf16 per patch row (16-byte aligned)
constexpr int kOffW1 = 0, kOffB1 = kN1 * kK1, kOffW2 = kOffB1 + kN1, kOffB2 = kOffW2 + kN2 * kK2;
constexpr int kOffW3 = kOffB2 + kN2, kOffB3 = kOffW3 + kN3 * kN2, kNParam = kOffB3 + kN3;
static_assert(kNParam == 11971, "parameter count");

// float -> bf16 bits, round to nearest even (v_cvt_pk_bf16_f32 on gfx950)
__device__ __forceinline__ unsigned short f2bf(float x) { return __builtin_bit_cast(unsigned short, (__bf16)x); }
__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float((unsigned)b << 16); }
__device__ __forceinline__ float bfr(float x) { return bf2f(f2bf(x)); }

// sigmoid with the hardware reciprocal (about 1 ulp; the network runs in
// bf16 anyway)
__device__ __forceinline__ float sigm(float z) { return __builtin_amdgcn_rcpf(1.f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
__device__ __forceinline__ float dsilu(float z)
{
    const float s = sigm(z);
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

// Offset of conv1 tap k = ((dx*3 + dy)*3 + dz)*3 + c inside a patch row,
// relative to the window origin; -1 for the padding k >= 81.  Computed, not
// looked up: the index differs across lanes.
__device__ __forceinline__ int tap_off(int k)
{
    const int blk = k / 9, r9 = k - 9 * blk;  // blk = dx*3 + dy
    const int dx = blk / 3, dy = blk - 3 * dx;
    return k < kK1 ? (dx * 5 + dy) * 9 + r9 : -1;
}

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
                const int off = tap_off(k0 + e);
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

// Layer 3, loss and the head of the backward pass; one thread per sample.
// With grad, each block of 256 samples also leaves the sums of its samples'
// head gradient terms (dW3, db3, db2 and the loss, kHeadVals values) in
// part[block]; ws_head_sum_kernel adds the blocks up in order (no atomics:
// the same sums every run).  Samples s >= nvalid are padding: they produce
// outputs but no gradient.
constexpr int kHeadVals = kN3 * kN2 + kN3 + kN2 + 1;
constexpr int kHeadRow = kN3 + kN2 + kN2 + 2;  // dy, bf16(h), dz, loss term, pad: an odd stride of floats

__global__ __launch_bounds__(256) void ws_head_kernel(int S, int nvalid, float inv_n, const float *__restrict__ Z2,
                                                      const float *__restrict__ H2, Net net,
                                                      const float *__restrict__ T, float *__restrict__ out,
                                                      float *__restrict__ part, unsigned short *__restrict__ dZ2,
                                                      unsigned short *__restrict__ dZ2T)
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
    if (!part) return;  // uniform: forward only
    __shared__ float sh[256][kHeadRow];
    const int t = threadIdx.x;
    float dy[kN3], lsum = 0.f;
#pragma unroll
    for (int m = 0; m < kN3; m++) {
        const float d = valid ? o[m] - T[(long)s * kN3 + m] : 0.f;
        lsum += d * d;
        dy[m] = 2.f * d * inv_n;
        sh[t][m] = dy[m];
    }
#pragma unroll
    for (int j = 0; j < kN2; j++) sh[t][kN3 + j] = h[j];
    sh[t][kN3 + 2 * kN2] = lsum * inv_n;
    // dZ2 = (dY W3) * silu'(Z2)
#pragma unroll
    for (int j = 0; j < kN2; j++) {
        float dh = 0.f;
#pragma unroll
        for (int m = 0; m < kN3; m++) dh = fmaf(dy[m], bfr(W3[m * kN2 + j]), dh);
        const float dz = live ? dh * dsilu(Z2[(long)s * kN2 + j]) : 0.f;
        sh[t][kN3 + kN2 + j] = dz;
        if (live) {
            const unsigned short b = f2bf(dz);
            dZ2[(long)s * kN2 + j] = b;
            dZ2T[(long)j * S + s] = b;
        }
    }
    __syncthreads();
    // the block's sums: thread v < kHeadVals takes value v over the 256 samples
    if (t < kHeadVals) {
        float acc = 0.f;
        if (t < kN3 * kN2) {  // dW3[m][j] = sum dy[m] bf16(h[j])
            const int m = t / kN2, j = t % kN2;
            for (int q = 0; q < 256; q++) acc = fmaf(sh[q][m], sh[q][kN3 + j], acc);
        } else {
            // db3 (dy), db2 (dz), the loss term
            const int c = t < kN3 * kN2 + kN3 ? t - kN3 * kN2 : t < kN3 * kN2 + kN3 + kN2 ? kN3 + kN2 + (t - kN3 * kN2 - kN3)
                                                                                            : kN3 + 2 * kN2;
            for (int q = 0; q < 256; q++) acc += sh[q][c];
        }
        part[(long)blockIdx.x * kHeadVals + t] = acc;
    }
}

// The head's gradient and loss from the blocks' sums, added in block order.
__global__ __launch_bounds__(256) void ws_head_sum_kernel(int nblocks, const float *__restrict__ part,
                                                          float *__restrict__ grad, float *__restrict__ loss)
{
    const int v = threadIdx.x;
    if (v >= kHeadVals) return;
    float acc = 0.f;
    for (int b = 0; b < nblocks; b++) acc += part[(long)b * kHeadVals + v];
    if (v < kN3 * kN2) grad[kOffW3 + v] += acc;
    else if (v < kN3 * kN2 + kN3) grad[kOffB3 + v - kN3 * kN2] += acc;
    else if (v < kHeadVals - 1) grad[kOffB2 + v - kN3 * kN2 - kN3] += acc;
    else loss[0] += acc;
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
// written transposed as dZ1T[c][s*9 + p]; db1.  Block (tile, y) takes the
// 16-column slices 6y .. 6y+5 of the 18.
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
    for (int t = 6 * blockIdx.y; t < 6 * blockIdx.y + 6; t++) {
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

// dW1[j][k] = sum_{(s,p)} dZ1[(s,p)][j] A1[(s,p)][k]: M = 32, N = 96, K = kr
// rows of the (s,p) range per block, whose sums go to part[block][j][k]
// (ws_w1_sum_kernel adds the blocks up in order).
__global__ __launch_bounds__(64) void ws_bwd1b_kernel(long K, int kr, const unsigned short *__restrict__ dZ1T,
                                                       const unsigned short *__restrict__ A1T, float *__restrict__ part)
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
                if (k < kK1) part[(long)blockIdx.x * (kN1 * kK1) + j * kK1 + k] = acc[m][t][i];
            }
}

// dW1 from ws_bwd1b_kernel's block sums, added in block order.
__global__ __launch_bounds__(256) void ws_w1_sum_kernel(int nblocks, const float *__restrict__ part,
                                                        float *__restrict__ grad)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= kN1 * kK1) return;
    float acc = 0.f;
    for (int b = 0; b < nblocks; b++) acc += part[(long)b * (kN1 * kK1) + v];
    grad[kOffW1 + v] += acc;
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


// ------------------------------------------------------------ use in FIRE

constexpr double kWsScale = 0.1;  // network output unit [A]

__device__ __forceinline__ int wrap_idx(int i, int n)
{
    i %= n;
    return i < 0 ? i + n : i;
}

__device__ __forceinline__ int search_prefix(const long long *__restrict__ prefix, int nitems, long long x)
{
    int lo = 0, hi = nitems;  // prefix[lo] <= x < prefix[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ void item_args(DevArgs &A, const BatchRelaxItem &it)
{
    A.n0 = it.n[0]; A.n1 = it.n[1]; A.n2 = it.n[2];
    A.nx = it.nx; A.ny = it.ny; A.nz = it.nz;
    for (int k = 0; k < 9; k++) A.dinv[k] = it.dinv[k];
}

// Patch of one probe start: total force (grid force + tip springs) at the
// 5x5x3 grid nodes around the node nearest to (sx, sy, sz), index
// ((x*5 + y)*3 + z)*3 + c, each component scaled by its own RMS over the
// patch (which keeps the zero set of every component, and with it the
// minimum, while the soft lateral components keep bf16 precision next to the
// stiff one), as bf16 (rounded once before and once after the scaling).
// The patch feeds bf16 matrix products, so it is computed in FP32 (node
// offsets from the tip are a few A); the 15 nodes of each x slab are loaded
// before any is used.  Returns the centre node position (FP64).
template <typename T4, int MODE>
__device__ __forceinline__ void ws_patch(const DevArgs &A, const T4 *__restrict__ grid, const double *dcell, double tx,
                                         double ty, double tz, double sx, double sy, double sz,
                                         unsigned short *row, double &nx_, double &ny_, double &nz_)
{
    double gx, gy, gz;
    to_grid(A, sx, sy, sz, gx, gy, gz);
    const int ix = (int)floor(gx + 0.5), iy = (int)floor(gy + 0.5), iz = (int)floor(gz + 0.5);
    nx_ = ix * dcell[0] + iy * dcell[3] + iz * dcell[6];
    ny_ = ix * dcell[1] + iy * dcell[4] + iz * dcell[7];
    nz_ = ix * dcell[2] + iy * dcell[5] + iz * dcell[8];
    // node offsets from the tip, FP32
    const float ox = (float)(nx_ - tx), oy = (float)(ny_ - ty), oz = (float)(nz_ - tz);
    float dc[9];
    for (int k = 0; k < 9; k++) dc[k] = (float)dcell[k];
    float R[9], ks[3], d0[3];
    for (int k = 0; k < 9; k++) R[k] = (float)A.rot[k];
    for (int k = 0; k < 3; k++) { ks[k] = (float)A.ks[k]; d0[k] = (float)A.d0[k]; }
    const float krad = (float)A.krad, lrad = (float)A.lrad, surf = MODE == 1 ? (float)A.surf_ff : 0.f;
    const int jz[3] = {wrap_idx(iz - 1, A.n2), wrap_idx(iz, A.n2), wrap_idx(iz + 1, A.n2)};
    float sumsq[3] = {0.f, 0.f, 0.f};
    for (int dx = -2; dx <= 2; dx++) {
        const int jx = wrap_idx(ix + dx, A.n0);
        float vx[15], vy[15], vz[15];
#pragma unroll
        for (int q = 0; q < 15; q++) {
            const int dy = q / 3 - 2;
            const unsigned idx = ((unsigned)jx * A.n1 + wrap_idx(iy + dy, A.n1)) * A.n2 + jz[q % 3];
            const T4 v = grid[idx];
            vx[q] = (float)v.x; vy[q] = (float)v.y; vz[q] = (float)v.z;
        }
#pragma unroll
        for (int q = 0; q < 15; q++) {
            const int dy = q / 3 - 2, dz = q % 3 - 1;
            // probe - tip at this node, then the springs (as total_force)
            const float ex = ox + dx * dc[0] + dy * dc[3] + dz * dc[6];
            const float ey = oy + dx * dc[1] + dy * dc[4] + dz * dc[7];
            const float ez = oz + dx * dc[2] + dy * dc[5] + dz * dc[8];
            const float qx = ex * R[0] + ey * R[1] + ez * R[2];
            const float qy = ex * R[3] + ey * R[4] + ez * R[5];
            const float qz = ex * R[6] + ey * R[7] + ez * R[8];
            const float il = rsqrtf(qx * qx + qy * qy + qz * qz);
            const float sr = krad * (1.f - lrad * il);
            const float ax = (qx - d0[0]) * ks[0] + qx * sr;
            const float ay = (qy - d0[1]) * ks[1] + qy * sr;
            const float az = (qz - d0[2]) * ks[2] + qz * sr + surf;
            const float f3[3] = {vx[q] + R[0] * ax + R[3] * ay + R[6] * az, vy[q] + R[1] * ax + R[4] * ay + R[7] * az,
                                 vz[q] + R[2] * ax + R[5] * ay + R[8] * az};
            const int e = (((dx + 2) * 5 + (dy + 2)) * 3 + (dz + 1)) * 3;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                sumsq[c] += f3[c] * f3[c];
                row[e + c] = f2bf(f3[c]);
            }
        }
    }
    float inv[3];
    for (int c = 0; c < 3; c++) {
        const float rms = sqrtf(sumsq[c] * (3.f / kPatch));
        inv[c] = rms > 0.f ? 1.f / rms : 0.f;
    }
    for (int e = 0; e < kPatch; e++) row[e] = f2bf(bf2f(row[e]) * inv[e % 3]);
    for (int e = kPatch; e < kPatchStride; e++) row[e] = 0;
}

template <typename T4, int MODE>
__global__ __launch_bounds__(256) void ws_samples_kernel(DevArgs A0, const BatchRelaxItem *__restrict__ items,
                                                         const long long *__restrict__ pos_prefix, int nitems,
                                                         const T4 *__restrict__ grids, const float *__restrict__ pos,
                                                         const float *__restrict__ truth, long long nsel,
                                                         const long long *__restrict__ sel,
                                                         unsigned short *__restrict__ patches,
                                                         float *__restrict__ targets, float jitter, unsigned seed)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nsel) return;
    const long long g = sel[q];
    const int m = search_prefix(pos_prefix, nitems, g);
    const BatchRelaxItem &it = items[m];
    DevArgs A = A0;
    item_args(A, it);
    const long long o = g - pos_prefix[m];
    const int s = (int)(o % it.nz);
    const long long col = o / it.nz;
    const int i = (int)(col / it.ny), j = (int)(col % it.ny);
    const double bx = it.p0[0] + i * it.da[0] + j * it.db[0];
    const double by = it.p0[1] + i * it.da[1] + j * it.db[1];
    const double bz = it.p0[2] + i * it.da[2] + j * it.db[2];
    const double tx = bx + s * it.dstep[0], ty = by + s * it.dstep[1], tz = bz + s * it.dstep[2];
    double sx, sy, sz;
    if (s == 0) {
        double ox, oy, oz;
        probe_offset<MODE>(A, ox, oy, oz);
        sx = tx + ox; sy = ty + oy; sz = tz + oz;
    } else {
        sx = (double)pos[3 * (g - 1)] + it.dstep[0];
        sy = (double)pos[3 * (g - 1) + 1] + it.dstep[1];
        sz = (double)pos[3 * (g - 1) + 2] + it.dstep[2];
    }
    if (jitter > 0.f) {
        // uniform offset in [-jitter, jitter) laterally and half that along z,
        // a function of the position index and the seed only
        unsigned h = ((unsigned)g ^ (unsigned)(g >> 32) * 0x85ebca6bu) * 0x9E3779B1u ^ seed;
        float u[3];
        for (int k = 0; k < 3; k++) {
            h ^= h >> 16; h *= 0x7feb352du; h ^= h >> 15; h *= 0x846ca68bu; h ^= h >> 16;
            u[k] = (float)(h >> 8) * (1.f / 16777216.f) * 2.f - 1.f;
            h += 0x6a09e667u;
        }
        sx += jitter * u[0]; sy += jitter * u[1]; sz += 0.5f * jitter * u[2];
    }
    double nx_, ny_, nz_;
    ws_patch<T4, MODE>(A, grids + it.grid_off, it.dcell, tx, ty, tz, sx, sy, sz, patches + q * kPatchStride, nx_, ny_, nz_);
    targets[3 * q] = (float)(((double)truth[3 * g] - nx_) / kWsScale);
    targets[3 * q + 1] = (float)(((double)truth[3 * g + 1] - ny_) / kWsScale);
    targets[3 * q + 2] = (float)(((double)truth[3 * g + 2] - nz_) / kWsScale);
}

// Step-0 estimates, persistent: each block of 256 threads stages the
// weights once and then takes groups of 8x8 pixel columns (a tile of one
// item; wave_prefix[k] = tiles of items < k, ngroups in all) until none are
// left; output pred[col_off + i*ny + j].  Per tile and network pass:
//   wave 0, lane l     nearest grid node of column l's estimate (first pass:
//                      the plain start, tip + probe offset), and the box of
//                      grid nodes that holds the tile's 64 patches;
//   all 4 waves        that box, loaded once into LDS (coalesced along z;
//                      direct gathers when it is larger than kBoxCap);
//   all 4 waves        the 75 patch nodes of every column, split 4 ways,
//                      from the box into LDS rows (bf16), per-component RMS
//                      scaling;
//   wave w             layers 1-3 for columns 16w..16w+15 (layers 1-2 on the
//                      matrix cores); only this wave reads what it writes,
//                      so a wave-scope fence orders it, not a block barrier.
// The relaxation's step-0 Newton run refines the estimate.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int wave_min_i(int v)
{
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ int wave_max_i(int v)
{
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

constexpr int kBoxCap = 1024;  // grid nodes of a staged box (x, y, z floats)

template <typename T4, int MODE>
__global__ __launch_bounds__(256) void ws_predict_kernel(DevArgs A0, Net net, const BatchRelaxItem *__restrict__ items,
                                                         const long long *__restrict__ wave_prefix, int nitems,
                                                         long long ngroups, const T4 *__restrict__ grids,
                                                         float *__restrict__ pred, int reps)
{
    // patch rows of 226 bf16 (113 dwords, odd: consecutive rows start in
    // different LDS banks)
    constexpr int kRow = 226;
    __shared__ __align__(16) unsigned short Pl[64 * kRow];
    // the staged box (3 planes of kBoxCap floats) is dead once the patch rows
    // are built; layers 1-2 then reuse its memory for H1l and H2l
    constexpr int kUnion = 3 * kBoxCap * 4 > 64 * 40 * 2 + 64 * 33 * 4 ? 3 * kBoxCap * 4 : 64 * 40 * 2 + 64 * 33 * 4;
    __shared__ __align__(16) unsigned char U[kUnion];
    unsigned short *H1l = (unsigned short *)U;            // rows 16w.. belong to wave w
    float *H2l = (float *)(U + 64 * 40 * 2);              // rows 16w.. belong to wave w
    float *Bx = (float *)U, *By = Bx + kBoxCap, *Bz = By + kBoxCap;
    __shared__ int box[6];                                // lo x, y, z; extent x, y, z (0: not staged)
    __shared__ __align__(16) unsigned short W1l[kN1 * kK1p];
    __shared__ __align__(16) unsigned short W2l[kN2 * kK2];
    __shared__ float W3l[kN3 * kN2];  // bf16-rounded
    __shared__ float red[4][64][3];
    __shared__ int nodes[64][3];
    __shared__ float toff[64][3];     // centre node - tip, FP32
    __shared__ double cnode[64][3];   // centre node position
    __shared__ double est[64][3];     // the estimates
    const int t = threadIdx.x;
    const int w = t >> 6, l = t & 63;
    const float *W3 = net.theta + kOffW3;
    const int q0 = (75 * w) / 4, q1 = (75 * (w + 1)) / 4;  // this wave's patch nodes
    for (int k = t; k < kN1 * kK1p / 8; k += 256) ((uint4 *)W1l)[k] = ((const uint4 *)net.W1b)[k];
    for (int k = t; k < kN2 * kK2 / 8; k += 256) ((uint4 *)W2l)[k] = ((const uint4 *)net.W2b)[k];
    for (int k = t; k < kN3 * kN2; k += 256) W3l[k] = bfr(W3[k]);
    // node deltas of this thread's patch nodes, packed (dx+2) | (dy+2) << 4 | (dz+1) << 8
    int nd[19];
#pragma unroll
    for (int k = 0; k < 19; k++) {
        const int q = q0 + k;
        nd[k] = (q / 15) | (((q / 3) % 5) << 4) | ((q % 3) << 8);
    }
    // conv1 tap offsets of this lane's A fragments (k0 = 32q + 8(l/16) + e)
    short toffs[3][8];
    for (int q = 0; q < 3; q++)
        for (int e = 0; e < 8; e++) toffs[q][e] = (short)tap_off(32 * q + 8 * (l >> 4) + e);
    double ox, oy, oz;
    probe_offset<MODE>(A0, ox, oy, oz);

    // a contiguous range of tiles per block: consecutive tiles mostly share
    // their item, which is then looked up once
    const long long per = (ngroups + gridDim.x - 1) / gridDim.x;
    const long long g0 = (long long)blockIdx.x * per, g1 = min(ngroups, g0 + per);
    int m = g0 < g1 ? search_prefix(wave_prefix, nitems, g0) : 0;
    long long mend = g0 < g1 ? wave_prefix[m + 1] : 0;
    for (long long grp = g0; grp < g1; grp++) {
        while (grp >= mend) {  // next item with tiles (items without any have equal prefixes)
            m++;
            mend = wave_prefix[m + 1];
        }
        const BatchRelaxItem &it = items[m];
        DevArgs A = A0;
        item_args(A, it);
        const int tiles_y = (it.ny + 7) / 8;
        const int tile = (int)(grp - wave_prefix[m]);
        const int i = (tile / tiles_y) * 8 + (l >> 3), j = (tile % tiles_y) * 8 + (l & 7);  // column l of the tile
        const bool live = i < it.nx && j < it.ny;
        const double tx = it.p0[0] + i * it.da[0] + j * it.db[0];
        const double ty = it.p0[1] + i * it.da[1] + j * it.db[1];
        const double tz = it.p0[2] + i * it.da[2] + j * it.db[2];
        const T4 *grid = grids + it.grid_off;
        float dc[9], R[9], ks[3], d0[3];
        for (int k = 0; k < 9; k++) { dc[k] = (float)it.dcell[k]; R[k] = (float)A.rot[k]; }
        for (int k = 0; k < 3; k++) { ks[k] = (float)A.ks[k]; d0[k] = (float)A.d0[k]; }
        const float krad = (float)A.krad, lrad = (float)A.lrad, surf = MODE == 1 ? (float)A.surf_ff : 0.f;
        __syncthreads();  // the previous group is done with every LDS buffer
        if (w == 0) { est[l][0] = tx + ox; est[l][1] = ty + oy; est[l][2] = tz + oz; }
        for (int rep = 0; rep < reps; rep++) {
            // ---- centre node of every column (wave 0)
            if (w == 0) {
                double gx, gy, gz;
                to_grid(A, est[l][0], est[l][1], est[l][2], gx, gy, gz);
                const int ix = (int)floor(gx + 0.5), iy = (int)floor(gy + 0.5), iz = (int)floor(gz + 0.5);
                const double nx_ = ix * it.dcell[0] + iy * it.dcell[3] + iz * it.dcell[6];
                const double ny_ = ix * it.dcell[1] + iy * it.dcell[4] + iz * it.dcell[7];
                const double nz_ = ix * it.dcell[2] + iy * it.dcell[5] + iz * it.dcell[8];
                cnode[l][0] = nx_; cnode[l][1] = ny_; cnode[l][2] = nz_;
                nodes[l][0] = wrap_idx(ix, A.n0); nodes[l][1] = wrap_idx(iy, A.n1); nodes[l][2] = wrap_idx(iz, A.n2);
                toff[l][0] = (float)(nx_ - tx); toff[l][1] = (float)(ny_ - ty); toff[l][2] = (float)(nz_ - tz);
                // box of the live columns' patches (unwrapped node indices)
                const int big = 1 << 30;
                const int lx = wave_min_i(live ? ix - 2 : big), hx = wave_max_i(live ? ix + 2 : -big);
                const int ly = wave_min_i(live ? iy - 2 : big), hy = wave_max_i(live ? iy + 2 : -big);
                const int lz = wave_min_i(live ? iz - 1 : big), hz = wave_max_i(live ? iz + 1 : -big);
                if (l == 0) {
                    const long ex = (long)hx - lx + 1, ey = (long)hy - ly + 1, ez = (long)hz - lz + 1;
                    const bool fits = ex > 0 && ey > 0 && ez > 0 && ex <= A.n0 && ey <= A.n1 && ez <= A.n2 &&
                                      ex * ey * ez <= kBoxCap;
                    box[0] = lx; box[1] = ly; box[2] = lz;
                    box[3] = fits ? (int)ex : 0; box[4] = fits ? (int)ey : 0; box[5] = fits ? (int)ez : 0;
                }
            }
            __syncthreads();  // (1) nodes and the box
            const int bx0 = box[0], by0 = box[1], bz0 = box[2], ex = box[3], ey = box[4], ez = box[5];
            const bool staged = ex > 0;
            if (staged) {
                const int vol = ex * ey * ez;
                for (int q = t; q < vol; q += 256) {
                    const int cz = q % ez, r = q / ez;
                    const int cy = r % ey, cx = r / ey;
                    const T4 v = grid[((unsigned)wrap_idx(bx0 + cx, A.n0) * A.n1 + wrap_idx(by0 + cy, A.n1)) * A.n2 +
                                      wrap_idx(bz0 + cz, A.n2)];
                    Bx[q] = (float)v.x; By[q] = (float)v.y; Bz[q] = (float)v.z;
                }
                __syncthreads();  // (1b) the box
            }
            // ---- patch nodes, split over the 4 waves
            {
                const int ix = nodes[l][0], iy = nodes[l][1], iz = nodes[l][2];  // wrapped into the grid
                // this column's patch origin inside the box (the box's lower
                // corner wraps to the same node as the column's, modulo n)
                const int ox_ = staged ? wrap_idx(ix - 2 - bx0, A.n0) : 0;
                const int oy_ = staged ? wrap_idx(iy - 2 - by0, A.n1) : 0;
                const int oz_ = staged ? wrap_idx(iz - 1 - bz0, A.n2) : 0;
                const float ex0 = toff[l][0], ey0 = toff[l][1], ez0 = toff[l][2];
                unsigned short *row = Pl + l * kRow;
                float ss[3] = {0.f, 0.f, 0.f};
                // loads in chunks of 7 nodes, each chunk issued before use
                for (int k0 = 0; k0 < 19; k0 += 7) {
                    float vx[7], vy[7], vz[7];
#pragma unroll
                    for (int k = 0; k < 7; k++) {
                        const int q = q0 + k0 + k;
                        if (k0 + k < 19 && q < q1) {
                            const int dx = (nd[k0 + k] & 15) - 2, dy = ((nd[k0 + k] >> 4) & 15) - 2, dz = (nd[k0 + k] >> 8) - 1;
                            if (!live) {
                                vx[k] = vy[k] = vz[k] = 0.f;  // a column off the scan (edge tile)
                            } else if (staged) {
                                const int bq = ((ox_ + dx + 2) * ey + (oy_ + dy + 2)) * ez + (oz_ + dz + 1);
                                vx[k] = Bx[bq]; vy[k] = By[bq]; vz[k] = Bz[bq];
                            } else {
                                int jx = ix + dx, jy = iy + dy, jz = iz + dz;  // |d| <= 2 < n: one wrap at most
                                jx += jx < 0 ? A.n0 : (jx >= A.n0 ? -A.n0 : 0);
                                jy += jy < 0 ? A.n1 : (jy >= A.n1 ? -A.n1 : 0);
                                jz += jz < 0 ? A.n2 : (jz >= A.n2 ? -A.n2 : 0);
                                const T4 *pv = grid + ((unsigned)jx * A.n1 + jy) * A.n2 + jz;
                                vx[k] = (float)pv->x; vy[k] = (float)pv->y; vz[k] = (float)pv->z;
                            }
                        }
                    }
#pragma unroll
                    for (int k = 0; k < 7; k++) {
                        const int q = q0 + k0 + k;
                        if (k0 + k < 19 && q < q1) {
                            const float dx = (float)((nd[k0 + k] & 15) - 2), dy = (float)(((nd[k0 + k] >> 4) & 15) - 2);
                            const float dz = (float)((nd[k0 + k] >> 8) - 1);
                            const float ex = ex0 + dx * dc[0] + dy * dc[3] + dz * dc[6];
                            const float ey = ey0 + dx * dc[1] + dy * dc[4] + dz * dc[7];
                            const float ez = ez0 + dx * dc[2] + dy * dc[5] + dz * dc[8];
                            const float qx = ex * R[0] + ey * R[1] + ez * R[2];
                            const float qy = ex * R[3] + ey * R[4] + ez * R[5];
                            const float qz = ex * R[6] + ey * R[7] + ez * R[8];
                            const float il = rsqrtf(qx * qx + qy * qy + qz * qz);
                            const float sr = krad * (1.f - lrad * il);
                            const float ax = (qx - d0[0]) * ks[0] + qx * sr;
                            const float ay = (qy - d0[1]) * ks[1] + qy * sr;
                            const float az = (qz - d0[2]) * ks[2] + qz * sr + surf;
                            const float f3[3] = {vx[k] + R[0] * ax + R[3] * ay + R[6] * az,
                                                 vy[k] + R[1] * ax + R[4] * ay + R[7] * az,
                                                 vz[k] + R[2] * ax + R[5] * ay + R[8] * az};
#pragma unroll
                            for (int c = 0; c < 3; c++) {
                                ss[c] += f3[c] * f3[c];
                                row[3 * q + c] = f2bf(f3[c]);
                            }
                        }
                    }
                }
                for (int c = 0; c < 3; c++) red[w][l][c] = ss[c];
            }
            __syncthreads();  // (2) patch rows and partial sums
            {
                float inv[3];
                for (int c = 0; c < 3; c++) {
                    const float rms = sqrtf((red[0][l][c] + red[1][l][c] + red[2][l][c] + red[3][l][c]) * (3.f / kPatch));
                    inv[c] = rms > 0.f ? 1.f / rms : 0.f;
                }
                unsigned short *row = Pl + l * kRow;
                if (!live) inv[0] = inv[1] = inv[2] = 0.f;
                for (int q = q0; q < q1; q++)
#pragma unroll
                    for (int c = 0; c < 3; c++) row[3 * q + c] = f2bf(bf2f(row[3 * q + c]) * inv[c]);
            }
            __syncthreads();  // (3) normalised rows
            // ---- layers 1 and 2: wave w takes columns 16w .. 16w+15
            f32x4 acc2[2] = {f32x4{0.f, 0.f, 0.f, 0.f}, f32x4{0.f, 0.f, 0.f, 0.f}};
            for (int p = 0; p < 9; p++) {
                const int poff = ((p / 3) * 5 + (p % 3)) * 9;
                f32x4 acc1[2] = {f32x4{0.f, 0.f, 0.f, 0.f}, f32x4{0.f, 0.f, 0.f, 0.f}};
                const unsigned short *arow = Pl + (16 * w + (l & 15)) * kRow + poff;
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    const int k0 = 32 * q + 8 * (l >> 4);
                    unsigned short v[8];
#pragma unroll
                    for (int e = 0; e < 8; e++) {
                        const int off = toffs[q][e];
                        v[e] = off >= 0 ? arow[off] : (unsigned short)0;
                    }
                    const bf16x8 a = pack8(v);
#pragma unroll
                    for (int tt = 0; tt < 2; tt++)
                        acc1[tt] = mfma(a, ld8(W1l + (16 * tt + (l & 15)) * kK1p + k0), acc1[tt]);
                }
                wave_sync();  // this wave's reads of H1l (previous p) are done
#pragma unroll
                for (int tt = 0; tt < 2; tt++) {
                    const int jj = 16 * tt + (l & 15);
                    const float bj = net.theta[kOffB1 + jj];
#pragma unroll
                    for (int ii = 0; ii < 4; ii++)
                        H1l[(16 * w + 4 * (l >> 4) + ii) * 40 + jj] = f2bf(silu(acc1[tt][ii] + bj));
                }
                wave_sync();  // H1l rows of this wave written
                const bf16x8 a2 = ld8(H1l + (16 * w + (l & 15)) * 40 + 8 * (l >> 4));
#pragma unroll
                for (int tt = 0; tt < 2; tt++)
                    acc2[tt] = mfma(a2, ld8(W2l + (16 * tt + (l & 15)) * kK2 + p * 32 + 8 * (l >> 4)), acc2[tt]);
            }
#pragma unroll
            for (int tt = 0; tt < 2; tt++) {
                const int jj = 16 * tt + (l & 15);
                const float bj = net.theta[kOffB2 + jj];
#pragma unroll
                for (int ii = 0; ii < 4; ii++) H2l[(16 * w + 4 * (l >> 4) + ii) * 33 + jj] = silu(acc2[tt][ii] + bj);
            }
            wave_sync();  // H2l rows of this wave written
            // ---- layer 3: lane l < 48 of wave w gives output l / 16 of column 16w + l % 16
            if (l < 48) {
                const int c = 16 * w + (l & 15), mm = l >> 4;
                float a = net.theta[kOffB3 + mm];
                for (int jj = 0; jj < kN2; jj++) a = fmaf(bfr(H2l[c * 33 + jj]), W3l[mm * kN2 + jj], a);
                est[c][mm] = cnode[c][mm] + kWsScale * (double)a;
            }
            __syncthreads();  // (4) estimates
        }
        if (w == 0 && live) {
            double rx = est[l][0], ry = est[l][1], rz = est[l][2];
            if (!(isfinite(rx) && isfinite(ry) && isfinite(rz))) { rx = tx + ox; ry = ty + oy; rz = tz + oz; }
            float *pp = pred + 3 * ((long long)it.col_off + (long long)i * it.ny + j);
            pp[0] = (float)rx; pp[1] = (float)ry; pp[2] = (float)rz;
        }
    }
}

static void shared_args(DevArgs &A, const ppafm_relax_params &rp, const ppafm_tip &tip)
{
    A = DevArgs{};
    A.max_iters = rp.max_iters;
    A.probe_start = rp.probe_start;
    for (int k = 0; k < 9; k++) A.rot[k] = tip.rot[k];
    A.conv_f2 = rp.conv_f2; A.dtmax = rp.dt; A.a0 = rp.damping;
    A.finc = rp.finc; A.fdec = rp.fdec; A.falpha = rp.falpha;
    A.dtmin = rp.dtmin_frac * rp.dt; A.f2safe = rp.f2safe;
    for (int k = 0; k < 3; k++) { A.ks[k] = tip.kspring[k]; A.d0[k] = tip.d0[k]; }
    A.krad = tip.krad; A.lrad = tip.lrad; A.surf_ff = tip.surf_ff;
}

}  // namespace

int launch_ws_predict(const WsNetView &nv, const RelaxBatchLaunch &L, const long long *wave_prefix, long long nwaves,
                      float *pred, int reps, hipStream_t st)
{
    PP_REQUIRE(reps >= 1 && reps <= 8, "network passes must be 1..8");
    PP_REQUIRE(L.rp.mode == 0 || L.rp.mode == 1, "relax mode must be 0 or 1");
    PP_REQUIRE(nwaves >= 0, "negative column count");
    if (nwaves == 0) return 0;
    DevArgs A;
    shared_args(A, L.rp, L.tip);
    const Net net{nv.theta, nv.W1b, nv.W2b, nullptr};
    int dev = 0, ncu = 0;
    PP_TRY_HIP(hipGetDevice(&dev));
    PP_TRY_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const long long nblk = std::min<long long>(nwaves, 2LL * std::max(ncu, 1));  // persistent: 2 blocks per CU
    dim3 grd((unsigned)nblk), blk(256);
    if (L.f64) {
        if (L.rp.mode == 0) hipLaunchKernelGGL((ws_predict_kernel<double4, 0>), grd, blk, 0, st, A, net, L.items, wave_prefix, L.nitems, nwaves, (const double4 *)L.grids, pred, reps);
        else hipLaunchKernelGGL((ws_predict_kernel<double4, 1>), grd, blk, 0, st, A, net, L.items, wave_prefix, L.nitems, nwaves, (const double4 *)L.grids, pred, reps);
    } else {
        if (L.rp.mode == 0) hipLaunchKernelGGL((ws_predict_kernel<float4, 0>), grd, blk, 0, st, A, net, L.items, wave_prefix, L.nitems, nwaves, (const float4 *)L.grids, pred, reps);
        else hipLaunchKernelGGL((ws_predict_kernel<float4, 1>), grd, blk, 0, st, A, net, L.items, wave_prefix, L.nitems, nwaves, (const float4 *)L.grids, pred, reps);
    }
    return 0;
}

int launch_ws_samples(const RelaxBatchLaunch &L, const long long *pos_prefix, const float *pos, const float *truth,
                      long long nsel, const long long *sel, unsigned short *patches, float *targets, float jitter,
                      unsigned seed, hipStream_t st)
{
    PP_REQUIRE(L.rp.mode == 0 || L.rp.mode == 1, "relax mode must be 0 or 1");
    if (nsel <= 0) return 0;
    DevArgs A;
    shared_args(A, L.rp, L.tip);
    const long long nblk = (nsel + 255) / 256;
    PP_REQUIRE(nblk < (1LL << 31), "too many samples");
    dim3 grd((unsigned)nblk), blk(256);
    if (L.f64) {
        if (L.rp.mode == 0) hipLaunchKernelGGL((ws_samples_kernel<double4, 0>), grd, blk, 0, st, A, L.items, pos_prefix, L.nitems, (const double4 *)L.grids, pos, truth, nsel, sel, patches, targets, jitter, seed);
        else hipLaunchKernelGGL((ws_samples_kernel<double4, 1>), grd, blk, 0, st, A, L.items, pos_prefix, L.nitems, (const double4 *)L.grids, pos, truth, nsel, sel, patches, targets, jitter, seed);
    } else {
        if (L.rp.mode == 0) hipLaunchKernelGGL((ws_samples_kernel<float4, 0>), grd, blk, 0, st, A, L.items, pos_prefix, L.nitems, (const float4 *)L.grids, pos, truth, nsel, sel, patches, targets, jitter, seed);
        else hipLaunchKernelGGL((ws_samples_kernel<float4, 1>), grd, blk, 0, st, A, L.items, pos_prefix, L.nitems, (const float4 *)L.grids, pos, truth, nsel, sel, patches, targets, jitter, seed);
    }
    return 0;
}

namespace {
}  // namespace

// --------------------------------------------------------------- model object

constexpr int kW1Rows = 576;  // (s, p) rows per ws_bwd1b_kernel block: 128 blocks for 8192 samples

struct WsModel {
    float *theta = nullptr, *grad = nullptr, *m1 = nullptr, *m2 = nullptr, *loss = nullptr;
    unsigned short *W1b = nullptr, *W2b = nullptr, *W2bT = nullptr;
    int cap = 0;  // samples of the scratch buffers (multiple of 64)
    float *Z1 = nullptr, *Z2 = nullptr, *H2 = nullptr;
    unsigned short *H1 = nullptr, *H1T = nullptr, *A1T = nullptr, *dZ2 = nullptr, *dZ2T = nullptr, *dZ1T = nullptr;
    unsigned short *Ppad = nullptr;  // padded copy of a partial last minibatch
    float *Tpad = nullptr;
    float *part = nullptr;   // per-block sums of the head's gradient terms
    float *part1 = nullptr;  // per-block sums of dW1 (ws_bwd1b_kernel)
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
    PP_TRY_HIP(hipMalloc((void **)&m.part, ((c + 255) / 256) * kHeadVals * 4));
    PP_TRY_HIP(hipMalloc((void **)&m.part1, (9 * c / kW1Rows) * (kN1 * kK1) * 4));
    return 0;
}

void ws_free(WsModel &m)
{
    void *p[] = {m.theta, m.grad, m.m1, m.m2, m.loss, m.W1b, m.W2b, m.W2bT, m.Z1, m.Z2, m.H2,
                 m.H1, m.H1T, m.A1T, m.dZ2, m.dZ2T, m.dZ1T, m.Ppad, m.Tpad, m.part, m.part1};
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
    const int nhb = (S + 255) / 256;
    hipLaunchKernelGGL(ws_head_kernel, dim3(nhb), dim3(256), 0, st, S, nvalid, inv_n, m.Z2, m.H2, net, T, out,
                       train ? m.part : nullptr, m.dZ2, m.dZ2T);
    if (!train) return 0;
    hipLaunchKernelGGL(ws_head_sum_kernel, dim3(1), dim3(256), 0, st, nhb, m.part, m.grad, m.loss);
    // dW2: K = S split into chunks of up to 512 samples
    int ks = 512;
    while (S % ks) ks >>= 1;
    hipLaunchKernelGGL(ws_bwd2_kernel, dim3(3, S / ks), dim3(64), 0, st, S, ks, m.dZ2T, m.H1T, m.grad);
    static_assert(kK2 / 16 == 18, "ws_bwd1a_kernel splits 18 column slices over 3 blocks");
    hipLaunchKernelGGL(ws_bwd1a_kernel, dim3(S / 64, 3), dim3(64), 0, st, S, m.dZ2, net, m.Z1, m.dZ1T, m.grad);
    // dW1: K = 9 S rows in blocks of kW1Rows (a divisor of 9 S as S is a multiple of 64)
    const long K = 9L * S;
    const int nb1 = (int)(K / kW1Rows);
    hipLaunchKernelGGL(ws_bwd1b_kernel, dim3((unsigned)nb1), dim3(64), 0, st, K, kW1Rows, m.dZ1T, m.A1T, m.part1);
    hipLaunchKernelGGL(ws_w1_sum_kernel, dim3((kN1 * kK1 + 255) / 256), dim3(256), 0, st, nb1, m.part1, m.grad);
    return 0;
}

}  // namespace pp

using namespace pp;

struct ppafm_ws_model {
    WsModel m;
};

namespace pp {
WsNetView ws_view(const ppafm_ws_model *w) { return WsNetView{w->m.theta, w->m.W1b, w->m.W2b}; }
}  // namespace pp

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

int ppafm_ws_get_state(const ppafm_ws_model *w, float *theta, float *m1, float *m2, long long *step)
{
    PP_REQUIRE(w && theta && m1 && m2 && step, "NULL argument");
    PP_TRY_HIP(hipDeviceSynchronize());
    PP_TRY_HIP(hipMemcpy(theta, w->m.theta, kNParam * 4, hipMemcpyDeviceToHost));
    PP_TRY_HIP(hipMemcpy(m1, w->m.m1, kNParam * 4, hipMemcpyDeviceToHost));
    PP_TRY_HIP(hipMemcpy(m2, w->m.m2, kNParam * 4, hipMemcpyDeviceToHost));
    *step = w->m.step;
    return 0;
}

int ppafm_ws_set_state(ppafm_ws_model *w, const float *theta, const float *m1, const float *m2, long long step)
{
    PP_REQUIRE(w && theta && m1 && m2, "NULL argument");
    PP_REQUIRE(step >= 0, "step must be >= 0");
    PP_TRY_HIP(hipMemcpy(w->m.theta, theta, kNParam * 4, hipMemcpyHostToDevice));
    PP_TRY_HIP(hipMemcpy(w->m.m1, m1, kNParam * 4, hipMemcpyHostToDevice));
    PP_TRY_HIP(hipMemcpy(w->m.m2, m2, kNParam * 4, hipMemcpyHostToDevice));
    w->m.step = (long)step;
    ws_refresh(w->m, nullptr);
    PP_TRY_HIP(hipDeviceSynchronize());
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
