// Peak-rate microbenchmarks for the rooflines: FP32 and FP64 vector FMA, and
// bf16 MFMA (v_mfma_f32_16x16x32_bf16).  Each thread runs independent
// dependency chains (16 FMA chains, or 8 MFMA accumulators per wave) long
// enough to hide latency, at 16 waves per CU; the values stay bounded
// (x <- 0.999 x + 0.001 tends to 1) and every thread stores its sum, so
// nothing is optimised away.  No counterpart in the reference.

#include "pp_common.hpp"

namespace pp {

namespace {

constexpr int kChains = 16;
constexpr int kAcc = 8;
constexpr int kThreadsPeak = 256;

typedef __bf16 bf16x8p __attribute__((ext_vector_type(8)));
typedef float f32x4p __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(kThreadsPeak) void fma32_peak_kernel(float *out, int iters)
{
    float a[kChains];
    const float base = (float)(threadIdx.x & 63) * 1e-3f;
#pragma unroll
    for (int k = 0; k < kChains; k++) a[k] = base + (float)k * 1e-2f;
    for (int it = 0; it < iters; it++)
#pragma unroll
        for (int k = 0; k < kChains; k++) a[k] = __fmaf_rn(a[k], 0.999f, 0.001f);
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < kChains; k++) s += a[k];
    out[(long)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(kThreadsPeak) void fma64_peak_kernel(double *out, int iters)
{
    double a[kChains];
    const double base = (double)(threadIdx.x & 63) * 1e-3;
#pragma unroll
    for (int k = 0; k < kChains; k++) a[k] = base + (double)k * 1e-2;
    for (int it = 0; it < iters; it++)
#pragma unroll
        for (int k = 0; k < kChains; k++) a[k] = __fma_rn(a[k], 0.999, 0.001);
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < kChains; k++) s += a[k];
    out[(long)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(kThreadsPeak) void mfma_peak_kernel(float *out, int iters)
{
    // operands: small bf16 values (0x3c00.. ~ 0.0078), different per lane
    uint4 u;
    const unsigned h = 0x3c00u + (threadIdx.x & 15);
    u.x = u.y = u.z = u.w = h | (h << 16);
    const bf16x8p a = __builtin_bit_cast(bf16x8p, u);
    const bf16x8p b = __builtin_bit_cast(bf16x8p, u);
    f32x4p acc[kAcc];
#pragma unroll
    for (int k = 0; k < kAcc; k++) acc[k] = f32x4p{0.f, 0.f, 0.f, 0.f};
    for (int it = 0; it < iters; it++)
#pragma unroll
        for (int k = 0; k < kAcc; k++) acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[k], 0, 0, 0);
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < kAcc; k++) s += acc[k][0] + acc[k][1] + acc[k][2] + acc[k][3];
    out[(long)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace

}  // namespace pp

using namespace pp;

extern "C" int ppafm_measure_peak(int kind, int iters, double *tflops)
{
    PP_REQUIRE(kind >= 0 && kind <= 2, "kind must be 0 (FP32 FMA), 1 (FP64 FMA) or 2 (bf16 MFMA)");
    PP_REQUIRE(iters >= 1 && tflops, "iters must be >= 1 and tflops not NULL");
    int dev = 0, ncu = 0;
    PP_TRY_HIP(hipGetDevice(&dev));
    PP_TRY_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int blocks = ncu * 4;  // 4 waves per block: 16 waves per CU
    const long nthreads = (long)blocks * kThreadsPeak;
    void *out = nullptr;
    PP_TRY_HIP(hipMalloc(&out, nthreads * sizeof(double)));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    float ms = 0.f;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        set_error("hipEventCreate failed");
        rc = -1;
    }
    for (int rep = 0; rep < 2 && rc == 0; rep++) {  // first launch warms up
        if (hipEventRecord(e0, nullptr) != hipSuccess) { rc = -1; break; }
        if (kind == 0) hipLaunchKernelGGL(fma32_peak_kernel, dim3(blocks), dim3(kThreadsPeak), 0, nullptr, (float *)out, iters);
        else if (kind == 1) hipLaunchKernelGGL(fma64_peak_kernel, dim3(blocks), dim3(kThreadsPeak), 0, nullptr, (double *)out, iters);
        else hipLaunchKernelGGL(mfma_peak_kernel, dim3(blocks), dim3(kThreadsPeak), 0, nullptr, (float *)out, iters);
        if (hipGetLastError() != hipSuccess || hipEventRecord(e1, nullptr) != hipSuccess ||
            hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            set_error("peak kernel launch or timing failed");
            rc = -1;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(out);
    if (rc) return rc;
    double flop;
    if (kind < 2) flop = 2.0 * kChains * (double)iters * (double)nthreads;
    else flop = 2.0 * 16 * 16 * 32 * kAcc * (double)iters * (double)(nthreads / 64);
    *tflops = flop / (ms * 1e-3) / 1e12;
    return 0;
}
