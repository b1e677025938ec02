// Valid-mode z correlation for Fz -> df.
//
// convolveZ (ppafm/ocl/cl/relax.cl:571-608) correlates each pixel column of
// the force stack with the df weights; Fz2df (ppafm/common.py:421-430) is
// the same with the weights reversed (np.convolve 'valid').  One thread per
// output value, weights in the kernel argument block, FP64 accumulation.
// HBM bound and tiny next to the relaxation; kept separate so it can read
// any channel of any force layout.

#include <algorithm>

#include "pp_common.hpp"

// Explicit __fmaf_rn where a fused multiply-add is wanted; no implicit
// contraction, so FP64 expressions round like the CPU reference.
#pragma clang fp contract(off)

namespace pp {

namespace {

constexpr int kMaxW = 64;
// LDS staging of the df convolution: about this many inputs per block, and
// columns longer than the byte budget use the unstaged kernel.
constexpr int kConvLdsInputs = 2048;
constexpr size_t kConvLdsBytes = 32768;

struct ConvArgs {
    long ncol;
    int nz, nzout, nw;
    int in_stride, in_step, in_offset;
    double scale;
    double w[kMaxW];
};

template <typename T>
__global__ __launch_bounds__(256) void convolve_z_kernel(ConvArgs A, const T *__restrict__ in, T *__restrict__ out)
{
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = A.ncol * A.nzout;
    if (gid >= total) return;
    const long c = gid / A.nzout;
    const int j = (int)(gid % A.nzout);
    const T *p = in + c * A.in_stride + A.in_offset + (long)j * A.in_step;
    double acc = 0;
    for (int k = 0; k < A.nw; k++) acc += (double)p[(long)k * A.in_step] * A.w[k];
    out[gid] = (T)(acc * A.scale);
}

// The same convolution with a block's columns staged in LDS: the block reads
// its columns' inputs once, with neighbouring lanes on neighbouring inputs,
// and every output then sums its nw terms from LDS in the order above, so
// the results are the same. The kernel above reads each input nw times
// through the vector cache.
template <typename T>
__global__ __launch_bounds__(256) void convolve_z_lds_kernel(ConvArgs A, int cols_per_block, const T *__restrict__ in,
                                                             T *__restrict__ out)
{
    extern __shared__ unsigned char conv_lds[];
    T *s = (T *)conv_lds;
    const long c0 = (long)blockIdx.x * cols_per_block;
    const int nc = (int)min((long)cols_per_block, A.ncol - c0);
    const int nin = nc * A.nz;
    const T *src = in + c0 * A.in_stride + A.in_offset;
    for (int i = threadIdx.x; i < nin; i += blockDim.x) {
        const int c = i / A.nz, z = i - c * A.nz;
        s[i] = src[(long)c * A.in_stride + (long)z * A.in_step];
    }
    __syncthreads();
    const int nout = nc * A.nzout;
    T *o = out + c0 * A.nzout;
    for (int i = threadIdx.x; i < nout; i += blockDim.x) {
        const int c = i / A.nzout, j = i - c * A.nzout;
        const T *p = s + c * A.nz + j;
        double acc = 0;
        for (int k = 0; k < A.nw; k++) acc += (double)p[k] * A.w[k];
        o[i] = (T)(acc * A.scale);
    }
}

// io.limit_vec_field (io.py:766-778) on a float64 (npts,3) device array.
__global__ __launch_bounds__(256) void limit_vec_kernel(double *__restrict__ F, long npts, double fmax)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    const double x = F[3 * i], y = F[3 * i + 1], z = F[3 * i + 2];
    const double fr = sqrt(x * x + y * y + z * z);
    if (fr > fmax) {
        const double s = fmax / fr;
        F[3 * i] = x * s; F[3 * i + 1] = y * s; F[3 * i + 2] = z * s;
    }
}

// Plain device copy (float4 granularity), used to calibrate HBM byte counters.
__global__ __launch_bounds__(256) void copy_kernel(const float4 *__restrict__ src, float4 *__restrict__ dst, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) dst[i] = src[i];
}

// Fill n 32-bit words with a value: clears counters and queues inside a
// captured graph, where memset nodes are avoided.
__global__ __launch_bounds__(64) void fill_u32_kernel(unsigned *__restrict__ p, long n, unsigned v)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}

}  // namespace

int launch_fill_u32(void *p, long n, unsigned v, hipStream_t st)
{
    PP_REQUIRE(n >= 0, "size must be >= 0");
    if (n == 0) return 0;
    const long nblk = std::min<long>((n + 63) / 64, 1024L);
    hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)nblk), dim3(64), 0, st, (unsigned *)p, n, v);
    return 0;
}

int launch_copy(const void *src, void *dst, long n16, hipStream_t st)
{
    PP_REQUIRE(n16 >= 0, "size must be >= 0");
    if (n16 == 0) return 0;
    const long nblk = std::min<long>((n16 + 255) / 256, 256L * 64);
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)nblk), dim3(256), 0, st, (const float4 *)src, (float4 *)dst, n16);
    return 0;
}

int launch_limit_vec_field(double *F, long npts, double fmax, hipStream_t st)
{
    PP_REQUIRE(npts >= 0, "npts must be >= 0");
    if (npts == 0) return 0;
    const long nblk = (npts + 255) / 256;
    PP_REQUIRE(nblk < (1L << 31), "array too large");
    hipLaunchKernelGGL(limit_vec_kernel, dim3((unsigned)nblk), dim3(256), 0, st, F, npts, fmax);
    return 0;
}

int launch_convolve_z(long ncol, int nz, int in_stride, int in_step, int in_offset, const void *in, int nw,
                      const double *w, double scale, void *out, int f64, hipStream_t st)
{
    PP_REQUIRE(nw >= 1 && nw <= kMaxW, "number of df weights must be 1..64");
    PP_REQUIRE(nz >= nw, "nz must be >= the number of weights");
    PP_REQUIRE(ncol >= 0, "ncol must be >= 0");
    PP_REQUIRE(in_step >= 1 && in_offset >= 0 && in_stride >= (long)(nz - 1) * in_step + in_offset + 1,
               "input strides inconsistent with nz");
    ConvArgs A;
    A.ncol = ncol; A.nz = nz; A.nzout = nz - nw + 1; A.nw = nw;
    A.in_stride = in_stride; A.in_step = in_step; A.in_offset = in_offset;
    A.scale = scale;
    for (int k = 0; k < kMaxW; k++) A.w[k] = k < nw ? w[k] : 0.0;
    const long total = ncol * A.nzout;
    if (total == 0) return 0;
    const size_t esz = f64 ? sizeof(double) : sizeof(float);
    if ((size_t)nz * esz <= kConvLdsBytes) {
        // Columns per block: about 2048 inputs, within the LDS budget.
        const int cpb = (int)std::max<long>(1, std::min<long>(ncol, kConvLdsInputs / nz));
        const long nb = (ncol + cpb - 1) / cpb;
        PP_REQUIRE(nb < (1L << 31), "too many columns");
        const size_t lds = (size_t)cpb * nz * esz;
        if (f64)
            hipLaunchKernelGGL(convolve_z_lds_kernel<double>, dim3((unsigned)nb), dim3(256), lds, st, A, cpb,
                               (const double *)in, (double *)out);
        else
            hipLaunchKernelGGL(convolve_z_lds_kernel<float>, dim3((unsigned)nb), dim3(256), lds, st, A, cpb,
                               (const float *)in, (float *)out);
        return 0;
    }
    const long nblk = (total + 255) / 256;
    PP_REQUIRE(nblk < (1L << 31), "too many columns");
    if (f64)
        hipLaunchKernelGGL(convolve_z_kernel<double>, dim3((unsigned)nblk), dim3(256), 0, st, A,
                           (const double *)in, (double *)out);
    else
        hipLaunchKernelGGL(convolve_z_kernel<float>, dim3((unsigned)nblk), dim3(256), 0, st, A,
                           (const float *)in, (float *)out);
    return 0;
}

}  // namespace pp
