// Shared pieces of the force-field grid kernels (ff_grid.hip: FP64 grids
// and the mixed FP32 kernel; ff_grid_f32.hip: the FP32 kernel): the brick
// geometry, the kernel arguments, the atom records staged in LDS, the FP64
// pair of the reference's operation order and the batch lookup.
#pragma once

#include "pp_common.hpp"

#pragma clang fp contract(off)

namespace pp {

constexpr int kColsX = 4, kColsY = 4;
constexpr int kZLanes = 16;
constexpr int kK = 4;  // z points per thread, strided by kZLanes
constexpr int kZBrick = kZLanes * kK;
constexpr int kThreads = kColsX * kColsY * kZLanes;  // 256
constexpr int kTile = 128;                           // atoms per LDS tile
constexpr int kFold = 32;                            // FP32 partial-sum length
constexpr double kMinFp64Radius = 0.2;               // r^-13 overflows FP32 below ~0.04 A

struct FFArgs {
    int n[3];
    int nbx, nby, nbz;
    int natoms;
    int ns;
    double dcell[9];
    float sw[4], sdz[4];
    float swsum, swdz;  // sum of the stencil weights, of weight x z offset
    float r2safe;
    double r2safe_d;
    float near2;  // pairs with |d|^2 below this go to FP64
    int mode;
    int with_lj;
    double probe_q;
    double fmax;
};

struct AtomLds {
    float4 hi;  // x, y, z (hi part, brick-local), charge * coulomb constant
    float4 lo;  // x, y, z (lo part), unused
    float4 lj;  // 6*c6, 12*c12, c6, c12
};

__device__ __forceinline__ float r2_f32(float dx, float dy, float dz, float r2safe)
{
    return __fmaf_rn(dx, dx, __fmaf_rn(dy, dy, __fmaf_rn(dz, dz, r2safe)));
}

__device__ __forceinline__ void clamp_vec(double &x, double &y, double &z, double fmax)
{
    if (fmax > 0) {
        double fr = sqrt(x * x + y * y + z * z);
        if (fr > fmax) {
            double s = fmax / fr;
            x *= s; y *= s; z *= s;
        }
    }
}

// FP64 contribution of one atom to one point: LJ at d and the Coulomb
// stencil at d + (0, 0, sdz[s]); ProbeParticle.cpp:262-278 / Forces.h:120-132,
// 237-244 in the order of the oracle.
template <bool LJ, int NS, bool WE>
__device__ __forceinline__ void pair_fp64(const FFArgs &A, double dx, double dy, double dz, double q, double c6,
                                          double c12, double aL[4], double aE[4])
{
    if (LJ) {
        const double ir2 = 1.0 / ((dx * dx + dy * dy + dz * dz) + A.r2safe_d);
        const double ir6 = ir2 * ir2 * ir2;
        const double e6 = c6 * ir6, e12 = c12 * ir6 * ir6;
        const double s = (12 * e12 - 6 * e6) * ir2;
        aL[0] += dx * s; aL[1] += dy * s; aL[2] += dz * s;
        if (WE) aL[3] += e12 - e6;
    }
#pragma unroll
    for (int s_ = 0; s_ < NS; s_++) {
        const double dzs = dz + (double)A.sdz[s_];
        const double ir2 = 1.0 / ((dx * dx + dy * dy + dzs * dzs) + A.r2safe_d);
        const double e = sqrt(ir2) * (q * (double)A.sw[s_]);
        const double s = e * ir2;
        aE[0] += dx * s; aE[1] += dy * s; aE[2] += dzs * s;
        if (WE) aE[3] += e;
    }
}

// ---- batches: one launch over the bricks of every molecule.  Block b finds
// its molecule by binary search over the brick prefix sums.
__device__ __forceinline__ int find_item(const long long *__restrict__ prefix, int nitems, long long x)
{
    int lo = 0, hi = nitems;  // prefix[lo] <= x < prefix[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ FFArgs batch_args(const FFArgs &A0, const BatchGridItem &it)
{
    FFArgs A = A0;
    for (int k = 0; k < 3; k++) A.n[k] = it.n[k];
    A.nbx = (it.n[0] + kColsX - 1) / kColsX;
    A.nby = (it.n[1] + kColsY - 1) / kColsY;
    A.nbz = (it.n[2] + kZBrick - 1) / kZBrick;
    A.natoms = it.natoms;
    for (int k = 0; k < 9; k++) A.dcell[k] = it.dcell[k];
    return A;
}


// One launch: a single grid (batch == nullptr) or every grid of a batch.
struct BatchLaunch {
    const BatchGridItem *items;
    const long long *prefix;
    int nitems;
    long long nbricks;
    bool qskip;  // some atoms are uncharged: skip their Coulomb terms
    bool ortho;  // every grid has orthogonal cell axes
};

// The FP32 grid kernel (ff_grid_f32.hip) for the kernel arguments A, on one
// grid (batch == nullptr) or on every grid of a batch.  z0: stencil point 0
// sits at zero z offset; ortho: the cell axes are orthogonal.
int launch_ff_f32(const FFArgs &A, bool z0, bool with_lj, bool with_energy, const AtomRec *atoms, void *out,
                  const BatchLaunch *batch, bool ortho, hipStream_t st);

}  // namespace pp
