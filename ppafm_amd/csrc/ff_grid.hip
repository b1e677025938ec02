// Force-field grid build on gfx950.
//
// Pair terms follow ppafm/cpp/Forces.h:120-132 (addAtomLJ) and :237-244
// (addAtomCoulomb) with the finite-difference probe stencils of
// ProbeParticle.cpp:193-209, or ppafm/ocl/cl/FF.cl:115-130 (getLJ,
// getCoulomb) for the OpenCL-mode model; the grid iteration replaces
// iterateGrid3D_omp (Grid.h:142-170) and evalLJC_QZs_noPos (FF.cl:366-423).
//
// Kernel ff_build_f32 (device path, FP32 VALU bound):
//   * a block owns a brick of 4x4 grid columns x 64 z points; lanes run along
//     z so every float4 store instruction writes 4 runs of 16 consecutive
//     points (z fastest);
//   * atoms are staged in LDS per tile; every lane reads the same atom
//     (LDS broadcast) and applies it to its 4 points;
//   * coordinates are relative to the brick centre and atom offsets are kept
//     as FP32 hi + lo, so the pair distance carries FP32 relative accuracy
//     even 40 A away from the grid origin;
//   * pair math is FP32 (one v_rsq per pair serves LJ and Coulomb), partial
//     sums are folded into FP64 accumulators every 32 atoms;
//   * pairs closer than 0.2 A (where r^-13 overflows FP32) are skipped in
//     the FP32 pass and added in FP64 by a second pass that runs only for
//     the flagged points;
//   * the epilogue applies the per-field clamp of io.limit_vec_field
//     (io.py:766-778) and combines FF = clamp(LJ) + q clamp(El)
//     (HighLevel.py:143-145) before the single FP32 store.
//
// The host ABI's FP64 accumulation into the caller's grids is in
// ff_models.hip.

#include "ff_grid.hpp"

// Explicit __fmaf_rn where a fused multiply-add is wanted; no implicit
// contraction, so FP64 expressions round like the CPU reference.
#pragma clang fp contract(off)

namespace pp {

namespace {


// LJ: include the LJ term.  NS: number of Coulomb stencil points.  Z0:
// stencil point 0 sits at zero z offset and reuses the LJ pair's rsq.
// WE: accumulate the energy channel.
template <bool LJ, int NS, bool Z0, bool WE, bool QS = false>
__device__ __forceinline__ void ff_brick_f32(const FFArgs &A, const AtomRec *__restrict__ atoms, float4 *__restrict__ out,
                                             int b, AtomLds *lds, double4 *ldsd)
{
    const int t = threadIdx.x;
    const int col = t / kZLanes;
    const int zl = t % kZLanes;
    const int bz = b % A.nbz; b /= A.nbz;
    const int by = b % A.nby;
    const int bx = b / A.nby;
    const int ix = bx * kColsX + (col % kColsX);
    const int iy = by * kColsY + (col / kColsX);
    const int z0 = bz * kZBrick;

    // brick centre (fractional grid index) and its Cartesian position
    const double cx = bx * kColsX + 0.5 * (kColsX - 1);
    const double cy = by * kColsY + 0.5 * (kColsY - 1);
    const double cz = z0 + 0.5 * (kZBrick - 1);
    const double *D = A.dcell;
    const double ccx = D[0] * cx + D[3] * cy + D[6] * cz;
    const double ccy = D[1] * cx + D[4] * cy + D[7] * cz;
    const double ccz = D[2] * cx + D[5] * cy + D[8] * cz;
    const double fx = ix - cx, fy = iy - cy;

    // this lane's points, brick-local
    float px[kK], py[kK], pz[kK];
#pragma unroll
    for (int k = 0; k < kK; k++) {
        const double fz = (z0 + zl + kZLanes * k) - cz;
        px[k] = (float)(D[0] * fx + D[3] * fy + D[6] * fz);
        py[k] = (float)(D[1] * fx + D[4] * fy + D[7] * fz);
        pz[k] = (float)(D[2] * fx + D[5] * fy + D[8] * fz);
    }
    // z sub-rows holding at least one valid point (uniform per block)
    const int nk = min(kK, (A.n[2] - z0 + kZLanes - 1) / kZLanes);
    const float near2s = A.near2 + A.r2safe;

    double aL[kK][4], aE[kK][4];  // FP64 accumulators Fx, Fy, Fz, E of the LJ and El fields
#pragma unroll
    for (int k = 0; k < kK; k++)
#pragma unroll
        for (int c = 0; c < 4; c++) { aL[k][c] = 0; aE[k][c] = 0; }

    for (int i0 = 0; i0 < A.natoms; i0 += kTile) {
        const int nt = min(kTile, A.natoms - i0);
        __syncthreads();
        for (int j = t; j < nt; j += kThreads) {
            const AtomRec a = atoms[i0 + j];
            const double rx = a.x - ccx, ry = a.y - ccy, rz = a.z - ccz;
            const float hx = (float)rx, hy = (float)ry, hz = (float)rz;
            AtomLds r;
            r.hi = make_float4(hx, hy, hz, (float)a.q);
            r.lo = make_float4((float)(rx - hx), (float)(ry - hy), (float)(rz - hz), 0.f);
            r.lj = make_float4((float)(6.0 * a.c6), (float)(12.0 * a.c12), (float)a.c6, (float)a.c12);
            lds[j] = r;
            ldsd[j] = make_double4(a.q, a.c6, a.c12, 0.0);
        }
        __syncthreads();

        for (int j0 = 0; j0 < nt; j0 += kFold) {
            const int j1 = min(nt, j0 + kFold);
            float sL[kK][4], sE[kK][4];
#pragma unroll
            for (int k = 0; k < kK; k++)
#pragma unroll
                for (int c = 0; c < 4; c++) { sL[k][c] = 0.f; sE[k][c] = 0.f; }

            for (int j = j0; j < j1; j++) {
                const float4 hi = lds[j].hi;
                const float4 lo = lds[j].lo;
                const float4 lj = lds[j].lj;
#pragma unroll
                for (int k = 0; k < kK; k++) {
                    if (k >= nk) break;
                    const float dx = (px[k] - hi.x) - lo.x;
                    const float dy = (py[k] - hi.y) - lo.y;
                    const float dz = (pz[k] - hi.z) - lo.z;
                    const float r2c = r2_f32(dx, dy, dz, A.r2safe);
                    const bool nr = r2c < near2s;
                    if (nr) {
                        // pair inside the FP64 radius: exact brick-local offsets
                        const double fz = (z0 + zl + kZLanes * k) - cz;
                        const double qx = D[0] * fx + D[3] * fy + D[6] * fz;
                        const double qy = D[1] * fx + D[4] * fy + D[7] * fz;
                        const double qz = D[2] * fx + D[5] * fy + D[8] * fz;
                        const double4 ad = ldsd[j];
                        pair_fp64<LJ, NS, WE>(A, qx - ((double)hi.x + (double)lo.x), qy - ((double)hi.y + (double)lo.y),
                                              qz - ((double)hi.z + (double)lo.z), ad.x, ad.y, ad.z, aL[k], aE[k]);
                    }
                    float ir0 = 0.f, ir20 = 0.f;
                    if (LJ || Z0) {
                        ir0 = __builtin_amdgcn_rsqf(nr ? 1.0e30f : r2c);
                        ir20 = ir0 * ir0;
                    }
                    if (LJ) {
                        const float ir6 = ir20 * ir20 * ir20;
                        const float s = __fmaf_rn(lj.y, ir6, -lj.x) * ir6 * ir20;
                        sL[k][0] = __fmaf_rn(dx, s, sL[k][0]);
                        sL[k][1] = __fmaf_rn(dy, s, sL[k][1]);
                        sL[k][2] = __fmaf_rn(dz, s, sL[k][2]);
                        if (WE) sL[k][3] = __fmaf_rn(__fmaf_rn(lj.w, ir6, -lj.z), ir6, sL[k][3]);
                    }
                    // QS: uniform over the wave (one atom for all lanes); an
                    // uncharged atom's Coulomb terms are exact zeros
                    if (!QS || hi.w != 0.f)
#pragma unroll
                    for (int s_ = 0; s_ < NS; s_++) {
                        float ir, ir2, dzs;
                        if (Z0 && s_ == 0) {
                            ir = ir0; ir2 = ir20; dzs = dz;
                        } else {
                            dzs = dz + A.sdz[s_];
                            const float r2 = r2_f32(dx, dy, dzs, A.r2safe);
                            ir = __builtin_amdgcn_rsqf(nr ? 1.0e30f : r2);
                            ir2 = ir * ir;
                        }
                        const float e = (hi.w * A.sw[s_]) * ir;
                        const float s = e * ir2;
                        sE[k][0] = __fmaf_rn(dx, s, sE[k][0]);
                        sE[k][1] = __fmaf_rn(dy, s, sE[k][1]);
                        sE[k][2] = __fmaf_rn(dzs, s, sE[k][2]);
                        if (WE) sE[k][3] += e;
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < kK; k++)
#pragma unroll
                for (int c = 0; c < 4; c++) { aL[k][c] += (double)sL[k][c]; aE[k][c] += (double)sE[k][c]; }
        }
    }

    if (ix >= A.n[0] || iy >= A.n[1]) return;
#pragma unroll
    for (int k = 0; k < kK; k++) {
        const int iz = z0 + zl + kZLanes * k;
        if (iz >= A.n[2]) break;
        double gx, gy, gz, e;
        if (A.mode == 0) {
            clamp_vec(aL[k][0], aL[k][1], aL[k][2], A.fmax);
            clamp_vec(aE[k][0], aE[k][1], aE[k][2], A.fmax);
            gx = aL[k][0] + A.probe_q * aE[k][0];
            gy = aL[k][1] + A.probe_q * aE[k][1];
            gz = aL[k][2] + A.probe_q * aE[k][2];
            e = aL[k][3] + A.probe_q * aE[k][3];
        } else {
            gx = aL[k][0] + aE[k][0];
            gy = aL[k][1] + aE[k][1];
            gz = aL[k][2] + aE[k][2];
            e = aL[k][3] + aE[k][3];
        }
        const long ig = ((long)ix * A.n[1] + iy) * A.n[2] + iz;
        out[ig] = make_float4((float)gx, (float)gy, (float)gz, WE ? (float)e : 0.f);
    }
}

// 2/sqrt(x): v_rsq_f64 (measured on gfx950: 5.2e-8 relative) refined by
// one Newton step, y (3 - x y^2), which is twice the step y + y (1/2 - x y^2 / 2)
// to 1/sqrt(x) (4.3e-15 relative) in three FP64 instructions instead of four
__device__ __forceinline__ double rsq2_nr(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    return y * __fma_rn(-(x * y), y, 3.0);
}

// FP64 pair with reciprocal square roots in place of IEEE divide and sqrt;
// the Coulomb point at zero z offset shares the LJ pair's rsq.  ORTHO (grid
// cells with orthogonal axes): dxy2 = dx^2 + dy^2 + R2SAFE is the same for
// all z points of a lane and comes from the caller.
// The pair works with Y = 2/r (rsq2_nr), and the factors of 2 go into the
// staged coefficients and the sums (exact powers of two, undone by the
// caller): q arrives as q/2 and c12 as 12 c12 / 64, c6 as 6 c6; then
//   LJ:      256 s = fma(12 c12 / 64, Y^6, -6 c6) Y^8   (s = (12 c12 r^-12 - 6 c6 r^-6) r^-2)
//            768 E = Y^6 fma(12 c12 / 64, Y^6, -12 c6)    (E = c12 r^-12 - c6 r^-6)
//   Coulomb: e = (q/2) sw Y = q sw / r,  4 e r^-2 = e Y^2
// so aL[0..2] sum 256 F_LJ, aL[3] 768 E_LJ, aE[0..2] 4 F_C and aE[3] E_C.
template <bool LJ, int NS, bool Z0, bool WE, bool QS = false, bool ORTHO = false>
__device__ __forceinline__ void pair_fp64_fast(const FFArgs &A, double dx, double dy, double dz, double q, double c6,
                                               double c12, double aL[4], double aE[4], double dxy2 = 0.0)
{
    const double r2 = ORTHO ? __fma_rn(dz, dz, dxy2) : __fma_rn(dx, dx, __fma_rn(dy, dy, __fma_rn(dz, dz, A.r2safe_d)));
    double ir0 = 0, ir20 = 0;  // Y = 2/r and Y^2
    if (LJ || Z0) {
        ir0 = rsq2_nr(r2);
        ir20 = ir0 * ir0;
    }
    if (LJ) {
        const double ir6 = ir20 * ir20 * ir20;
        const double s = __fma_rn(c12, ir6, -c6) * (ir6 * ir20);
        aL[0] = __fma_rn(dx, s, aL[0]); aL[1] = __fma_rn(dy, s, aL[1]); aL[2] = __fma_rn(dz, s, aL[2]);
        if (WE) aL[3] = __fma_rn(ir6, __fma_rn(c12, ir6, -2.0 * c6), aL[3]);
    }
    // QS (batches holding uncharged atoms): every lane of the brick reads the
    // same atom, so this branch is uniform; an uncharged atom's Coulomb terms
    // are exact zeros and are skipped
    if (!QS || q != 0.0)
#pragma unroll
    for (int s_ = 0; s_ < NS; s_++) {
        double ir, ir2, dzs;
        if (Z0 && s_ == 0) {
            ir = ir0; ir2 = ir20; dzs = dz;
        } else {
            dzs = dz + (double)A.sdz[s_];
            ir = rsq2_nr(ORTHO ? __fma_rn(dzs, dzs, dxy2) : __fma_rn(dx, dx, __fma_rn(dy, dy, __fma_rn(dzs, dzs, A.r2safe_d))));
            ir2 = ir * ir;
        }
        // one Coulomb point (NS == 1): q arrives multiplied by sw[0] already
        const double e = ir * (NS == 1 ? q : q * (double)A.sw[s_]);
        const double s = e * ir2;
        aE[0] = __fma_rn(dx, s, aE[0]); aE[1] = __fma_rn(dy, s, aE[1]); aE[2] = __fma_rn(dzs, s, aE[2]);
        if (WE) aE[3] += e;
    }
}

// Exact variant: every pair in FP64 (pair_fp64) and a double4 grid, for
// callers that need the grid to FP64 accuracy (the FIRE stop rule turns
// grid perturbations of ~1e-7 into stopping-point changes of >1e-3 A on
// ill-conditioned strokes).  Same brick mapping as ff_build_f32.
template <bool LJ, int NS, bool Z0, bool WE, bool QS = false, bool ORTHO = false>
__device__ __forceinline__ void ff_brick_f64(const FFArgs &A, const AtomRec *__restrict__ atoms, double4 *__restrict__ out,
                                             int b, double4 *lpos, double2 *llj)
{
    const int t = threadIdx.x;
    const int col = t / kZLanes;
    const int zl = t % kZLanes;
    const int bz = b % A.nbz; b /= A.nbz;
    const int by = b % A.nby;
    const int bx = b / A.nby;
    const int ix = bx * kColsX + (col % kColsX);
    const int iy = by * kColsY + (col / kColsX);
    const int z0 = bz * kZBrick;
    const double cx = bx * kColsX + 0.5 * (kColsX - 1);
    const double cy = by * kColsY + 0.5 * (kColsY - 1);
    const double cz = z0 + 0.5 * (kZBrick - 1);
    const double *D = A.dcell;
    const double ccx = D[0] * cx + D[3] * cy + D[6] * cz;
    const double ccy = D[1] * cx + D[4] * cy + D[7] * cz;
    const double ccz = D[2] * cx + D[5] * cy + D[8] * cz;
    const double fx = ix - cx, fy = iy - cy;
    double px[kK], py[kK], pz[kK];
#pragma unroll
    for (int k = 0; k < kK; k++) {
        const double fz = (z0 + zl + kZLanes * k) - cz;
        px[k] = D[0] * fx + D[3] * fy + D[6] * fz;
        py[k] = D[1] * fx + D[4] * fy + D[7] * fz;
        pz[k] = D[2] * fx + D[5] * fy + D[8] * fz;
    }
    const int nk = min(kK, (A.n[2] - z0 + kZLanes - 1) / kZLanes);
    double aL[kK][4], aE[kK][4];
#pragma unroll
    for (int k = 0; k < kK; k++)
#pragma unroll
        for (int c = 0; c < 4; c++) { aL[k][c] = 0; aE[k][c] = 0; }

    for (int i0 = 0; i0 < A.natoms; i0 += kTile) {
        const int nt = min(kTile, A.natoms - i0);
        __syncthreads();
        for (int j = t; j < nt; j += kThreads) {
            const AtomRec a = atoms[i0 + j];
            lpos[j] = make_double4(a.x - ccx, a.y - ccy, a.z - ccz, 0.5 * (NS == 1 ? a.q * (double)A.sw[0] : a.q));
            llj[j] = make_double2(6.0 * a.c6, (12.0 * a.c12) * (1.0 / 64.0));  // the scaling pair_fp64_fast expects
        }
        __syncthreads();
        if constexpr (ORTHO) {
            // orthogonal cell axes: the lane's points share x and y, so one
            // dx, dy and dx^2 + dy^2 per atom.  Bricks with all kK points in
            // the grid (all but the top ones) take a loop without the
            // per-point test, two atoms per pass
            if (nk == kK) {
#pragma unroll 2
                for (int j = 0; j < nt; j++) {
                    const double4 a = lpos[j];
                    const double2 c = llj[j];
                    const double dx = px[0] - a.x, dy = py[0] - a.y;
                    const double dxy2 = __fma_rn(dx, dx, __fma_rn(dy, dy, A.r2safe_d));
#pragma unroll
                    for (int k = 0; k < kK; k++)
                        pair_fp64_fast<LJ, NS, Z0, WE, QS, true>(A, dx, dy, pz[k] - a.z, a.w, c.x, c.y, aL[k], aE[k], dxy2);
                }
            } else {
                for (int j = 0; j < nt; j++) {
                    const double4 a = lpos[j];
                    const double2 c = llj[j];
                    const double dx = px[0] - a.x, dy = py[0] - a.y;
                    const double dxy2 = __fma_rn(dx, dx, __fma_rn(dy, dy, A.r2safe_d));
#pragma unroll
                    for (int k = 0; k < kK; k++) {
                        if (k >= nk) break;
                        pair_fp64_fast<LJ, NS, Z0, WE, QS, true>(A, dx, dy, pz[k] - a.z, a.w, c.x, c.y, aL[k], aE[k], dxy2);
                    }
                }
            }
        } else {
            for (int j = 0; j < nt; j++) {
                const double4 a = lpos[j];
                const double2 c = llj[j];
#pragma unroll
                for (int k = 0; k < kK; k++) {
                    if (k >= nk) break;
                    pair_fp64_fast<LJ, NS, Z0, WE, QS>(A, px[k] - a.x, py[k] - a.y, pz[k] - a.z, a.w, c.x, c.y, aL[k], aE[k]);
                }
            }
        }
    }

    if (ix >= A.n[0] || iy >= A.n[1]) return;
#pragma unroll
    for (int k = 0; k < kK; k++) {
        const int iz = z0 + zl + kZLanes * k;
        if (iz >= A.n[2]) break;
        // the pairs summed 256 F_LJ, 768 E_LJ and 4 F_C (pair_fp64_fast)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            aL[k][c] *= 1.0 / 256.0;
            aE[k][c] *= 0.25;
        }
        if (WE) aL[k][3] *= 1.0 / 768.0;
        double gx, gy, gz, e;
        if (A.mode == 0) {
            clamp_vec(aL[k][0], aL[k][1], aL[k][2], A.fmax);
            clamp_vec(aE[k][0], aE[k][1], aE[k][2], A.fmax);
            gx = aL[k][0] + A.probe_q * aE[k][0];
            gy = aL[k][1] + A.probe_q * aE[k][1];
            gz = aL[k][2] + A.probe_q * aE[k][2];
            e = aL[k][3] + A.probe_q * aE[k][3];
        } else {
            gx = aL[k][0] + aE[k][0];
            gy = aL[k][1] + aE[k][1];
            gz = aL[k][2] + aE[k][2];
            e = aL[k][3] + aE[k][3];
        }
        out[((long)ix * A.n[1] + iy) * A.n[2] + iz] = make_double4(gx, gy, gz, WE ? e : 0.0);
    }
}

template <bool LJ, int NS, bool Z0, bool WE>
__global__ __launch_bounds__(kThreads) void ff_build_f32(FFArgs A, const AtomRec *__restrict__ atoms,
                                                         float4 *__restrict__ out)
{
    __shared__ AtomLds lds[kTile];
    __shared__ double4 ldsd[kTile];  // q, c6, c12 in FP64 for the FP64 pairs
    ff_brick_f32<LJ, NS, Z0, WE>(A, atoms, out, blockIdx.x, lds, ldsd);
}

template <bool LJ, int NS, bool Z0, bool WE, bool ORTHO>
__global__ __launch_bounds__(kThreads) void ff_build_f64(FFArgs A, const AtomRec *__restrict__ atoms,
                                                         double4 *__restrict__ out)
{
    __shared__ double4 lpos[kTile];  // brick-local x, y, z, q
    __shared__ double2 llj[kTile];   // c6, c12
    ff_brick_f64<LJ, NS, Z0, WE, false, ORTHO>(A, atoms, out, blockIdx.x, lpos, llj);
}

template <bool LJ, int NS, bool Z0, bool WE, bool QS>
__global__ __launch_bounds__(kThreads) void ff_build_f32_batch(FFArgs A0, const BatchGridItem *__restrict__ items,
                                                               const long long *__restrict__ prefix, int nitems,
                                                               const AtomRec *__restrict__ atoms, float4 *__restrict__ grid)
{
    __shared__ AtomLds lds[kTile];
    __shared__ double4 ldsd[kTile];
    const long long bidx = blockIdx.x;
    const int m = find_item(prefix, nitems, bidx);
    const BatchGridItem it = items[m];
    const FFArgs A = batch_args(A0, it);
    ff_brick_f32<LJ, NS, Z0, WE, QS>(A, atoms + it.atom_off, grid + it.grid_off, (int)(bidx - prefix[m]), lds, ldsd);
}

template <bool LJ, int NS, bool Z0, bool WE, bool QS, bool ORTHO>
__global__ __launch_bounds__(kThreads) void ff_build_f64_batch(FFArgs A0, const BatchGridItem *__restrict__ items,
                                                               const long long *__restrict__ prefix, int nitems,
                                                               const AtomRec *__restrict__ atoms, double4 *__restrict__ grid)
{
    __shared__ double4 lpos[kTile];
    __shared__ double2 llj[kTile];
    const long long bidx = blockIdx.x;
    const int m = find_item(prefix, nitems, bidx);
    const BatchGridItem it = items[m];
    const FFArgs A = batch_args(A0, it);
    ff_brick_f64<LJ, NS, Z0, WE, QS, ORTHO>(A, atoms + it.atom_off, grid + it.grid_off, (int)(bidx - prefix[m]), lpos, llj);
}



template <bool LJ, int NS, bool Z0, bool WE, bool QS>
void launch_batch_t(const FFArgs &A, const AtomRec *atoms, void *out, bool f64, const BatchLaunch *batch, hipStream_t st)
{
    dim3 block(kThreads), grid((unsigned)batch->nbricks);
    if (f64 && batch->ortho)
        hipLaunchKernelGGL((ff_build_f64_batch<LJ, NS, Z0, WE, QS, true>), grid, block, 0, st, A, batch->items,
                           batch->prefix, batch->nitems, atoms, (double4 *)out);
    else if (f64)
        hipLaunchKernelGGL((ff_build_f64_batch<LJ, NS, Z0, WE, QS, false>), grid, block, 0, st, A, batch->items,
                           batch->prefix, batch->nitems, atoms, (double4 *)out);
    else
        hipLaunchKernelGGL((ff_build_f32_batch<LJ, NS, Z0, WE, QS>), grid, block, 0, st, A, batch->items, batch->prefix,
                           batch->nitems, atoms, (float4 *)out);
}

template <bool LJ, int NS, bool Z0, bool WE>
void launch_t(const FFArgs &A, const AtomRec *atoms, void *out, bool f64, const BatchLaunch *batch, hipStream_t st)
{
    dim3 block(kThreads);
    if (batch) {
        if (LJ && NS > 0 && batch->qskip) launch_batch_t<LJ, NS, Z0, WE, (LJ && (NS > 0))>(A, atoms, out, f64, batch, st);
        else launch_batch_t<LJ, NS, Z0, WE, false>(A, atoms, out, f64, batch, st);
        return;
    }
    dim3 grid((unsigned)A.nbx * A.nby * A.nbz);
    if (f64 && ortho_cell(A.dcell))
        hipLaunchKernelGGL((ff_build_f64<LJ, NS, Z0, WE, true>), grid, block, 0, st, A, atoms, (double4 *)out);
    else if (f64)
        hipLaunchKernelGGL((ff_build_f64<LJ, NS, Z0, WE, false>), grid, block, 0, st, A, atoms, (double4 *)out);
    else
        hipLaunchKernelGGL((ff_build_f32<LJ, NS, Z0, WE>), grid, block, 0, st, A, atoms, (float4 *)out);
}

template <bool LJ, bool WE>
int launch_ns(int ns, bool z0, const FFArgs &A, const AtomRec *atoms, void *out, bool f64, const BatchLaunch *b,
              hipStream_t st)
{
    switch (ns) {
    case 0: launch_t<LJ, 0, false, WE>(A, atoms, out, f64, b, st); break;
    case 1: z0 ? launch_t<LJ, 1, true, WE>(A, atoms, out, f64, b, st) : launch_t<LJ, 1, false, WE>(A, atoms, out, f64, b, st); break;
    case 2: z0 ? launch_t<LJ, 2, true, WE>(A, atoms, out, f64, b, st) : launch_t<LJ, 2, false, WE>(A, atoms, out, f64, b, st); break;
    case 3: z0 ? launch_t<LJ, 3, true, WE>(A, atoms, out, f64, b, st) : launch_t<LJ, 3, false, WE>(A, atoms, out, f64, b, st); break;
    case 4: z0 ? launch_t<LJ, 4, true, WE>(A, atoms, out, f64, b, st) : launch_t<LJ, 4, false, WE>(A, atoms, out, f64, b, st); break;
    default: set_error("n_stencil must be 0..4"); return -2;
    }
    return 0;
}

int first_zero_index(const ppafm_ff_model &m)
{
    for (int s = 0; s < m.n_stencil; s++)
        if (m.stencil_w[s] != 0 && m.stencil_dz[s] == 0) return s;
    return -1;
}

}  // namespace

// Model part of the kernel arguments (stencil compaction, R2SAFE, FP64
// radius); returns whether stencil point 0 sits at zero z offset.
static bool model_args(const ppafm_ff_model &m, FFArgs &A)
{
    // stencil points with zero weight are dropped; the first zero-offset point
    // goes first so that it can share the LJ pair's rsq
    const int zi = first_zero_index(m);
    int ns = 0;
    if (zi >= 0) { A.sw[ns] = (float)m.stencil_w[zi]; A.sdz[ns] = 0.f; ns++; }
    for (int s = 0; s < m.n_stencil; s++)
        if (m.stencil_w[s] != 0 && s != zi) {
            A.sw[ns] = (float)m.stencil_w[s]; A.sdz[ns] = (float)m.stencil_dz[s]; ns++;
        }
    for (int s = ns; s < 4; s++) { A.sw[s] = 0.f; A.sdz[s] = 0.f; }
    A.ns = ns;
    double ws = 0, wdz = 0;
    for (int s = 0; s < ns; s++) { ws += (double)A.sw[s]; wdz += (double)A.sw[s] * (double)A.sdz[s]; }
    A.swsum = (float)ws;
    A.swdz = (float)wdz;
    A.mode = m.mode;
    A.r2safe_d = m.mode == 0 ? (double)1.0e-8f : (double)1.0e-6f;  // Forces.h:12 / FF.cl:1
    A.r2safe = (float)A.r2safe_d;
    A.with_lj = m.with_lj;
    A.probe_q = m.probe_q;
    A.fmax = m.fmax;
    // FP64 radius: at least 0.2 A beyond the farthest stencil offset, so no
    // FP32 pair (LJ or shifted Coulomb point) comes closer than 0.2 A
    double maxdz = 0;
    for (int s = 0; s < ns; s++) maxdz = fmax(maxdz, fabs((double)A.sdz[s]));
    const double rmin = kMinFp64Radius + maxdz;
    const double rn = m.fp64_radius > rmin ? m.fp64_radius : rmin;
    A.near2 = (float)(rn * rn);
    return zi >= 0;
}

static int dispatch(const ppafm_ff_model &m, const FFArgs &A, bool z0, const AtomRec *atoms, void *out, bool f64,
                    const BatchLaunch *b, hipStream_t st)
{
    const bool we = m.with_energy != 0;
    // float4 grids without an FP64 radius: the FP32 kernel (ff_grid_f32.hip)
    if (!f64 && !(m.fp64_radius > 0))
        return launch_ff_f32(A, z0, m.with_lj != 0, we, atoms, out, b, b ? b->ortho : ortho_cell(A.dcell), st);
    if (m.with_lj) {
        if (we) return launch_ns<true, true>(A.ns, z0, A, atoms, out, f64, b, st);
        return launch_ns<true, false>(A.ns, z0, A, atoms, out, f64, b, st);
    }
    if (we) return launch_ns<false, true>(A.ns, z0, A, atoms, out, f64, b, st);
    return launch_ns<false, false>(A.ns, z0, A, atoms, out, f64, b, st);
}

int launch_ff_build(const GridGeom &g, const ppafm_ff_model &m, int natoms, const AtomRec *atoms_dev, void *out,
                    bool f64, hipStream_t st)
{
    PP_REQUIRE(g.n[0] > 0 && g.n[1] > 0 && g.n[2] > 0, "grid dimensions must be positive");
    PP_REQUIRE(m.n_stencil >= 0 && m.n_stencil <= 4, "n_stencil must be 0..4");
    PP_REQUIRE(natoms >= 0, "natoms must be >= 0");
    FFArgs A;
    for (int i = 0; i < 3; i++) A.n[i] = g.n[i];
    A.nbx = (g.n[0] + kColsX - 1) / kColsX;
    A.nby = (g.n[1] + kColsY - 1) / kColsY;
    A.nbz = (g.n[2] + kZBrick - 1) / kZBrick;
    PP_REQUIRE((long)A.nbx * A.nby * A.nbz < (1L << 31), "grid too large for one launch");
    A.natoms = natoms;
    for (int i = 0; i < 9; i++) A.dcell[i] = g.dcell[i];
    const bool z0 = model_args(m, A);
    return dispatch(m, A, z0, atoms_dev, out, f64, nullptr, st);
}

long long ff_bricks(const int n[3])
{
    return (long long)((n[0] + kColsX - 1) / kColsX) * ((n[1] + kColsY - 1) / kColsY) * ((n[2] + kZBrick - 1) / kZBrick);
}

int launch_ff_build_batch(const ppafm_ff_model &m, int nitems, const BatchGridItem *items_dev,
                          const long long *prefix_dev, long long nbricks, const AtomRec *atoms_dev, void *grids,
                          bool f64, bool qskip, bool ortho, hipStream_t st)
{
    PP_REQUIRE(m.n_stencil >= 0 && m.n_stencil <= 4, "n_stencil must be 0..4");
    PP_REQUIRE(nbricks >= 0 && nbricks < (1LL << 31), "batch too large for one launch");
    if (nbricks == 0) return 0;
    FFArgs A = {};
    const bool z0 = model_args(m, A);
    BatchLaunch b{items_dev, prefix_dev, nitems, nbricks, qskip, ortho};
    return dispatch(m, A, z0, atoms_dev, grids, f64, &b, st);
}

}  // namespace pp
