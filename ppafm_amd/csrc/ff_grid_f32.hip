// FP32 force-field grid kernel on gfx950 (the float4 grid of the OpenCL
// path, FF.cl:366-423 evalLJC_QZs_noPos, and an FP32 variant of the
// ProbeParticle grid, Forces.h:120-132 / 237-244): VALU bound.
//
//   * Brick mapping as the FP64 kernel (4x4 grid columns x 64 z points per
//     256-thread block, 4 z points per lane, z fastest stores).
//   * Atoms are staged through LDS, brick-local, as FP32 hi + lo parts
//     (x = hi + lo within FP32 rounding of the FP64 offset), so a pair
//     distance keeps FP32 relative accuracy 40 A from the grid origin.
//   * One v_rsq_f32 per pair serves the LJ term and the Coulomb point at
//     zero z offset; orthogonal cells share dx, dy and dx^2 + dy^2 + R2SAFE
//     across a lane's 4 z points.
//   * Coulomb stencils with z offsets (tip charges Qs at QZs, or the pz /
//     dz2 probe kinds) are finite differences whose terms cancel by orders
//     of magnitude (the dz2-like -10/+20/-10 tip: ~10^3).  They are summed
//     relative to the zero-offset distance r0 without forming the
//     differences in floating point: with W = sum w_s,
//       sum w_s / r_s   = W / r0   + sum w_s (1/r_s - 1/r0),
//       1/r_s - 1/r0    = (r0^2 - r_s^2) / (r0 r_s (r0 + r_s)),
//       r0^2 - r_s^2    = -dz_s (2 dz + dz_s)            (exact offsets),
//       1/r_s^3 - 1/r0^3 = (1/r_s - 1/r0)(1/r_s^2 + 1/(r_s r0) + 1/r0^2),
//     and the z offsets' own term sum w_s dz_s / r_s^3 the same way, so the
//     FP32 rounding is relative to the stencil's sum, not to its terms.
//   * The OpenCL model (mode 1) sums LJ and Coulomb into one force scale per
//     pair (one set of 3 FMAs); the ProbeParticle model (mode 0) keeps the
//     two fields apart for their separate clamps (io.py:766-778).
//   * FP32 partial sums over 32 atoms are folded into FP64 accumulators.
//   * Overflow guard: where a pair comes so close that an FP32 intermediate
//     overflows (r^-14 of the repulsion, below ~0.005 A), the point's sums
//     are no longer finite; such points (rare: an atom on a grid node) are
//     summed again over all atoms in FP64 after the loop.  Every other pair
//     is FP32.
// The epilogue clamps and combines as the FP64 kernel does.

#include <type_traits>

#include "ff_grid.hpp"

#pragma clang fp contract(off)

namespace pp {

namespace {

// SEP: separate LJ and Coulomb accumulators (mode 0), else one (mode 1).
template <bool LJ, int NS, bool Z0, bool WE, bool QS, bool ORTHO, bool SEP>
__device__ __forceinline__ void brick_f32(const FFArgs &A, const AtomRec *__restrict__ atoms, float4 *__restrict__ out,
                                          int b, AtomLds *lds)
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
    float px[kK], py[kK], pz[kK];
#pragma unroll
    for (int k = 0; k < kK; k++) {
        const double fz = (z0 + zl + kZLanes * k) - cz;
        px[k] = (float)(D[0] * fx + D[3] * fy + D[6] * fz);
        py[k] = (float)(D[1] * fx + D[4] * fy + D[7] * fz);
        pz[k] = (float)(D[2] * fx + D[5] * fy + D[8] * fz);
    }
    const int nk = min(kK, (A.n[2] - z0 + kZLanes - 1) / kZLanes);

    double aL[kK][4], aE[kK][4];  // aE: mode 0 only
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
        }
        __syncthreads();

        for (int j0 = 0; j0 < nt; j0 += kFold) {
            const int j1 = min(nt, j0 + kFold);
            float sL[kK][4], sE[kK][4];
#pragma unroll
            for (int k = 0; k < kK; k++)
#pragma unroll
                for (int c = 0; c < 4; c++) { sL[k][c] = 0.f; sE[k][c] = 0.f; }
            // one atom against the lane's points; all: every point of the
            // lane is inside the grid (all bricks but the top ones), so
            // no per-point test
            auto atom = [&](int j, auto all) {
                const float4 hi = lds[j].hi;
                const float4 lo = lds[j].lo;
                const float4 lj = lds[j].lj;
                // uniform over the block (every lane reads the same atom)
                const bool charged = NS > 0 && (!QS || hi.w != 0.f);
                float dx0 = 0.f, dy0 = 0.f, dxy2 = 0.f;
                if (ORTHO) {
                    dx0 = (px[0] - hi.x) - lo.x;
                    dy0 = (py[0] - hi.y) - lo.y;
                    dxy2 = __fmaf_rn(dx0, dx0, __fmaf_rn(dy0, dy0, A.r2safe));
                }
#pragma unroll
                for (int k = 0; k < kK; k++) {
                    if (!decltype(all)::value && k >= nk) break;
                    const float dx = ORTHO ? dx0 : (px[k] - hi.x) - lo.x;
                    const float dy = ORTHO ? dy0 : (py[k] - hi.y) - lo.y;
                    // (pz - hi) is exact for a close atom (the operands are
                    // within a factor of 2), which the near pairs' r^-13 needs
                    const float dz = (pz[k] - hi.z) - lo.z;
                    const float r2 = ORTHO ? __fmaf_rn(dz, dz, dxy2) : r2_f32(dx, dy, dz, A.r2safe);
                    float ir0 = 0.f, ir20 = 0.f;
                    if (LJ || NS > 0) {
                        ir0 = __builtin_amdgcn_rsqf(r2);
                        ir20 = ir0 * ir0;
                    }
                    // s: force over distance at zero z offset; e: energy;
                    // sc, ec, fzc: the Coulomb stencil's force scale, energy
                    // and extra z force
                    // s = s0 * ir8 and e = e0 * ir6; the combined sum folds
                    // these products into the additions of the Coulomb terms
                    float s0 = 0.f, e0 = 0.f, ir6 = 0.f, ir8 = 0.f, sc = 0.f, ec = 0.f, fzc = 0.f;
                    if (LJ) {
                        ir6 = ir20 * ir20 * ir20;
                        ir8 = ir6 * ir20;
                        s0 = __fmaf_rn(lj.y, ir6, -lj.x);
                        if (WE) e0 = __fmaf_rn(lj.w, ir6, -lj.z);
                    }
                    if (charged && NS == 1 && Z0) {  // one point at zero offset
                        ec = (hi.w * A.sw[0]) * ir0;
                        sc = ec * ir20;
                    } else if (charged) {  // stencil with offsets, cancellation free
                        const float r0 = r2 * ir0;
                        float s1 = 0.f, s3 = 0.f, s3z = 0.f;
#pragma unroll
                        for (int s_ = Z0 ? 1 : 0; s_ < NS; s_++) {
                            const float dzs = dz + A.sdz[s_];
                            const float r2s = ORTHO ? __fmaf_rn(dzs, dzs, dxy2) : r2_f32(dx, dy, dzs, A.r2safe);
                            const float irs = __builtin_amdgcn_rsqf(r2s);
                            const float rs = r2s * irs;
                            const float dr2 = -A.sdz[s_] * (dz + dz + A.sdz[s_]);  // r0^2 - rs^2
                            const float d1 = (dr2 * ir0) * (irs * __builtin_amdgcn_rcpf(r0 + rs));
                            const float d3 = d1 * __fmaf_rn(irs, irs, __fmaf_rn(irs, ir0, ir20));
                            s1 = __fmaf_rn(A.sw[s_], d1, s1);
                            s3 = __fmaf_rn(A.sw[s_], d3, s3);
                            s3z = __fmaf_rn(A.sw[s_] * A.sdz[s_], d3, s3z);
                        }
                        const float ir30 = ir0 * ir20;
                        ec = hi.w * __fmaf_rn(A.swsum, ir0, s1);
                        sc = hi.w * __fmaf_rn(A.swsum, ir30, s3);
                        fzc = hi.w * __fmaf_rn(A.swdz, ir30, s3z);
                    }
                    if (SEP) {
                        if (LJ) {
                            const float s = s0 * ir8;
                            sL[k][0] = __fmaf_rn(dx, s, sL[k][0]);
                            sL[k][1] = __fmaf_rn(dy, s, sL[k][1]);
                            sL[k][2] = __fmaf_rn(dz, s, sL[k][2]);
                            if (WE) sL[k][3] = __fmaf_rn(e0, ir6, sL[k][3]);
                        }
                        if (charged) {
                            sE[k][0] = __fmaf_rn(dx, sc, sE[k][0]);
                            sE[k][1] = __fmaf_rn(dy, sc, sE[k][1]);
                            sE[k][2] = __fmaf_rn(dz, sc, NS == 1 && Z0 ? sE[k][2] : sE[k][2] + fzc);
                            if (WE) sE[k][3] += ec;
                        }
                    } else if (LJ || charged) {
                        // s0 ir8 + ec ir20 = (s0 ir6 + ec) ir20 (one Coulomb point)
                        const float st = LJ ? (NS == 1 && Z0 ? __fmaf_rn(s0, ir6, ec) * ir20 : __fmaf_rn(s0, ir8, sc)) : sc;
                        sL[k][0] = __fmaf_rn(dx, st, sL[k][0]);
                        sL[k][1] = __fmaf_rn(dy, st, sL[k][1]);
                        sL[k][2] = __fmaf_rn(dz, st, NS == 1 && Z0 ? sL[k][2] : sL[k][2] + fzc);
                        if (WE) sL[k][3] += LJ ? __fmaf_rn(e0, ir6, ec) : ec;
                    }
                }
                        };
            if (nk == kK) {
                for (int j = j0; j < j1; j++) atom(j, std::true_type{});
            } else {
                for (int j = j0; j < j1; j++) atom(j, std::false_type{});
            }
#pragma unroll
            for (int k = 0; k < kK; k++)
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    aL[k][c] += (double)sL[k][c];
                    if (SEP) aE[k][c] += (double)sE[k][c];
                }
        }
    }

    if (ix >= A.n[0] || iy >= A.n[1]) return;
#pragma unroll
    for (int k = 0; k < kK; k++) {
        const int iz = z0 + zl + kZLanes * k;
        if (iz >= A.n[2]) break;
        bool finite = true;
#pragma unroll
        for (int c = 0; c < 4; c++) finite = finite && isfinite(aL[k][c]) && (!SEP || isfinite(aE[k][c]));
        if (!finite) {
            // overflow guard: this point again, every pair in FP64
            const double fz = (z0 + zl + kZLanes * k) - cz;
            const double qx = ccx + (D[0] * fx + D[3] * fy + D[6] * fz);
            const double qy = ccy + (D[1] * fx + D[4] * fy + D[7] * fz);
            const double qz = ccz + (D[2] * fx + D[5] * fy + D[8] * fz);
#pragma unroll
            for (int c = 0; c < 4; c++) { aL[k][c] = 0; aE[k][c] = 0; }
            for (int j = 0; j < A.natoms; j++) {
                const AtomRec a = atoms[j];
                pair_fp64<LJ, NS, WE>(A, qx - a.x, qy - a.y, qz - a.z, a.q, a.c6, a.c12, aL[k], SEP ? aE[k] : aL[k]);
            }
        }
        double gx, gy, gz, en;
        if (SEP) {
            clamp_vec(aL[k][0], aL[k][1], aL[k][2], A.fmax);
            clamp_vec(aE[k][0], aE[k][1], aE[k][2], A.fmax);
            gx = aL[k][0] + A.probe_q * aE[k][0];
            gy = aL[k][1] + A.probe_q * aE[k][1];
            gz = aL[k][2] + A.probe_q * aE[k][2];
            en = aL[k][3] + A.probe_q * aE[k][3];
        } else {
            gx = aL[k][0]; gy = aL[k][1]; gz = aL[k][2]; en = aL[k][3];
        }
        out[((long)ix * A.n[1] + iy) * A.n[2] + iz] = make_float4((float)gx, (float)gy, (float)gz, WE ? (float)en : 0.f);
    }
}

template <bool LJ, int NS, bool Z0, bool WE, bool ORTHO, bool SEP>
__global__ __launch_bounds__(kThreads) void ff_f32_kernel(FFArgs A, const AtomRec *__restrict__ atoms,
                                                          float4 *__restrict__ out)
{
    __shared__ AtomLds lds[kTile];
    brick_f32<LJ, NS, Z0, WE, false, ORTHO, SEP>(A, atoms, out, blockIdx.x, lds);
}

template <bool LJ, int NS, bool Z0, bool WE, bool QS, bool ORTHO, bool SEP>
__global__ __launch_bounds__(kThreads) void ff_f32_batch_kernel(FFArgs A0, const BatchGridItem *__restrict__ items,
                                                                const long long *__restrict__ prefix, int nitems,
                                                                const AtomRec *__restrict__ atoms, float4 *__restrict__ grid)
{
    __shared__ AtomLds lds[kTile];
    const long long bidx = blockIdx.x;
    const int m = find_item(prefix, nitems, bidx);
    const BatchGridItem it = items[m];
    const FFArgs A = batch_args(A0, it);
    brick_f32<LJ, NS, Z0, WE, QS, ORTHO, SEP>(A, atoms + it.atom_off, grid + it.grid_off, (int)(bidx - prefix[m]), lds);
}

template <bool LJ, int NS, bool Z0, bool WE, bool ORTHO, bool SEP>
void launch_o(const FFArgs &A, const AtomRec *atoms, void *out, const BatchLaunch *b, hipStream_t st)
{
    dim3 block(kThreads);
    if (b) {
        dim3 grid((unsigned)b->nbricks);
        if (LJ && NS > 0 && b->qskip)
            hipLaunchKernelGGL((ff_f32_batch_kernel<LJ, NS, Z0, WE, (LJ && NS > 0), ORTHO, SEP>), grid, block, 0, st, A,
                               b->items, b->prefix, b->nitems, atoms, (float4 *)out);
        else
            hipLaunchKernelGGL((ff_f32_batch_kernel<LJ, NS, Z0, WE, false, ORTHO, SEP>), grid, block, 0, st, A, b->items,
                               b->prefix, b->nitems, atoms, (float4 *)out);
        return;
    }
    dim3 grid((unsigned)A.nbx * A.nby * A.nbz);
    hipLaunchKernelGGL((ff_f32_kernel<LJ, NS, Z0, WE, ORTHO, SEP>), grid, block, 0, st, A, atoms, (float4 *)out);
}

template <bool LJ, int NS, bool Z0, bool WE>
void launch_s(const FFArgs &A, bool ortho, const AtomRec *atoms, void *out, const BatchLaunch *b, hipStream_t st)
{
    const bool sep = A.mode == 0;
    if (ortho) {
        if (sep) launch_o<LJ, NS, Z0, WE, true, true>(A, atoms, out, b, st);
        else launch_o<LJ, NS, Z0, WE, true, false>(A, atoms, out, b, st);
    } else {
        if (sep) launch_o<LJ, NS, Z0, WE, false, true>(A, atoms, out, b, st);
        else launch_o<LJ, NS, Z0, WE, false, false>(A, atoms, out, b, st);
    }
}

template <bool LJ, bool WE>
int launch_ns(int ns, bool z0, const FFArgs &A, bool ortho, const AtomRec *atoms, void *out, const BatchLaunch *b,
              hipStream_t st)
{
    switch (ns) {
    case 0: launch_s<LJ, 0, false, WE>(A, ortho, atoms, out, b, st); break;
    case 1: z0 ? launch_s<LJ, 1, true, WE>(A, ortho, atoms, out, b, st) : launch_s<LJ, 1, false, WE>(A, ortho, atoms, out, b, st); break;
    case 2: z0 ? launch_s<LJ, 2, true, WE>(A, ortho, atoms, out, b, st) : launch_s<LJ, 2, false, WE>(A, ortho, atoms, out, b, st); break;
    case 3: z0 ? launch_s<LJ, 3, true, WE>(A, ortho, atoms, out, b, st) : launch_s<LJ, 3, false, WE>(A, ortho, atoms, out, b, st); break;
    case 4: z0 ? launch_s<LJ, 4, true, WE>(A, ortho, atoms, out, b, st) : launch_s<LJ, 4, false, WE>(A, ortho, atoms, out, b, st); break;
    default: set_error("n_stencil must be 0..4"); return -2;
    }
    return 0;
}

}  // namespace

int launch_ff_f32(const FFArgs &A, bool z0, bool with_lj, bool with_energy, const AtomRec *atoms, void *out,
                  const BatchLaunch *batch, bool ortho, hipStream_t st)
{
    if (with_lj) {
        if (with_energy) return launch_ns<true, true>(A.ns, z0, A, ortho, atoms, out, batch, st);
        return launch_ns<true, false>(A.ns, z0, A, ortho, atoms, out, batch, st);
    }
    if (with_energy) return launch_ns<false, true>(A.ns, z0, A, ortho, atoms, out, batch, st);
    return launch_ns<false, false>(A.ns, z0, A, ortho, atoms, out, batch, st);
}

}  // namespace pp
