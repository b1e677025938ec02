// FP64 force-field models of the reference C ABI on gfx950.
//
// Pair terms restated from ppafm/cpp (each rounds like the reference: FP64,
// the reference's operation order, no FMA contraction):
//   LJ                 addAtomLJ                 Forces.h:120-132
//   LJ in (R, E)       addAtomLJ_RE              Forces.h:108-117
//   bare r^-6          addAtomVdW_noDamp         Forces.h:135-141
//   damped vdW         addAtomVdW_dampConst      Forces.h:144-148
//   vdW + damping fn   addAtomVdW_addDamp<R2_func, R4_func, invR4_func,
//                      invR8_func>               Forces.h:150-199
//   DFT-D3             addAtomDFTD3              Forces.h:202-222
//   Morse              addAtomMorse              Forces.h:225-232
//   Coulomb s/pz/dz2   addAtomCoulomb + stencils Forces.h:237-244,
//                                                ProbeParticle.cpp:178-209
//   densities          addAtom_Gauss/Slater/splineR4 ProbeParticle.cpp:227-256
//   OpenCL vdW         addvdW without damping and with the invR4 / invR8
//                      damping (FF.cl:548-665: R2SAFE 1e-6f there)
// and the entry points built on them:
//   grids      getLennardJonesFF, getVdWFF, getVdWFF_RE, getDFTD3FF,
//              getMorseFF, getCoulombFF, get{Gauss,Slater}Density,
//              getDensityR4spline (evalCell + iterateGrid3D_omp,
//              ProbeParticle.cpp:262-278, 508-631): one thread per grid
//              point, atoms and their coefficients staged through LDS in
//              tiles, += into the caller's grids;
//   points     getInPoints_LJ (ProbeParticle.cpp:379-395), evalRadialFF
//              (556-575): one thread per point;
//   D3         computeD3Coeffs (397-506): one thread per atom, its
//              coordination number an O(N) loop over the other atoms;
//   utilities  subsample_uniform_spline / subsample_nonuniform_spline
//              (752-784), test_force (786-803), test_eigen3x3 (805-815).

#include <cmath>

#include "pp_common.hpp"
#include "tip_spline.hpp"

#pragma clang fp contract(off)

namespace pp {

namespace {

// Forces.h:12 writes R2SAFE as the float literal 1.0e-8f.
constexpr double kR2Safe = (double)1.0e-8f;
// Forces.h:16 (float literal 400.0f, exact in double)
constexpr double kR2D3Cutoff = 400.0;

struct v3d {
    double x, y, z;
};

__device__ __forceinline__ double norm2(v3d d) { return d.x * d.x + d.y * d.y + d.z * d.z; }
__device__ __forceinline__ void add_mul(v3d &f, v3d d, double s) { f.x += d.x * s; f.y += d.y * s; f.z += d.z * s; }

__device__ __forceinline__ double pair_lj(v3d d, v3d &f, double c6, double c12)
{
    const double ir2 = 1.0 / (norm2(d) + kR2Safe);
    const double ir6 = ir2 * ir2 * ir2;
    const double e6 = c6 * ir6, e12 = c12 * ir6 * ir6;
    add_mul(f, d, (12 * e12 - 6 * e6) * ir2);
    return e12 - e6;
}

__device__ __forceinline__ double pair_lj_re(v3d d, v3d &f, double r0, double e0)
{
    const double ir2 = 1 / (norm2(d) + kR2Safe);
    const double u2 = r0 * r0 * ir2;
    const double u6 = u2 * u2 * u2;
    const double e6 = e0 * u6, e12 = e6 * u6;
    add_mul(f, d, 12 * (e12 - e6) * ir2);
    return e12 - 2 * e6;
}

__device__ __forceinline__ double pair_inv_r6(v3d d, v3d &f, double c6)
{
    const double ir2 = 1 / norm2(d);
    const double ir4 = ir2 * ir2;
    const double e = -c6 * ir4 * ir2;
    add_mul(f, d, e * 6 * ir2);
    return e;
}

__device__ __forceinline__ double pair_vdw_const(v3d d, v3d &f, double c6, double adamp)
{
    double r2 = norm2(d);
    r2 *= r2;
    r2 *= r2;
    add_mul(f, d, -6 * c6 / (r2 + adamp * c6));
    return 0;
}

// Damping functions of u2 = r^2 / R^2 and their derivatives d/du2
// (R2_func, R4_func, invR4_func, invR8_func, Forces.h:150-199).
__device__ __forceinline__ double damp_fn(int kind, double u2, double &dd, double r2safe = kR2Safe)
{
    if (kind == 1) {
        if (u2 > 1) { dd = 0; return 0; }
        dd = -2;
        return 1 - u2;
    }
    if (kind == 2) {
        if (u2 > 1) { dd = 0; return 0; }
        const double e = 1 - u2;
        dd = -4 * e;
        return e * e;
    }
    const double ir2 = 1 / (u2 + r2safe);
    if (kind == 3) {
        const double ir4 = ir2 * ir2;
        dd = -4 * ir4 * ir2;
        return ir4;
    }
    double ir8 = ir2 * ir2;
    ir8 *= ir8;
    dd = -8 * ir8 * ir2;
    return ir8;
}

__device__ __forceinline__ double pair_vdw_damp(int kind, v3d d, v3d &f, double r, double e0, double adamp,
                                               double r2safe = kR2Safe)
{
    const double r2 = norm2(d);
    const double ir2 = 1. / (r * r);
    const double u2 = r2 * ir2;
    const double u4 = u2 * u2;
    double dd;
    const double dfn = damp_fn(kind, u2, dd, r2safe);
    const double e = 1. / (u4 * u2 + dfn * adamp);
    const double en = -2 * e0 * e;
    add_mul(f, d, en * e * (6 * u4 + dd * adamp) * ir2);
    return en;
}

// addvdW of the OpenCL force field (FF.cl:548-665) without damping: the
// R2SAFE there is the float literal 1e-6f.
constexpr double kR2SafeOcl = (double)1.0e-6f;

__device__ __forceinline__ double pair_vdw_nodamp_ocl(v3d d, v3d &f, double c6)
{
    const double ir2 = 1.0 / (norm2(d) + kR2SafeOcl);
    const double ir6 = ir2 * ir2 * ir2;
    const double e = -c6 * ir6;
    add_mul(f, d, 6.0 * e * ir2);
    return e;
}

__device__ __forceinline__ double pair_d3(v3d d, v3d &f, double c6, double c8, double r06, double r08)
{
    const double r2 = norm2(d);
    if (r2 > kR2D3Cutoff) return 0;
    const double r4 = r2 * r2, r6 = r4 * r2, r8 = r6 * r2;
    const double d6 = 1.0 / (r6 + r06), d8 = 1.0 / (r8 + r08);
    const double e6 = -c6 * d6, e8 = -c8 * d8;
    const double f6 = e6 * d6 * 6 * r4, f8 = e8 * d8 * 8 * r6;
    add_mul(f, d, f6 + f8);
    return e6 + e8;
}

__device__ __forceinline__ double pair_morse(v3d d, v3d &f, double r0, double eps, double alpha)
{
    const double r = sqrt(norm2(d) + kR2Safe);
    const double ex = exp(alpha * (r - r0));
    const double e = eps * (ex * ex - 2 * ex);
    const double fr = eps * 2 * alpha * (ex - ex * ex);
    add_mul(f, d, fr / r);
    return e;
}

__device__ __forceinline__ double pair_coul(v3d d, v3d &f, double kqq)
{
    const double ir2 = 1.0 / (norm2(d) + kR2Safe);
    const double ir = sqrt(ir2);
    const double e = ir * kqq;
    add_mul(f, d, e * ir2);
    return e;
}

// The pz / dz2 probes: finite-difference stencils 0.1 A along z.
__device__ __forceinline__ double pair_coul_kind(int kind, v3d d, v3d &fout, double kqq)
{
    if (kind == 0) return pair_coul(d, fout, kqq);
    v3d g{0, 0, 0};
    double e = 0;
    if (kind == 1) {
        d.z -= 0.1;     e += pair_coul(d, g, -kqq);
        d.z += 2 * 0.1; e += pair_coul(d, g, kqq);
        fout.x += g.x * 10.0; fout.y += g.y * 10.0; fout.z += g.z * 10.0;
        return e * 10.0;
    }
    e += pair_coul(d, g, -2 * kqq);
    d.z -= 0.1;     e += pair_coul(d, g, kqq);
    d.z += 2 * 0.1; e += pair_coul(d, g, kqq);
    fout.x += g.x * 100.0; fout.y += g.y * 100.0; fout.z += g.z * 100.0;
    return e * 100.0;
}

__device__ __forceinline__ double dens_gauss(v3d d, double amp, double sigma)
{
    const double is = 1 / sigma;
    const double r2 = norm2(d);
    const double t = r2 * is;
    return amp * exp(-0.5 * (t * t));
}

__device__ __forceinline__ double dens_slater(v3d d, double amp, double lam)
{
    const double beta = 1 / lam;
    const double r = sqrt(norm2(d));
    return amp * exp(-beta * r);
}

__device__ __forceinline__ double dens_r4(v3d d, double amp, double rc)
{
    const double s2 = rc * rc;
    double r2 = norm2(d);
    if (r2 > s2) return 0;
    r2 /= s2;
    const double rf = 1 - r2;
    return amp * rf * rf;
}

// One pair term of model m (FFModelF64) with coefficients c[0..3] and the
// model parameter par; accumulates the force into f, returns the energy.
__device__ __forceinline__ double pair_model(int m, v3d d, v3d &f, const double c[4], double par)
{
    switch (m) {
    case kModelLJ: return pair_lj(d, f, c[0], c[1]);
    case kModelCoulS: return pair_coul_kind(0, d, f, c[0]);
    case kModelCoulPz: return pair_coul_kind(1, d, f, c[0]);
    case kModelCoulDz2: return pair_coul_kind(2, d, f, c[0]);
    case kModelVdW: return pair_vdw_const(d, f, c[0], par);
    case kModelVdWR2: return pair_vdw_damp(1, d, f, c[0], c[1], par);
    case kModelVdWR4: return pair_vdw_damp(2, d, f, c[0], c[1], par);
    case kModelVdWInvR4: return pair_vdw_damp(3, d, f, c[0], c[1], par);
    case kModelVdWInvR8: return pair_vdw_damp(4, d, f, c[0], c[1], par);
    case kModelD3: return pair_d3(d, f, c[0], c[1], c[2], c[3]);
    case kModelMorse: return pair_morse(d, f, c[0], c[1], par);
    case kModelGauss: return dens_gauss(d, c[0], c[1]);
    case kModelSlater: return dens_slater(d, c[0], c[1]);
    case kModelR4: return dens_r4(d, c[0], c[1]);
    case kModelLJRE: return pair_lj_re(d, f, c[0], c[1]);
    case kModelInvR6: return pair_inv_r6(d, f, c[0]);
    case kModelVdWNoDampOcl: return pair_vdw_nodamp_ocl(d, f, c[0]);
    case kModelVdWInvR4Ocl: return pair_vdw_damp(3, d, f, c[0], c[1], par, kR2SafeOcl);
    case kModelVdWInvR8Ocl: return pair_vdw_damp(4, d, f, c[0], c[1], par, kR2SafeOcl);
    default: return 0;
    }
}

constexpr int kTileM = 256;

struct AccumArgs {
    int n[3];
    long npts;
    int natoms, ncoef, model;
    int fstride, estride;  // doubles between consecutive points of F and E
    double par;
    double dcell[9];
};

// evalCell<...> over the grid (ProbeParticle.cpp:262-278) for one model:
// F (npts,3) and E (npts) += where non-NULL.  The model is uniform, so the
// switch in pair_model is a scalar branch.
__global__ __launch_bounds__(256) void ff_accum_models(AccumArgs A, const double *__restrict__ pos,
                                                       const double *__restrict__ coefs, double *__restrict__ F,
                                                       double *__restrict__ E)
{
    __shared__ double4 sx[kTileM];
    __shared__ double4 sc[kTileM];
    const long ig = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = ig < A.npts;
    const long igc = live ? ig : 0;
    const int ic = (int)(igc % A.n[2]);
    const long r = igc / A.n[2];
    const int ib = (int)(r % A.n[1]);
    const int ia = (int)(r / A.n[1]);
    const double *D = A.dcell;
    // iterateGrid3D_omp (Grid.h:142-170): pos0 + a*ia + b*ib + c*ic
    const double px = D[0] * ia + D[3] * ib + D[6] * ic;
    const double py = D[1] * ia + D[4] * ib + D[7] * ic;
    const double pz = D[2] * ia + D[5] * ib + D[8] * ic;
    v3d f{0, 0, 0};
    double e = 0;
    for (int i0 = 0; i0 < A.natoms; i0 += kTileM) {
        const int nt = min(kTileM, A.natoms - i0);
        __syncthreads();
        for (int j = threadIdx.x; j < nt; j += blockDim.x) {
            const long a = i0 + j;
            sx[j] = make_double4(pos[3 * a], pos[3 * a + 1], pos[3 * a + 2], 0.0);
            double c[4] = {0, 0, 0, 0};
            for (int k = 0; k < A.ncoef; k++) c[k] = coefs[a * A.ncoef + k];
            sc[j] = make_double4(c[0], c[1], c[2], c[3]);
        }
        __syncthreads();
        for (int j = 0; j < nt; j++) {
            const double4 a = sx[j], cc = sc[j];
            const double c[4] = {cc.x, cc.y, cc.z, cc.w};
            e += pair_model(A.model, v3d{px - a.x, py - a.y, pz - a.z}, f, c, A.par);
        }
    }
    if (!live) return;
    if (F) {
        double *fp = F + ig * A.fstride;
        fp[0] += f.x; fp[1] += f.y; fp[2] += f.z;
    }
    if (E) E[ig * A.estride] += e;
}

// getInPoints_LJ (ProbeParticle.cpp:379-395): LJ at listed points, with the
// separation taken atom minus point as there; FEs (npoints,4) = (F, E).
__global__ __launch_bounds__(256) void lj_points_f64(long npoints, const double *__restrict__ pts, int natoms,
                                                     const double *__restrict__ pos, const double *__restrict__ cljs,
                                                     double *__restrict__ fes)
{
    __shared__ double4 sx[kTileM];
    __shared__ double2 sc[kTileM];
    const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = ip < npoints;
    const v3d p = live ? v3d{pts[3 * ip], pts[3 * ip + 1], pts[3 * ip + 2]} : v3d{0, 0, 0};
    v3d f{0, 0, 0};
    double e = 0;
    for (int i0 = 0; i0 < natoms; i0 += kTileM) {
        const int nt = min(kTileM, natoms - i0);
        __syncthreads();
        for (int j = threadIdx.x; j < nt; j += blockDim.x) {
            const long a = i0 + j;
            sx[j] = make_double4(pos[3 * a], pos[3 * a + 1], pos[3 * a + 2], 0.0);
            sc[j] = make_double2(cljs[2 * a], cljs[2 * a + 1]);
        }
        __syncthreads();
        for (int j = 0; j < nt; j++) {
            const double4 a = sx[j];
            const double2 c = sc[j];
            e += pair_lj(v3d{a.x - p.x, a.y - p.y, a.z - p.z}, f, c.x, c.y);
        }
    }
    if (!live) return;
    double *o = fes + 4 * ip;
    o[0] = f.x; o[1] = f.y; o[2] = f.z; o[3] = e;
}

// evalRadialFF (ProbeParticle.cpp:556-575): one pair term at separations
// (r, 0, 0); Es = energy, Fs = x force.
__global__ __launch_bounds__(256) void radial_f64(long n, const double *__restrict__ rs, int model, double c0,
                                                  double c1, double par, double *__restrict__ es,
                                                  double *__restrict__ fs)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    v3d f{0, 0, 0};
    const double c[4] = {c0, c1, 0, 0};
    const double e = model < 0 ? 0.0 : pair_model(model, v3d{rs[i], 0, 0}, f, c, par);
    fs[i] = f.x;
    es[i] = e;
}

constexpr int kMaxRefCN = 5;   // ProbeParticle.cpp:28
constexpr int kMaxD3Elem = 94;  // ProbeParticle.cpp:29

struct D3Args {
    int natoms, elem_pp;
    double k1, k2, k3, s6, s8, a1, a2;
};

// computeD3Coeffs (ProbeParticle.cpp:397-506): per atom, the coordination
// number over all other atoms, Gaussian weights of the element's and the
// probe's reference coordination numbers (stored in FP32 as there), C6 as
// their normalised combination, C8 = 3 r4r2_i r4r2_pp C6, and the
// Becke-Johnson damping radii.  Reference coordination numbers < 0 mark
// unused slots; the probe's unused slots get weight 0 (the reference leaves
// them uninitialised).
__global__ __launch_bounds__(256) void d3_coeffs_kernel(D3Args A, const double *__restrict__ pos,
                                                        const int *__restrict__ elems, const double *__restrict__ r_cov,
                                                        const double *__restrict__ ref_cn,
                                                        const double *__restrict__ ref_c6,
                                                        const double *__restrict__ r4r2, double *__restrict__ out)
{
    const int ia = blockIdx.x * blockDim.x + threadIdx.x;
    if (ia >= A.natoms) return;
    const double k12 = -A.k1 * A.k2;
    const int pp = A.elem_pp - 1;
    double lpp[kMaxRefCN];
    for (int i = 0; i < kMaxRefCN; i++) {
        const double cn = ref_cn[pp * kMaxRefCN + i];
        lpp[i] = cn >= 0.0 ? exp(A.k3 * cn * cn) : 0.0;
    }
    const v3d p{pos[3 * ia], pos[3 * ia + 1], pos[3 * ia + 2]};
    const int el = elems[ia] - 1;
    const double rc = r_cov[el];
    double cn = 0;
    for (int j = 0; j < A.natoms; j++) {
        if (j == ia) continue;
        const v3d d{pos[3 * j] - p.x, pos[3 * j + 1] - p.y, pos[3 * j + 2] - p.z};
        const double dist = sqrt(norm2(d));
        const double r = r_cov[elems[j] - 1] + rc;
        cn += 1.0 / (1.0 + exp(k12 * r / dist + A.k1));
    }
    float L[kMaxRefCN * kMaxRefCN];
    double norm = 0;
    int na = 0;
    for (; na < kMaxRefCN; na++) {
        const double ra = ref_cn[el * kMaxRefCN + na];
        if (ra < 0.0) break;
        const double dc = ra - cn;
        const double la = exp(A.k3 * dc * dc);
        for (int b = 0; b < kMaxRefCN; b++) {
            const double lab = la * lpp[b];
            norm += lab;
            L[na * kMaxRefCN + b] = (float)lab;
        }
    }
    if (norm == 0 && na > 0) {
        // coordination beyond every reference: all weight on the highest
        const int a = na - 1;
        for (int b = 0; b < kMaxRefCN; b++) {
            norm += lpp[b];
            L[a * kMaxRefCN + b] = (float)lpp[b];
        }
    }
    const long pair = (long)el * kMaxD3Elem * kMaxRefCN * kMaxRefCN + (long)pp * kMaxRefCN * kMaxRefCN;
    double c6 = 0;
    for (int a = 0; a < na; a++)
        for (int b = 0; b < kMaxRefCN; b++) c6 += L[a * kMaxRefCN + b] * ref_c6[pair + a * kMaxRefCN + b];
    c6 /= norm;
    const double qq = 3 * r4r2[el] * r4r2[pp];
    const double c8 = qq * c6;
    const double r0 = A.a1 * sqrt(qq) + A.a2;
    const double r02 = r0 * r0, r06 = r02 * r02 * r02, r08 = r06 * r02;
    double *o = out + 4 * ia;
    o[0] = c6 * A.s6; o[1] = c8 * A.s8; o[2] = r06; o[3] = r08;
}

__global__ __launch_bounds__(256) void spline_uniform_kernel(double x0, double dx, int n, const double *__restrict__ ydys,
                                                             long m, const double *__restrict__ xs,
                                                             double *__restrict__ ys)
{
    const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const double denom = 1 / dx;
    double u = (xs[j] - x0) * denom;
    int i = (int)u;
    // beyond the table the reference reads past its ends; here the end
    // intervals extrapolate
    i = i < 0 ? 0 : (i > n - 2 ? n - 2 : i);
    u -= i;
    const int k = i << 1;
    ys[j] = hermite(u, ydys[k], ydys[k + 2], ydys[k + 1] * dx, ydys[k + 3] * dx);
}

// subsample_nonuniform_spline: the interval search carries over from one
// sample to the next (a state machine over the samples), so one lane walks
// them in order.
__global__ __launch_bounds__(64) void spline_nonuniform_kernel(int n, const double *__restrict__ kx,
                                                               const double *__restrict__ ydys, long m,
                                                               const double *__restrict__ xs, double *__restrict__ ys)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int i = 0;
    double x0 = 0, x1 = -1e300, dx = 1, denom = 1;
    for (long j = 0; j < m; j++) {
        const double x = xs[j];
        if (x > x1) {
            i = find_index(i, n - i, x, kx);
            if (i > n - 2) i = n - 2;
            x0 = kx[i];
            x1 = kx[i + 1];
            dx = x1 - x0;
            denom = 1 / dx;
        }
        const double u = (x - x0) * denom;
        const int k = i << 1;
        ys[j] = hermite(u, ydys[k], ydys[k + 2], ydys[k + 1] * dx, ydys[k + 3] * dx);
    }
}

struct TestForceArgs {
    int type, n;
    double r0[3], dr[3], R[3];
    double krad, lrad;
    int sn;
    const double *sxs, *sydys;
};

// test_force (ProbeParticle.cpp:786-803): the radial tip force along
// r0 + k dr (r advanced by repeated addition as there), type 1 spline,
// type 2 harmonic; other types give 0 (the reference leaves them unset).
__global__ __launch_bounds__(64) void test_force_kernel(TestForceArgs A, double *__restrict__ fs)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    v3d r{A.r0[0], A.r0[1], A.r0[2]};
    for (int i = 0; i < A.n; i++) {
        const v3d d{r.x - A.R[0], r.y - A.R[1], r.z - A.R[2]};
        v3d f{0, 0, 0};
        if (A.type == 1) {
            const double x = sqrt(norm2(d));
            const double s = tip_spline_force_scale(x, A.sn, A.sxs, A.sydys);
            f = v3d{d.x * s, d.y * s, d.z * s};
        } else if (A.type == 2) {
            const double l = sqrt(norm2(d));
            const double s = A.krad * (l - A.lrad) / l;
            f = v3d{d.x * s, d.y * s, d.z * s};
        }
        fs[3 * i] = f.x; fs[3 * i + 1] = f.y; fs[3 * i + 2] = f.z;
        r.x += A.dr[0]; r.y += A.dr[1]; r.z += A.dr[2];
    }
}

}  // namespace

int launch_ff_models(const GridGeom &g, int model, double par, int natoms, const double *pos, const double *coefs,
                     int ncoef, double *F, double *E, hipStream_t st)
{
    return launch_ff_models_strided(g, model, par, natoms, pos, coefs, ncoef, F, 3, E, 1, st);
}

int launch_ff_models_strided(const GridGeom &g, int model, double par, int natoms, const double *pos,
                             const double *coefs, int ncoef, double *F, int fstride, double *E, int estride,
                             hipStream_t st)
{
    PP_REQUIRE(model >= 0 && model < kModelCount, "unknown force-field model");
    PP_REQUIRE(ncoef >= 0 && ncoef <= 4, "at most 4 coefficients per atom");
    AccumArgs A;
    for (int i = 0; i < 3; i++) A.n[i] = g.n[i];
    A.npts = (long)g.n[0] * g.n[1] * g.n[2];
    A.natoms = natoms;
    A.ncoef = ncoef;
    A.model = model;
    A.fstride = fstride;
    A.estride = estride;
    A.par = par;
    for (int i = 0; i < 9; i++) A.dcell[i] = g.dcell[i];
    if (A.npts == 0 || (!F && !E)) return 0;
    const long nblk = (A.npts + 255) / 256;
    PP_REQUIRE(nblk < (1L << 31), "grid too large for one launch");
    hipLaunchKernelGGL(ff_accum_models, dim3((unsigned)nblk), dim3(256), 0, st, A, pos, coefs, F, E);
    return 0;
}

int launch_lj_points(long npoints, const double *pts, int natoms, const double *pos, const double *cljs, double *fes,
                     hipStream_t st)
{
    if (npoints <= 0) return 0;
    const long nblk = (npoints + 255) / 256;
    PP_REQUIRE(nblk < (1L << 31), "too many points");
    hipLaunchKernelGGL(lj_points_f64, dim3((unsigned)nblk), dim3(256), 0, st, npoints, pts, natoms, pos, cljs, fes);
    return 0;
}

int launch_radial(long n, const double *rs, int model, double c0, double c1, double par, double *es, double *fs,
                  hipStream_t st)
{
    if (n <= 0) return 0;
    const long nblk = (n + 255) / 256;
    PP_REQUIRE(nblk < (1L << 31), "too many points");
    hipLaunchKernelGGL(radial_f64, dim3((unsigned)nblk), dim3(256), 0, st, n, rs, model, c0, c1, par, es, fs);
    return 0;
}

int launch_d3_coeffs(int natoms, const double *pos, const int *elems, const double *r_cov, const double *ref_cn,
                     const double *ref_c6, const double *r4r2, const double k[3], const double params[4], int elem_pp,
                     double *out, hipStream_t st)
{
    if (natoms <= 0) return 0;
    D3Args A;
    A.natoms = natoms;
    A.elem_pp = elem_pp;
    A.k1 = k[0]; A.k2 = k[1]; A.k3 = -k[2];
    A.s6 = params[0]; A.s8 = params[1]; A.a1 = params[2]; A.a2 = params[3];
    hipLaunchKernelGGL(d3_coeffs_kernel, dim3((natoms + 255) / 256), dim3(256), 0, st, A, pos, elems, r_cov, ref_cn,
                       ref_c6, r4r2, out);
    return 0;
}

int launch_spline_uniform(double x0, double dx, int n, const double *ydys, long m, const double *xs, double *ys,
                          hipStream_t st)
{
    PP_REQUIRE(n >= 2, "a spline needs at least 2 knots");
    if (m <= 0) return 0;
    const long nblk = (m + 255) / 256;
    PP_REQUIRE(nblk < (1L << 31), "too many samples");
    hipLaunchKernelGGL(spline_uniform_kernel, dim3((unsigned)nblk), dim3(256), 0, st, x0, dx, n, ydys, m, xs, ys);
    return 0;
}

int launch_spline_nonuniform(int n, const double *kx, const double *ydys, long m, const double *xs, double *ys,
                             hipStream_t st)
{
    PP_REQUIRE(n >= 2, "a spline needs at least 2 knots");
    if (m <= 0) return 0;
    hipLaunchKernelGGL(spline_nonuniform_kernel, dim3(1), dim3(64), 0, st, n, kx, ydys, m, xs, ys);
    return 0;
}

int launch_test_force(int type, int n, const double r0[3], const double dr[3], const double R[3], double krad,
                      double lrad, int sn, const double *sxs, const double *sydys, double *fs, hipStream_t st)
{
    PP_REQUIRE(type != 1 || (sn >= 2 && sxs && sydys), "type 1 needs a tip spline of at least 2 knots");
    if (n <= 0) return 0;
    TestForceArgs A;
    A.type = type; A.n = n;
    for (int k = 0; k < 3; k++) { A.r0[k] = r0[k]; A.dr[k] = dr[k]; A.R[k] = R[k]; }
    A.krad = krad; A.lrad = lrad;
    A.sn = sn; A.sxs = sxs; A.sydys = sydys;
    hipLaunchKernelGGL(test_force_kernel, dim3(1), dim3(64), 0, st, A, fs);
    return 0;
}

}  // namespace pp
