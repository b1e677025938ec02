// FIRE relaxation of the probe particle on gfx950.
//
// Semantics restated from ppafm/cpp/ProbeParticle.cpp (mode 0):
//   getPPforce 282-294, relaxProbe 297-313, FIREstate::move 142-174,
//   RELAX::move 84-88, relaxTipStroke 633-676, relaxTipStrokes_omp 682-701,
//   periodic trilinear interpolation Grid.h:66-111;
// and from ppafm/ocl/cl/relax.cl (mode 1):
//   tipForce 37-41, update_FIRE 105-126, relaxStrokesTilted 391-464, with
//   exact software interpolation in place of the image sampler (interpFE
//   64-73 samples the same corners with hardware-rounded weights).
//
// Execution model: one lane per pixel column, one wave (64 lanes = 8x8
// columns) per block.  Each lane runs its own state machine over
// (z step, FIRE iteration): when its probe converges it writes the step's
// outputs and starts the next step immediately, so a wave runs for
// max over lanes of (sum over z of iterations) instead of the lock-step
// sum over z of the per-step maximum.  Every loop trip is one 8-corner
// gather (the dominant cost); the FIRE state, the tip force and the
// interpolation weights are FP64, the grid values FP32 (device path) or
// FP64 (host ABI path, bit-compatible with the CPU reference).
// Blocks are remapped so that each XCD works on a contiguous band of tiles.

#include <cmath>
#include <cstdlib>

#include "pp_common.hpp"
#include "tip_spline.hpp"

// No FMA contraction in this file: the FP64 path then rounds every operation
// like the CPU reference (x86-64 baseline code has no FMA either).
#pragma clang fp contract(off)

namespace pp {

namespace {

struct RelaxArgs {
    int n[3];
    double dinv[9];
    int mode, alg, max_iters, probe_start;
    double conv_f2, dt, damping, finc, fdec, falpha, dtmin, f2safe;
    double rot[9], d0[3], ks[3], krad, lrad, surf_ff;
    int nx, ny, nz;
    int tiles_x, tiles_y;
    double p0[3], da[3], db[3], dstep[3];
    const double *tips;
    void *pos;
    int pos_f32;
    void *force;
    int *iters;
    unsigned long long *counters;
    // radial tip force from a spline (forceRSpline) instead of the harmonic
    // spring when spline_n > 0 (host ABI only)
    int spline_n;
    const double *spline_xs, *spline_ydys;
};

struct d3 {
    double x, y, z;
};

__device__ __forceinline__ double dot3(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// One axis of SET_GRID_IND (Grid.h:66-69) with the same branches, so the
// FP64 path rounds like the reference.  The modulo runs only off-grid.
__device__ __forceinline__ void wrap_axis(double r, int n, int &lo, int &hi, double &wlo, double &whi)
{
    int i;
    if (r >= 0) {
        i = (int)r;
        whi = r - i;
        wlo = 1 - whi;
        if (i >= n) i %= n;
    } else {
        const int t = (int)r;
        wlo = t - r;
        whi = 1 - wlo;
        i = t % n + n - 1;
    }
    lo = i;
    hi = (i + 1 == n) ? 0 : i + 1;
}

struct Corners {
    long idx[8];
    double w[8];
};

// Corner order and weights of interpolate3DvecWrap (Grid.h:102-107).
__device__ __forceinline__ Corners corners(const int n[3], d3 g)
{
    int ix, jx, iy, jy, iz, jz;
    double mx, tx, my, ty, mz, tz;
    wrap_axis(g.x, n[0], ix, jx, mx, tx);
    wrap_axis(g.y, n[1], iy, jy, my, ty);
    wrap_axis(g.z, n[2], iz, jz, mz, tz);
    const long ny = n[1], nz = n[2];
    const long bx0 = (long)ix * ny, bx1 = (long)jx * ny;
    Corners c;
    c.idx[0] = (bx0 + iy) * nz + iz; c.idx[1] = (bx0 + iy) * nz + jz;
    c.idx[2] = (bx0 + jy) * nz + iz; c.idx[3] = (bx0 + jy) * nz + jz;
    c.idx[4] = (bx1 + jy) * nz + iz; c.idx[5] = (bx1 + jy) * nz + jz;
    c.idx[6] = (bx1 + iy) * nz + iz; c.idx[7] = (bx1 + iy) * nz + jz;
    const double mxmy = mx * my, mxty = mx * ty, txmy = tx * my, txty = tx * ty;
    c.w[0] = mxmy * mz; c.w[1] = mxmy * tz; c.w[2] = mxty * mz; c.w[3] = mxty * tz;
    c.w[4] = txty * mz; c.w[5] = txty * tz; c.w[6] = txmy * mz; c.w[7] = txmy * tz;
    return c;
}

// The corner values of the cell a probe sampled last (key: its first corner
// index), kept in registers across FIRE iterations: a probe that stays in
// that cell samples without memory accesses, with the same values.
template <class V>
struct CornerCache {
    long key = -1;
    V v[8];
};

struct d3v {
    double x, y, z;
};

// FP32 float4 grid (Fx, Fy, Fz, E): all 8 loads issued before any use.
struct GridF32 {
    using Cache = CornerCache<float4>;
    const float4 *__restrict__ g;
    __device__ __forceinline__ void sample(const int n[3], d3 gc, double out[4]) const
    {
        Cache C;
        sample(n, gc, out, C);
    }
    __device__ __forceinline__ void sample(const int n[3], d3 gc, double out[4], Cache &C) const
    {
        const Corners c = corners(n, gc);
        if (c.idx[0] != C.key) {
#pragma unroll
            for (int k = 0; k < 8; k++) C.v[k] = g[c.idx[k]];
            C.key = c.idx[0];
        }
        const float4 *v = C.v;
        out[0] = v[0].x * c.w[0]; out[1] = v[0].y * c.w[0]; out[2] = v[0].z * c.w[0]; out[3] = v[0].w * c.w[0];
#pragma unroll
        for (int k = 1; k < 8; k++) {
            out[0] += v[k].x * c.w[k]; out[1] += v[k].y * c.w[k];
            out[2] += v[k].z * c.w[k]; out[3] += v[k].w * c.w[k];
        }
    }
};

// FP64 double4 grid of the exact device path.
struct GridD4 {
    using Cache = CornerCache<double4>;
    const double4 *__restrict__ g;
    __device__ __forceinline__ void sample(const int n[3], d3 gc, double out[4]) const
    {
        Cache C;
        sample(n, gc, out, C);
    }
    __device__ __forceinline__ void sample(const int n[3], d3 gc, double out[4], Cache &C) const
    {
        const Corners c = corners(n, gc);
        if (c.idx[0] != C.key) {
#pragma unroll
            for (int k = 0; k < 8; k++) C.v[k] = g[c.idx[k]];
            C.key = c.idx[0];
        }
        const double4 *v = C.v;
        out[0] = v[0].x * c.w[0]; out[1] = v[0].y * c.w[0]; out[2] = v[0].z * c.w[0]; out[3] = v[0].w * c.w[0];
#pragma unroll
        for (int k = 1; k < 8; k++) {
            out[0] += v[k].x * c.w[k]; out[1] += v[k].y * c.w[k];
            out[2] += v[k].z * c.w[k]; out[3] += v[k].w * c.w[k];
        }
    }
};

// FP64 (nx,ny,nz,3) grid of the host ABI.
struct GridF64 {
    using Cache = CornerCache<d3v>;
    const double *__restrict__ g;
    __device__ __forceinline__ void sample(const int n[3], d3 gc, double out[4]) const
    {
        Cache C;
        sample(n, gc, out, C);
    }
    __device__ __forceinline__ void sample(const int n[3], d3 gc, double out[4], Cache &C) const
    {
        const Corners c = corners(n, gc);
        if (c.idx[0] != C.key) {
#pragma unroll
            for (int k = 0; k < 8; k++) C.v[k] = d3v{g[3 * c.idx[k]], g[3 * c.idx[k] + 1], g[3 * c.idx[k] + 2]};
            C.key = c.idx[0];
        }
        const d3v *v = C.v;
        out[0] = v[0].x * c.w[0]; out[1] = v[0].y * c.w[0]; out[2] = v[0].z * c.w[0]; out[3] = 0;
#pragma unroll
        for (int k = 1; k < 8; k++) {
            out[0] += v[k].x * c.w[k]; out[1] += v[k].y * c.w[k]; out[2] += v[k].z * c.w[k];
        }
    }
};

__device__ __forceinline__ d3 tip_pos(const RelaxArgs &A, int i, int j, int s)
{
    if (A.tips) {
        const double *p = A.tips + 3 * (((long)i * A.ny + j) * A.nz + s);
        return d3{p[0], p[1], p[2]};
    }
    return d3{A.p0[0] + i * A.da[0] + j * A.db[0] + s * A.dstep[0],
              A.p0[1] + i * A.da[1] + j * A.db[1] + s * A.dstep[1],
              A.p0[2] + i * A.da[2] + j * A.db[2] + s * A.dstep[2]};
}

template <class Grid, bool F64OUT>
__global__ __launch_bounds__(64) void relax_kernel(RelaxArgs A, Grid G)
{
    // XCD-aware tile order: hardware places block b on XCD b % 8; give each
    // XCD a contiguous band of tiles so neighbouring tiles share an L2.
    const int ntiles = A.tiles_x * A.tiles_y;
    const int per_xcd = (ntiles + 7) / 8;
    const int b = blockIdx.x;
    const int tile = (b % 8) * per_xcd + b / 8;
    if (tile >= ntiles) return;
    const int ti = tile / A.tiles_y, tj = tile % A.tiles_y;
    const int lane = threadIdx.x;
    const int i = ti * 8 + lane / 8, j = tj * 8 + lane % 8;
    const bool live = i < A.nx && j < A.ny;

    const int nz = A.nz;
    const bool m0 = A.mode == 0;
    unsigned long long nevals = 0, npos = 0;

    int s = live ? 0 : nz;  // z step
    d3 rtip = live ? tip_pos(A, i, j, 0) : d3{0, 0, 0};
    d3 d0w;  // equilibrium offset in world frame (mode 1: R^T d0)
    if (m0) {
        d0w = d3{A.d0[0], A.d0[1], A.d0[2]};
    } else {
        d0w = d3{A.rot[0] * A.d0[0] + A.rot[3] * A.d0[1] + A.rot[6] * A.d0[2],
                 A.rot[1] * A.d0[0] + A.rot[4] * A.d0[1] + A.rot[7] * A.d0[2],
                 A.rot[2] * A.d0[0] + A.rot[5] * A.d0[1] + A.rot[8] * A.d0[2]};
    }
    d3 r = d3{rtip.x + d0w.x, rtip.y + d0w.y, rtip.z + d0w.z};
    if (m0 && A.probe_start == 1) {
        // relaxTipStroke applies the continuation shift at step 0 too (rTip unchanged)
        const d3 drp{r.x - rtip.x, r.y - rtip.y, r.z - rtip.z};
        r = d3{rtip.x + drp.x, rtip.y + drp.y, rtip.z + drp.z};
    }
    d3 v = d3{0, 0, 0};
    const double dtmax = A.dt, a0 = A.damping, dtmin = A.dtmin;
    double dt = dtmax, ac = a0;
    int it = 0;        // force evaluations in this step
    int phase = (m0 && A.max_iters <= 0) ? 1 : 0;
    int itr = 0;       // iteration index reported for the step
    const d3 dinv0{A.dinv[0], A.dinv[1], A.dinv[2]}, dinv1{A.dinv[3], A.dinv[4], A.dinv[5]},
        dinv2{A.dinv[6], A.dinv[7], A.dinv[8]};
    const d3 RA{A.rot[0], A.rot[1], A.rot[2]}, RB{A.rot[3], A.rot[4], A.rot[5]}, RC{A.rot[6], A.rot[7], A.rot[8]};
    typename Grid::Cache cache;

    while (s < nz) {
        const d3 gc{dot3(r, dinv0), dot3(r, dinv1), dot3(r, dinv2)};
        double fe[4];
        G.sample(A.n, gc, fe, cache);
        nevals++;
        const long o = ((long)i * A.ny + j) * nz + s;
        bool step_done = false;
        if (phase == 1) {
            // mode 0: relaxed position and the grid force sampled there
            if (A.pos) {
                if (A.pos_f32) {
                    float *p = (float *)A.pos + 3 * o;
                    p[0] = (float)r.x; p[1] = (float)r.y; p[2] = (float)r.z;
                } else {
                    double *p = (double *)A.pos + 3 * o;
                    p[0] = r.x; p[1] = r.y; p[2] = r.z;
                }
            }
            if (A.force) {
                if (F64OUT) {
                    double *p = (double *)A.force + 3 * o;
                    p[0] = fe[0]; p[1] = fe[1]; p[2] = fe[2];
                } else {
                    ((float4 *)A.force)[o] = make_float4((float)fe[0], (float)fe[1], (float)fe[2], (float)fe[3]);
                }
            }
            if (A.iters) A.iters[o] = itr;
            step_done = true;
        } else {
            d3 f{fe[0], fe[1], fe[2]};
            const d3 dp{r.x - rtip.x, r.y - rtip.y, r.z - rtip.z};
            if (m0) {
                // getPPforce: radial spring (or spline) then lateral spring
                const double l = sqrt(dot3(dp, dp));
                const double sr = A.spline_n ? tip_spline_force_scale(l, A.spline_n, A.spline_xs, A.spline_ydys)
                                             : A.krad * (l - A.lrad) / l;
                f.x += dp.x * sr; f.y += dp.y * sr; f.z += dp.z * sr;
                const d3 dl{dp.x - A.d0[0], dp.y - A.d0[1], dp.z - A.d0[2]};
                f.x += dl.x * A.ks[0]; f.y += dl.y * A.ks[1]; f.z += dl.z * A.ks[2];
            } else {
                // tipForce in the tip frame, rotated back (relax.cl:430-435)
                const d3 dq{dot3(dp, RA), dot3(dp, RB), dot3(dp, RC)};
                const double rl = sqrt(dot3(dq, dq));
                const double sr = A.krad * (rl - A.lrad) / rl;
                const d3 ft{(dq.x - A.d0[0]) * A.ks[0] + dq.x * sr, (dq.y - A.d0[1]) * A.ks[1] + dq.y * sr,
                            (dq.z - A.d0[2]) * A.ks[2] + dq.z * sr};
                f.x += RA.x * ft.x + RB.x * ft.y + RC.x * ft.z + RC.x * A.surf_ff;
                f.y += RA.y * ft.x + RB.y * ft.y + RC.y * ft.z + RC.y * A.surf_ff;
                f.z += RA.z * ft.x + RB.z * ft.y + RC.z * ft.z + RC.z * A.surf_ff;
            }
            const double ff = dot3(f, f);
            if (A.alg == 1) {
                const double vv = dot3(v, v), vf = dot3(f, v);
                if (vf < 0) {
                    v = d3{0, 0, 0};
                    dt = m0 ? dt * A.fdec : fmax(dtmin, dt * A.fdec);
                    ac = a0;
                } else {
                    const double cf = m0 ? ac * sqrt(vv / ff) : ac * sqrt(vv / (ff + A.f2safe));
                    const double cv = 1 - ac;
                    v.x *= cv; v.y *= cv; v.z *= cv;
                    v.x += f.x * cf; v.y += f.y * cf; v.z += f.z * cf;
                    dt = fmin(dt * A.finc, dtmax);
                    ac = ac * A.falpha;
                }
                v.x += f.x * dt; v.y += f.y * dt; v.z += f.z * dt;
                r.x += v.x * dt; r.y += v.y * dt; r.z += v.z * dt;
            } else {
                const double c = 1 - A.damping;
                v.x *= c; v.y *= c; v.z *= c;
                v.x += f.x * A.dt; v.y += f.y * A.dt; v.z += f.z * A.dt;
                r.x += v.x * A.dt; r.y += v.y * A.dt; r.z += v.z * A.dt;
            }
            it++;
            const bool conv = ff < A.conv_f2;
            if (conv || it >= A.max_iters) {
                if (m0) {
                    itr = conv ? it - 1 : A.max_iters;
                    phase = 1;
                } else {
                    // relaxStrokesTilted output: last evaluated force in the tip frame, path
                    if (A.force) {
                        const d3 fg{fe[0], fe[1], fe[2]};
                        ((float4 *)A.force)[o] = make_float4((float)dot3(fg, RA), (float)dot3(fg, RB),
                                                             (float)dot3(fg, RC), (float)fe[3]);
                    }
                    if (A.pos) {
                        if (A.pos_f32) {
                            float *p = (float *)A.pos + 3 * o;
                            p[0] = (float)r.x; p[1] = (float)r.y; p[2] = (float)r.z;
                        } else {
                            double *p = (double *)A.pos + 3 * o;
                            p[0] = r.x; p[1] = r.y; p[2] = r.z;
                        }
                    }
                    if (A.iters) A.iters[o] = it;
                    step_done = true;
                }
            }
        }
        if (step_done) {
            npos++;
            s++;
            if (s < nz) {
                const d3 nt = tip_pos(A, i, j, s);
                if (m0) {
                    if (A.probe_start == 0) {
                        r = d3{nt.x + A.d0[0], nt.y + A.d0[1], nt.z + A.d0[2]};
                    } else if (A.probe_start == 1) {
                        const d3 drp{r.x - rtip.x, r.y - rtip.y, r.z - rtip.z};
                        r = d3{nt.x + drp.x, nt.y + drp.y, nt.z + drp.z};
                    }
                    dt = dtmax;
                    ac = a0;
                    phase = A.max_iters <= 0 ? 1 : 0;
                } else {
                    r = d3{r.x + (nt.x - rtip.x), r.y + (nt.y - rtip.y), r.z + (nt.z - rtip.z)};
                }
                rtip = nt;
                v = d3{0, 0, 0};
                it = 0;
            }
        }
    }

    if (A.counters) {
        // wave reduction, then one atomic per wave
        for (int off = 32; off > 0; off >>= 1) {
            nevals += __shfl_xor(nevals, off, 64);
            npos += __shfl_xor(npos, off, 64);
        }
        if (lane == 0) {
            atomicAdd(&A.counters[0], nevals);
            atomicAdd(&A.counters[1], npos);
        }
    }
}

// ------------------------------------------------------------- stiffness
//
// stiffnessMatrix (ProbeParticle.cpp:703-750): at each relaxed position r
// the rows of the dynamical matrix are central differences of the total
// probe force (getPPforce, 282-294) under displacements of +-ddisp along x,
// y, z, D_k = -(F(r + h e_k) - F(r - h e_k)) / 2h, made symmetric; its
// eigenvalues are written in ascending order and, on request, the unit
// eigenvectors of the smallest, middle and largest.  The displacement is
// applied and undone on r in place, in the reference's order, so the
// displaced positions round as there.  One lane per position; the six force
// evaluations are six independent 8-corner gathers.

// getPPforce with the harmonic radial spring or the tip spline, in the
// reference's order.
template <class Grid>
__device__ __forceinline__ d3 pp_force_m0(const RelaxArgs &A, const Grid &G, d3 rtip, d3 r)
{
    const d3 gc{dot3(r, d3{A.dinv[0], A.dinv[1], A.dinv[2]}), dot3(r, d3{A.dinv[3], A.dinv[4], A.dinv[5]}),
                dot3(r, d3{A.dinv[6], A.dinv[7], A.dinv[8]})};
    double fe[4];
    G.sample(A.n, gc, fe);
    d3 f{fe[0], fe[1], fe[2]};
    const d3 dp{r.x - rtip.x, r.y - rtip.y, r.z - rtip.z};
    const double l = sqrt(dot3(dp, dp));
    const double sr = A.spline_n ? tip_spline_force_scale(l, A.spline_n, A.spline_xs, A.spline_ydys)
                                 : A.krad * (l - A.lrad) / l;
    f.x += dp.x * sr; f.y += dp.y * sr; f.z += dp.z * sr;
    const d3 dl{dp.x - A.d0[0], dp.y - A.d0[1], dp.z - A.d0[2]};
    f.x += dl.x * A.ks[0]; f.y += dl.y * A.ks[1]; f.z += dl.z * A.ks[2];
    return f;
}

// Eigenvalues of the symmetric matrix [[xx,xy,xz],[xy,yy,yz],[xz,yz,zz]],
// ascending, by the trigonometric solution of the characteristic cubic
// (O. K. Smith, Comm. ACM 4(4):168, 1961): with q = tr/3,
// p = sqrt(|A - qI|_F^2 / 6) and B = (A - qI)/p, det(B)/2 = cos(3 phi) and
// the roots are q + 2p cos(phi + 2k pi/3).  p comes from a sum of squares,
// so it does not cancel when eigenvalues nearly coincide.  NaN inputs give
// NaN eigenvalues.
__device__ __forceinline__ void sym3_eigenvalues(double xx, double yy, double zz, double xy, double xz, double yz,
                                                 double ev[3])
{
    const double q = (xx + yy + zz) / 3.0;
    const double dx = xx - q, dy = yy - q, dz = zz - q;
    const double p2 = dx * dx + dy * dy + dz * dz + 2.0 * (xy * xy + xz * xz + yz * yz);
    if (p2 == 0.0) {
        ev[0] = ev[1] = ev[2] = q;
        return;
    }
    const double p = sqrt(p2 / 6.0), ip = 1.0 / p;
    const double bxx = dx * ip, byy = dy * ip, bzz = dz * ip, bxy = xy * ip, bxz = xz * ip, byz = yz * ip;
    double r = 0.5 * (bxx * (byy * bzz - byz * byz) - bxy * (bxy * bzz - byz * bxz) + bxz * (bxy * byz - byy * bxz));
    r = r > 1.0 ? 1.0 : (r < -1.0 ? -1.0 : r);  // rounding only; NaN passes through
    const double phi = acos(r) / 3.0;
    const double two_pi_3 = 2.0943951023931954923;  // 2 pi / 3
    double hi = q + 2.0 * p * cos(phi);
    double lo = q + 2.0 * p * cos(phi + two_pi_3);
    double mid = 3.0 * q - hi - lo;
    // the construction orders them; the swaps only guard rounding ties
    double t;
    if (lo > mid) { t = lo; lo = mid; mid = t; }
    if (mid > hi) { t = mid; mid = hi; hi = t; }
    if (lo > mid) { t = lo; lo = mid; mid = t; }
    ev[0] = lo; ev[1] = mid; ev[2] = hi;
}

// Unit eigenvector for eigenvalue lam of the matrix with rows a, b, c: the
// longest of the three pairwise cross products of the rows of (A - lam I),
// which are orthogonal to its null space (the selection and sign rule of
// Mat3::eigenvec, Mat3.h:484-500, so the vectors come out with the
// reference's signs).  A fully degenerate matrix (all three products zero)
// gives the unit axis `axis`.
__device__ __forceinline__ d3 eigenvector3(d3 a, d3 b, d3 c, double lam, int axis)
{
    const d3 r0{a.x - lam, a.y, a.z}, r1{b.x, b.y - lam, b.z}, r2{c.x, c.y, c.z - lam};
    const d3 c01{r0.y * r1.z - r0.z * r1.y, r0.z * r1.x - r0.x * r1.z, r0.x * r1.y - r0.y * r1.x};
    const d3 c02{r0.y * r2.z - r0.z * r2.y, r0.z * r2.x - r0.x * r2.z, r0.x * r2.y - r0.y * r2.x};
    const d3 c12{r1.y * r2.z - r1.z * r2.y, r1.z * r2.x - r1.x * r2.z, r1.x * r2.y - r1.y * r2.x};
    const double d0 = dot3(c01, c01), d1 = dot3(c02, c02), d2 = dot3(c12, c12);
    d3 v = c01;
    double dm = d0;
    if (d1 > dm) { dm = d1; v = c02; }
    if (d2 > dm) { dm = d2; v = c12; }
    if (dm == 0.0) return d3{axis == 0 ? 1.0 : 0.0, axis == 1 ? 1.0 : 0.0, axis == 2 ? 1.0 : 0.0};
    const double s = 1.0 / sqrt(dm);
    return d3{v.x * s, v.y * s, v.z * s};
}

struct StiffOut {
    const void *pos;
    int pos_f32;
    double ddisp;
    int which;
    double *evals, *ev1, *ev2, *ev3;
};

template <class Grid>
__global__ __launch_bounds__(256) void stiffness_kernel(RelaxArgs A, Grid G, StiffOut O)
{
    const long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long np = (long)A.nx * A.ny * A.nz;
    if (o >= np) return;
    const int s = (int)(o % A.nz);
    const long col = o / A.nz;
    const int j = (int)(col % A.ny), i = (int)(col / A.ny);
    const d3 rtip = tip_pos(A, i, j, s);
    d3 r;
    if (O.pos_f32) {
        const float *p = (const float *)O.pos + 3 * o;
        r = d3{p[0], p[1], p[2]};
    } else {
        const double *p = (const double *)O.pos + 3 * o;
        r = d3{p[0], p[1], p[2]};
    }
    const double h = O.ddisp, h2 = 2 * O.ddisp, sc = -0.5 / O.ddisp;
    // displaced coordinates as the reference's in-place updates round them:
    // x - h, then + 2h, then - h back (x1), before y is displaced
    const double xm = r.x - h, xp = xm + h2, x1 = xp - h;
    const double ym = r.y - h, yp = ym + h2, y1 = yp - h;
    const double zm = r.z - h, zp = zm + h2;
    d3 a, b, c;
    // one axis per trip: its two gathers in flight together, registers
    // reused across trips (k is uniform, so the branches are scalar)
#pragma unroll 1
    for (int k = 0; k < 3; k++) {
        d3 lo, hi;
        if (k == 0) { lo = d3{xm, r.y, r.z}; hi = d3{xp, r.y, r.z}; }
        else if (k == 1) { lo = d3{x1, ym, r.z}; hi = d3{x1, yp, r.z}; }
        else { lo = d3{x1, y1, zm}; hi = d3{x1, y1, zp}; }
        const d3 f1 = pp_force_m0(A, G, rtip, lo), f2 = pp_force_m0(A, G, rtip, hi);
        const d3 row{(f2.x - f1.x) * sc, (f2.y - f1.y) * sc, (f2.z - f1.z) * sc};
        if (k == 0) a = row;
        else if (k == 1) b = row;
        else c = row;
    }
    // rows a, b, c are the responses to x, y, z displacements; symmetric part
    const double xy = 0.5 * (a.y + b.x), yz = 0.5 * (b.z + c.y), xz = 0.5 * (c.x + a.z);
    double ev[3];
    sym3_eigenvalues(a.x, b.y, c.z, xy, xz, yz, ev);
    double *e = O.evals + 3 * o;
    e[0] = ev[0]; e[1] = ev[1]; e[2] = ev[2];
    double *outs[3] = {O.ev1, O.ev2, O.ev3};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        if (O.which > k) {
            const d3 v = eigenvector3(d3{a.x, xy, xz}, d3{xy, b.y, yz}, d3{xz, yz, c.z}, ev[k], k);
            double *w = outs[k] + 3 * o;
            w[0] = v.x; w[1] = v.y; w[2] = v.z;
        }
    }
}

// test_eigen3x3 (ProbeParticle.cpp:805-815): the eigenvalues of the
// symmetric matrix on and above the diagonal of m (row-major) in the order
// of Mat3::eigenvals (Mat3.h:459-482) -- largest, smallest, middle when the
// largest element of m is positive, the reverse of the first two when it is
// negative, NaN when it is zero -- then the eigenvector of each from the
// full rows of m.  out: 3 eigenvalues, then 3 vectors.
__global__ __launch_bounds__(64) void eigen3x3_kernel(const double *__restrict__ m, double *__restrict__ out)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double amax = m[0];
    for (int k = 1; k < 9; k++) amax = m[k] > amax ? m[k] : amax;
    double ev[3];
    sym3_eigenvalues(m[0], m[4], m[8], m[1], m[2], m[5], ev);
    double o[3];
    if (amax > 0) { o[0] = ev[2]; o[1] = ev[0]; o[2] = ev[1]; }
    else if (amax < 0) { o[0] = ev[0]; o[1] = ev[2]; o[2] = ev[1]; }
    else { o[0] = o[1] = o[2] = __builtin_nan(""); }
    const d3 a{m[0], m[1], m[2]}, b{m[3], m[4], m[5]}, c{m[6], m[7], m[8]};
    for (int k = 0; k < 3; k++) {
        out[k] = o[k];
        const d3 v = eigenvector3(a, b, c, o[k], k);
        out[3 + 3 * k] = v.x; out[4 + 3 * k] = v.y; out[5 + 3 * k] = v.z;
    }
}

RelaxArgs make_args(const RelaxLaunch &L)
{
    RelaxArgs A;
    for (int k = 0; k < 3; k++) A.n[k] = L.g.n[k];
    for (int k = 0; k < 9; k++) A.dinv[k] = L.g.dinv[k];
    A.mode = L.rp.mode; A.alg = L.rp.alg; A.max_iters = L.rp.max_iters; A.probe_start = L.rp.probe_start;
    A.conv_f2 = L.rp.conv_f2; A.dt = L.rp.dt; A.damping = L.rp.damping; A.finc = L.rp.finc; A.fdec = L.rp.fdec;
    A.falpha = L.rp.falpha; A.dtmin = L.rp.dtmin_frac * L.rp.dt; A.f2safe = L.rp.f2safe;
    for (int k = 0; k < 9; k++) A.rot[k] = L.tip.rot[k];
    for (int k = 0; k < 3; k++) { A.d0[k] = L.tip.d0[k]; A.ks[k] = L.tip.kspring[k]; }
    A.krad = L.tip.krad; A.lrad = L.tip.lrad; A.surf_ff = L.tip.surf_ff;
    A.nx = L.scan.nx; A.ny = L.scan.ny; A.nz = L.scan.nz;
    A.tiles_x = (A.nx + 7) / 8; A.tiles_y = (A.ny + 7) / 8;
    for (int k = 0; k < 3; k++) {
        A.p0[k] = L.scan.p0[k]; A.da[k] = L.scan.da[k]; A.db[k] = L.scan.db[k]; A.dstep[k] = L.scan.dstep[k];
    }
    A.tips = L.tips;
    A.pos = L.out.pos; A.pos_f32 = L.out.pos_f32; A.force = L.out.force; A.iters = L.out.iters;
    A.counters = L.out.counters;
    A.spline_n = L.spline_n; A.spline_xs = L.spline_xs; A.spline_ydys = L.spline_ydys;
    return A;
}

int check(const RelaxLaunch &L)
{
    PP_REQUIRE(L.scan.nx >= 0 && L.scan.ny >= 0 && L.scan.nz >= 0, "scan sizes must be >= 0");
    PP_REQUIRE(L.g.n[0] > 0 && L.g.n[1] > 0 && L.g.n[2] > 0, "grid dimensions must be positive");
    PP_REQUIRE(L.rp.mode == 0 || L.rp.mode == 1, "relax mode must be 0 or 1");
    PP_REQUIRE(L.rp.alg == 0 || L.rp.alg == 1, "relaxAlg must be 0 (leap-frog) or 1 (FIRE)");
    PP_REQUIRE(L.rp.mode == 0 || L.rp.alg == 1, "mode 1 supports FIRE only");
    PP_REQUIRE(L.rp.max_iters >= 0 && L.rp.max_iters <= 1000000, "max_iters out of range");
    PP_REQUIRE(L.rp.mode == 0 || L.rp.max_iters >= 1, "mode 1 needs max_iters >= 1");
    PP_REQUIRE(L.rp.probe_start >= -1 && L.rp.probe_start <= 1, "probeStart must be -1, 0 or 1");
    PP_REQUIRE(L.spline_n == 0 || (L.spline_n >= 2 && L.spline_xs && L.spline_ydys),
               "a tip spline needs at least 2 knots");
    PP_REQUIRE(L.spline_n == 0 || L.rp.mode == 0, "tip splines apply to mode 0");
    return 0;
}

}  // namespace

int launch_relax_dev(const RelaxLaunch &L, const void *grid, bool f64, hipStream_t st)
{
    int e = check(L);
    if (e) return e;
    PP_REQUIRE(L.spline_n == 0, "tip splines are supported on the host ABI's FP64 grid only");
    static const bool generic = getenv("PPAFM_RELAX_GENERIC") != nullptr;
    if (!generic) {
        e = launch_relax_fast(L, grid, f64, st);
        if (e <= 0) return e;
    }
    RelaxArgs A = make_args(L);
    const long nt = (long)A.tiles_x * A.tiles_y;
    if (nt == 0 || A.nz == 0) return 0;
    const long nblk = ((nt + 7) / 8) * 8;
    PP_REQUIRE(nblk < (1L << 31), "scan too large");
    if (f64)
        hipLaunchKernelGGL((relax_kernel<GridD4, false>), dim3((unsigned)nblk), dim3(64), 0, st, A,
                           GridD4{(const double4 *)grid});
    else
        hipLaunchKernelGGL((relax_kernel<GridF32, false>), dim3((unsigned)nblk), dim3(64), 0, st, A,
                           GridF32{(const float4 *)grid});
    return 0;
}

int launch_relax_f64grid(const RelaxLaunch &L, const double *grid, double *pos, double *force, int *iters,
                         hipStream_t st)
{
    int e = check(L);
    if (e) return e;
    PP_REQUIRE(L.rp.mode == 0, "the FP64 grid path implements mode 0");
    RelaxArgs A = make_args(L);
    A.pos = pos; A.pos_f32 = 0; A.force = force; A.iters = iters;
    const long nt = (long)A.tiles_x * A.tiles_y;
    if (nt == 0 || A.nz == 0) return 0;
    const long nblk = ((nt + 7) / 8) * 8;
    PP_REQUIRE(nblk < (1L << 31), "scan too large");
    hipLaunchKernelGGL((relax_kernel<GridF64, true>), dim3((unsigned)nblk), dim3(64), 0, st, A, GridF64{grid});
    return 0;
}

int launch_stiffness(const RelaxLaunch &L, const void *grid, int grid_kind, const void *pos, int pos_f32,
                     double ddisp, int which, double *evals, double *ev1, double *ev2, double *ev3, hipStream_t st)
{
    PP_REQUIRE(L.scan.nx >= 0 && L.scan.ny >= 0 && L.scan.nz >= 0, "scan sizes must be >= 0");
    PP_REQUIRE(L.g.n[0] > 0 && L.g.n[1] > 0 && L.g.n[2] > 0, "grid dimensions must be positive");
    PP_REQUIRE(grid_kind >= 0 && grid_kind <= 2, "grid_kind must be 0 (float4), 1 (double4) or 2 (double3)");
    PP_REQUIRE(std::isfinite(ddisp) && ddisp != 0.0, "ddisp must be finite and non-zero");
    PP_REQUIRE(L.spline_n == 0 || (L.spline_n >= 2 && L.spline_xs && L.spline_ydys),
               "a tip spline needs at least 2 knots");
    const long np = (long)L.scan.nx * L.scan.ny * L.scan.nz;
    if (np == 0) return 0;
    PP_REQUIRE(grid && pos && evals, "null grid, position or eigenvalue pointer");
    PP_REQUIRE(which < 1 || ev1, "which >= 1 needs the first eigenvector output");
    PP_REQUIRE(which < 2 || ev2, "which >= 2 needs the second eigenvector output");
    PP_REQUIRE(which < 3 || ev3, "which >= 3 needs the third eigenvector output");
    RelaxArgs A = make_args(L);
    const StiffOut O{pos, pos_f32, ddisp, which, evals, ev1, ev2, ev3};
    const long nblk = (np + 255) / 256;
    PP_REQUIRE(nblk < (1L << 31), "too many positions");
    if (grid_kind == 0)
        hipLaunchKernelGGL((stiffness_kernel<GridF32>), dim3((unsigned)nblk), dim3(256), 0, st, A,
                           GridF32{(const float4 *)grid}, O);
    else if (grid_kind == 1)
        hipLaunchKernelGGL((stiffness_kernel<GridD4>), dim3((unsigned)nblk), dim3(256), 0, st, A,
                           GridD4{(const double4 *)grid}, O);
    else
        hipLaunchKernelGGL((stiffness_kernel<GridF64>), dim3((unsigned)nblk), dim3(256), 0, st, A,
                           GridF64{(const double *)grid}, O);
    return 0;
}

}  // namespace pp

namespace pp {

int launch_eigen3x3(const double *m, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(eigen3x3_kernel, dim3(1), dim3(64), 0, st, m, out);
    return 0;
}

}  // namespace pp
