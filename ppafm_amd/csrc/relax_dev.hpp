// Device helpers of the FIRE relaxation shared by relax_dev.hip and
// ws_batch.hip: kernel arguments, periodic trilinear sampling (Grid.h:66-69
// semantics), the tip springs and one FIRE move.  Contraction is off in the
// files that include this, so every kernel rounds identically.
#pragma once

#include "pp_common.hpp"

#pragma clang fp contract(off)

namespace pp {

namespace {

struct DevArgs {
    int n0, n1, n2;
    int nx, ny, nz;
    int tiles_y, ntiles;
    int max_iters;
    int probe_start;
    double dinv[9];
    double conv_f2, dtmax, a0, finc, fdec, falpha, dtmin, f2safe;
    double ks[3], d0[3], krad, lrad, surf_ff;
    double rot[9];
    double p0[3], da[3], db[3], dstep[3];
    const double *tips;
    void *pos;
    int pos_f32;
    float4 *force;
    int *iters;
    unsigned char *status = nullptr;  // per-position status (store_status), optional
    unsigned long long *counters;
    // warm starts: step-0 estimates per pixel column (3 floats, nullptr: the
    // plain start), Newton iterations allowed from the continuation (steps
    // > 0) and at step 0, FIRE iterations between Newton tries (0: none),
    // squared step length that counts as converged, cap of one Newton step
    const float *pred;
    int newton_max, newton_max0, newton_retry;
    double newton_tol2, newton_cap;
    double newton_extrap2;  // > 0: steps >= 2 start Newton at the second-order prediction when its
                            // shift is at most sqrt(newton_extrap2)
    double ws_res2;         // > 0: step 0 falls back to plain FIRE where |F|^2 at the estimate exceeds it
};

// 1/sqrt(x): v_rsq_f64 (5.2e-8 relative on gfx950) and one Newton step
// (4.3e-15 relative, measured with a probe kernel)
__device__ __forceinline__ double rsq_nr(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x * y;
    return fma(y, fma(-h, y, 0.5), y);
}

// grid index coordinates of a position (rows of dinv).  ORTHO: dinv is
// diagonal, and the diagonal products alone are the same numbers for finite
// positions (the zero terms add a signed zero, which no later operation
// tells apart)
template <bool ORTHO = false>
__device__ __forceinline__ void to_grid(const DevArgs &A, double rx, double ry, double rz, double &gx, double &gy,
                                        double &gz)
{
    if (ORTHO) {
        gx = rx * A.dinv[0];
        gy = ry * A.dinv[4];
        gz = rz * A.dinv[8];
        return;
    }
    gx = fma(rz, A.dinv[2], fma(ry, A.dinv[1], rx * A.dinv[0]));
    gy = fma(rz, A.dinv[5], fma(ry, A.dinv[4], rx * A.dinv[3]));
    gz = fma(rz, A.dinv[8], fma(ry, A.dinv[7], rx * A.dinv[6]));
}

// periodic cell index and weights along one axis (Grid.h:66-69 semantics)
__device__ __forceinline__ void axis(double r, int n, int &lo, int &hi, double &w)
{
    double fl = floor(r);
    w = r - fl;
    int i = (int)fl;
    if ((unsigned)i >= (unsigned)n) {
        i %= n;
        if (i < 0) i += n;
    }
    lo = i;
    hi = (i + 1 == n) ? 0 : i + 1;
}

// trilinear sum of the 8 corners (order: x, y, z bits 2, 1, 0)
template <typename T4>
__device__ __forceinline__ void interp8(const T4 v[8], double tx, double ty, double tz, double out[4], bool want_e)
{
    const double mx = 1 - tx, my = 1 - ty, mz = 1 - tz;
    const double w00 = mx * my, w01 = mx * ty, w10 = tx * my, w11 = tx * ty;
    const double w[8] = {w00 * mz, w00 * tz, w01 * mz, w01 * tz, w10 * mz, w10 * tz, w11 * mz, w11 * tz};
    out[0] = (double)v[0].x * w[0];
    out[1] = (double)v[0].y * w[0];
    out[2] = (double)v[0].z * w[0];
    out[3] = want_e ? (double)v[0].w * w[0] : 0.0;
#pragma unroll
    for (int k = 1; k < 8; k++) {
        out[0] = fma((double)v[k].x, w[k], out[0]);
        out[1] = fma((double)v[k].y, w[k], out[1]);
        out[2] = fma((double)v[k].z, w[k], out[2]);
        if (want_e) out[3] = fma((double)v[k].w, w[k], out[3]);
    }
}

// the 8 corners around grid coordinates (gx, gy, gz) and the fractional
// position in their cell
template <typename T4>
__device__ __forceinline__ void corners(const T4 *__restrict__ g, const DevArgs &A, double gx, double gy, double gz,
                                        T4 v[8], double &ux, double &uy, double &uz)
{
    int ix, jx, iy, jy, iz, jz;
    axis(gx, A.n0, ix, jx, ux);
    axis(gy, A.n1, iy, jy, uy);
    axis(gz, A.n2, iz, jz, uz);
    const unsigned c00 = ((unsigned)ix * A.n1 + iy) * A.n2, c01 = ((unsigned)ix * A.n1 + jy) * A.n2;
    const unsigned c10 = ((unsigned)jx * A.n1 + iy) * A.n2, c11 = ((unsigned)jx * A.n1 + jy) * A.n2;
    v[0] = g[c00 + iz]; v[1] = g[c00 + jz]; v[2] = g[c01 + iz]; v[3] = g[c01 + jz];
    v[4] = g[c10 + iz]; v[5] = g[c10 + jz]; v[6] = g[c11 + iz]; v[7] = g[c11 + jz];
}

// The 8 corners of the cell a probe last sampled, kept in registers across
// its FIRE iterations: while the probe stays in that cell (most iterations
// of a converging step) the sample needs no memory access.  The values are
// the grid's, so results equal those of gather().
template <typename T4>
struct CellCache {
    int ix = -1, iy = -1, iz = -1;
    T4 v[8];
};

typedef double pp_d2 __attribute__((ext_vector_type(2)));
typedef float pp_f4 __attribute__((ext_vector_type(4)));

// One grid corner through a buffer resource r (the grid's base in scalar
// registers, a 32-bit byte offset per load: no 64-bit address arithmetic
// per corner; the grid must be under 4 GiB).  WANT_W: the energy component
// too (mode 1); mode 0 reads x, y, z.
template <typename T4, bool WANT_W>
__device__ __forceinline__ T4 buf_corner(__amdgpu_buffer_rsrc_t r, unsigned idx)
{
    if constexpr (sizeof(T4) == 32) {
        const pp_d2 xy = __builtin_bit_cast(pp_d2, __builtin_amdgcn_raw_buffer_load_b128(r, idx * 32u, 0, 0));
        if constexpr (WANT_W) {
            const pp_d2 zw = __builtin_bit_cast(pp_d2, __builtin_amdgcn_raw_buffer_load_b128(r, idx * 32u + 16u, 0, 0));
            return T4{xy.x, xy.y, zw.x, zw.y};
        } else {
            const double z = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, idx * 32u + 16u, 0, 0));
            return T4{xy.x, xy.y, z, 0.0};
        }
    } else {
        const pp_f4 v = __builtin_bit_cast(pp_f4, __builtin_amdgcn_raw_buffer_load_b128(r, idx * 16u, 0, 0));
        return T4{v.x, v.y, v.z, WANT_W ? v.w : 0.f};
    }
}

// corners() through the cache: C.v holds the corners of the cell of
// (gx, gy, gz) afterwards; (ux, uy, uz) is the position in it.  BUF: the
// loads go through a buffer resource on g (buf_corner; the caller makes
// sure the grid is under 4 GiB).
template <typename T4, bool BUF = false, bool WANT_W = true>
__device__ __forceinline__ void corners_cached(const T4 *__restrict__ g, const DevArgs &A, double gx, double gy,
                                               double gz, CellCache<T4> &C, double &ux, double &uy, double &uz)
{
    int ix, jx, iy, jy, iz, jz;
    axis(gx, A.n0, ix, jx, ux);
    axis(gy, A.n1, iy, jy, uy);
    axis(gz, A.n2, iz, jz, uz);
    if (ix != C.ix || iy != C.iy || iz != C.iz) {
        const unsigned c00 = ((unsigned)ix * A.n1 + iy) * A.n2, c01 = ((unsigned)ix * A.n1 + jy) * A.n2;
        const unsigned c10 = ((unsigned)jx * A.n1 + iy) * A.n2, c11 = ((unsigned)jx * A.n1 + jy) * A.n2;
        if constexpr (BUF) {
            const __amdgpu_buffer_rsrc_t r =
                __builtin_amdgcn_make_buffer_rsrc(const_cast<T4 *>(g), (short)0, (int)0xFFFFFFFF, 0x00020000);
            C.v[0] = buf_corner<T4, WANT_W>(r, c00 + iz); C.v[1] = buf_corner<T4, WANT_W>(r, c00 + jz);
            C.v[2] = buf_corner<T4, WANT_W>(r, c01 + iz); C.v[3] = buf_corner<T4, WANT_W>(r, c01 + jz);
            C.v[4] = buf_corner<T4, WANT_W>(r, c10 + iz); C.v[5] = buf_corner<T4, WANT_W>(r, c10 + jz);
            C.v[6] = buf_corner<T4, WANT_W>(r, c11 + iz); C.v[7] = buf_corner<T4, WANT_W>(r, c11 + jz);
        } else {
            C.v[0] = g[c00 + iz]; C.v[1] = g[c00 + jz]; C.v[2] = g[c01 + iz]; C.v[3] = g[c01 + jz];
            C.v[4] = g[c10 + iz]; C.v[5] = g[c10 + jz]; C.v[6] = g[c11 + iz]; C.v[7] = g[c11 + jz];
        }
        C.ix = ix;
        C.iy = iy;
        C.iz = iz;
    }
}

template <typename T4, bool BUF = false, bool WANT_E = true>
__device__ __forceinline__ void gather_cached(const T4 *__restrict__ g, const DevArgs &A, double gx, double gy,
                                              double gz, double out[4], CellCache<T4> &C)
{
    double ux, uy, uz;
    corners_cached<T4, BUF, WANT_E>(g, A, gx, gy, gz, C, ux, uy, uz);
    interp8(C.v, ux, uy, uz, out, WANT_E);
}

template <typename T4>
__device__ __forceinline__ void gather(const T4 *__restrict__ g, const DevArgs &A, double gx, double gy, double gz,
                                       double out[4], bool want_e)
{
    T4 v[8];
    double tx, ty, tz;
    corners(g, A, gx, gy, gz, v, tx, ty, tz);
    interp8(v, tx, ty, tz, out, want_e);
}

// FIRE state of one probe
struct Fire {
    double rx, ry, rz, vx, vy, vz, dt, ac;
};

// Total force on the probe at S.r: grid force fe + tip springs
// (ProbeParticle.cpp:282-294 / relax.cl:37-41 in the tip frame).
template <int MODE>
__device__ __forceinline__ void total_force(const DevArgs &A, const double fe[4], double tx, double ty, double tz,
                                            double rx, double ry, double rz, double &fx, double &fy, double &fz)
{
    fx = fe[0]; fy = fe[1]; fz = fe[2];
    const double dx = rx - tx, dy = ry - ty, dz = rz - tz;
    if (MODE == 0) {
        const double l2 = fma(dx, dx, fma(dy, dy, dz * dz));
        const double il = rsq_nr(l2);
        const double sr = A.krad * fma(l2, il, -A.lrad) * il;
        fx = fx + fma(dx, sr, (dx - A.d0[0]) * A.ks[0]);
        fy = fy + fma(dy, sr, (dy - A.d0[1]) * A.ks[1]);
        fz = fz + fma(dz, sr, (dz - A.d0[2]) * A.ks[2]);
    } else {
        const double qx = fma(dz, A.rot[2], fma(dy, A.rot[1], dx * A.rot[0]));
        const double qy = fma(dz, A.rot[5], fma(dy, A.rot[4], dx * A.rot[3]));
        const double qz = fma(dz, A.rot[8], fma(dy, A.rot[7], dx * A.rot[6]));
        const double l2 = fma(qx, qx, fma(qy, qy, qz * qz));
        const double il = rsq_nr(l2);
        const double sr = A.krad * fma(l2, il, -A.lrad) * il;
        const double ax = fma(qx, sr, (qx - A.d0[0]) * A.ks[0]);
        const double ay = fma(qy, sr, (qy - A.d0[1]) * A.ks[1]);
        const double az = fma(qz, sr, (qz - A.d0[2]) * A.ks[2]) + A.surf_ff;
        fx = fx + fma(A.rot[6], az, fma(A.rot[3], ay, A.rot[0] * ax));
        fy = fy + fma(A.rot[7], az, fma(A.rot[4], ay, A.rot[1] * ax));
        fz = fz + fma(A.rot[8], az, fma(A.rot[5], ay, A.rot[2] * ax));
    }
}

// One FIRE move under force f (ProbeParticle.cpp:142-174 / relax.cl:105-126).
// Returns |f|^2.
template <int MODE>
__device__ __forceinline__ double fire_update(const DevArgs &A, double fx, double fy, double fz, Fire &S)
{
    const double ff = fma(fx, fx, fma(fy, fy, fz * fz));
    const double vv = fma(S.vx, S.vx, fma(S.vy, S.vy, S.vz * S.vz));
    const double vf = fma(S.vx, fx, fma(S.vy, fy, S.vz * fz));
    if (vf < 0) {
        S.vx = S.vy = S.vz = 0;
        S.dt = MODE == 0 ? S.dt * A.fdec : fmax(A.dtmin, S.dt * A.fdec);
        S.ac = A.a0;
    } else {
        // alpha |v| / |f| = alpha v^2 / sqrt(v^2 f^2): one reciprocal square root
        const double cf = vv > 0 ? S.ac * (vv * rsq_nr(vv * (MODE == 0 ? ff : ff + A.f2safe))) : 0.0;
        const double cv = 1 - S.ac;
        S.vx = fma(fx, cf, S.vx * cv);
        S.vy = fma(fy, cf, S.vy * cv);
        S.vz = fma(fz, cf, S.vz * cv);
        S.dt = fmin(S.dt * A.finc, A.dtmax);
        S.ac *= A.falpha;
    }
    S.vx = fma(fx, S.dt, S.vx);
    S.vy = fma(fy, S.dt, S.vy);
    S.vz = fma(fz, S.dt, S.vz);
    S.rx = fma(S.vx, S.dt, S.rx);
    S.ry = fma(S.vy, S.dt, S.ry);
    S.rz = fma(S.vz, S.dt, S.rz);
    return ff;
}

template <int MODE>
__device__ __forceinline__ double fire_move(const DevArgs &A, const double fe[4], double tx, double ty, double tz,
                                            Fire &S)
{
    double fx, fy, fz;
    total_force<MODE>(A, fe, tx, ty, tz, S.rx, S.ry, S.rz, fx, fy, fz);
    return fire_update<MODE>(A, fx, fy, fz, S);
}

// One Newton step toward the zero of the total force on the trilinear grid
// field: J dr = -F with the exact Jacobian of the interpolant in the current
// cell (derivatives of the trilinear weights mapped by dinv) plus the
// springs' Jacobian, from the cell's 8 corners v, the fractional position
// (ux, uy, uz) of r in it and the grid force fe there (interp8), all FP64.
// A step longer than max_step is cut to that length (damped Newton).
// Returns false, leaving r unchanged, when the step is not finite; st2 =
// squared length of the full step, f2 = squared total force at the old r.  `stable` tells whether the Jacobian at r
// has all eigenvalues in the left half plane (Routh-Hurwitz on its
// characteristic polynomial l^3 + a2 l^2 + a1 l + a0): only then is a zero of
// the force there a minimum that damped dynamics settle in, not a saddle or a
// maximum such as the point straight above a repulsive atom.
template <typename T4, int MODE, bool ORTHO = false>
__device__ __forceinline__ bool newton_from(const DevArgs &A, const T4 v[8], double ux, double uy, double uz,
                                            const double fe[4], double tx, double ty, double tz, double &rx,
                                            double &ry, double &rz, double max_step, bool &stable, double &st2,
                                            double &f2)
{
    double dT[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // dT[i][a] = dF_i / du_a
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int bx = k >> 2, by = (k >> 1) & 1, bz = k & 1;
        const double wx = bx ? ux : 1 - ux, wy = by ? uy : 1 - uy, wz = bz ? uz : 1 - uz;
        const double d0 = (bx ? 1.0 : -1.0) * wy * wz, d1 = wx * (by ? 1.0 : -1.0) * wz, d2 = wx * wy * (bz ? 1.0 : -1.0);
        const double c[3] = {(double)v[k].x, (double)v[k].y, (double)v[k].z};
#pragma unroll
        for (int i = 0; i < 3; i++) {
            dT[i][0] = fma(c[i], d0, dT[i][0]);
            dT[i][1] = fma(c[i], d1, dT[i][1]);
            dT[i][2] = fma(c[i], d2, dT[i][2]);
        }
    }
    double J[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)  // ORTHO: the diagonal alone, the same numbers (as to_grid)
            J[i][j] = ORTHO ? dT[i][j] * A.dinv[4 * j]
                            : fma(dT[i][2], A.dinv[6 + j], fma(dT[i][1], A.dinv[3 + j], dT[i][0] * A.dinv[j]));
    double fx, fy, fz;
    total_force<MODE>(A, fe, tx, ty, tz, rx, ry, rz, fx, fy, fz);
    f2 = fma(fx, fx, fma(fy, fy, fz * fz));
    // springs in the tip frame q = R d: a(q) = ks (q - d0) + krad (|q| - L) q / |q|,
    // force R^T a, Jacobian R^T (da/dq) R
    const double dx = rx - tx, dy = ry - ty, dz = rz - tz;
    const double *R = A.rot;
    const double q[3] = {MODE == 0 ? dx : fma(dz, R[2], fma(dy, R[1], dx * R[0])),
                         MODE == 0 ? dy : fma(dz, R[5], fma(dy, R[4], dx * R[3])),
                         MODE == 0 ? dz : fma(dz, R[8], fma(dy, R[7], dx * R[6]))};
    const double l2 = fma(q[0], q[0], fma(q[1], q[1], q[2] * q[2]));
    const double il = rsq_nr(l2);
    const double a1 = A.krad * (1.0 - A.lrad * il), a2 = A.krad * A.lrad * il * il * il;
    double Dq[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) Dq[i][j] = (i == j ? a1 + A.ks[i] : 0.0) + a2 * q[i] * q[j];
    if (MODE == 0) {  // ProbeParticle tips are not rotated
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) J[i][j] += Dq[i][j];
    } else {
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                double acc = 0;
#pragma unroll
                for (int a = 0; a < 3; a++)
#pragma unroll
                    for (int b = 0; b < 3; b++) acc = fma(R[3 * a + i] * Dq[a][b], R[3 * b + j], acc);
                J[i][j] += acc;
            }
    }
    // solve J dr = -f by the adjugate
    const double C00 = J[1][1] * J[2][2] - J[1][2] * J[2][1], C01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const double C02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    const double det = J[0][0] * C00 + J[0][1] * C01 + J[0][2] * C02;
    {
        const double b2 = -(J[0][0] + J[1][1] + J[2][2]);
        const double b1 = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) + (J[0][0] * J[2][2] - J[0][2] * J[2][0]) + C00;
        const double b0 = -det;
        stable = b2 > 0 && b0 > 0 && b2 * b1 > b0;
    }
    if (!(fabs(det) > 1e-300)) return false;
    const double id = 1.0 / det;
    const double b0 = -fx, b1 = -fy, b2 = -fz;
    const double r0 = (C00 * b0 + (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * b1 + (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * b2) * id;
    const double r1 = (C01 * b0 + (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * b1 + (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * b2) * id;
    const double r2 = (C02 * b0 + (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * b1 + (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * b2) * id;
    st2 = fma(r0, r0, fma(r1, r1, r2 * r2));
    if (!(st2 < 1e6)) return false;  // not finite (or absurd)
    // damped: a step longer than max_step is cut to that length
    const double sc = st2 > max_step * max_step ? max_step * rsq_nr(st2) : 1.0;
    rx = fma(r0, sc, rx); ry = fma(r1, sc, ry); rz = fma(r2, sc, rz);
    return true;
}

template <typename T4, int MODE>
__device__ __forceinline__ bool newton_step(const DevArgs &A, const T4 *__restrict__ g, double tx, double ty,
                                            double tz, double &rx, double &ry, double &rz, double max_step,
                                            bool &stable)
{
    double gx, gy, gz, ux, uy, uz, st2, f2, fe[4];
    to_grid(A, rx, ry, rz, gx, gy, gz);
    T4 v[8];
    corners(g, A, gx, gy, gz, v, ux, uy, uz);
    interp8(v, ux, uy, uz, fe, false);
    return newton_from<T4, MODE>(A, v, ux, uy, uz, fe, tx, ty, tz, rx, ry, rz, max_step, stable, st2, f2);
}

__device__ __forceinline__ void store_pos(const DevArgs &A, long o, double rx, double ry, double rz)
{
    if (!A.pos) return;
    if (A.pos_f32) {
        float *p = (float *)A.pos + 3 * o;
        p[0] = (float)rx; p[1] = (float)ry; p[2] = (float)rz;
    } else {
        double *p = (double *)A.pos + 3 * o;
        p[0] = rx; p[1] = ry; p[2] = rz;
    }
}

// Per-position status, the optional output the reference has no counterpart
// for (relaxTipStroke's itr > maxIters test, ProbeParticle.cpp:658-661, can
// never fire): bit 0 = the iteration cap ended the step without meeting the
// stop rule, bit 1 = a non-finite position or grid force (a NaN grid or a
// diverging probe); 0 = converged.
__device__ __forceinline__ void store_status(const DevArgs &A, long o, bool capped, double rx, double ry, double rz,
                                             const double fe[4])
{
    if (!A.status) return;
    const bool finite = isfinite(rx) && isfinite(ry) && isfinite(rz) && isfinite(fe[0]) && isfinite(fe[1]) && isfinite(fe[2]);
    A.status[o] = (unsigned char)((capped ? 1 : 0) | (finite ? 0 : 2));
}

__device__ __forceinline__ unsigned char status_byte(bool capped, double rx, double ry, double rz, const double fe[4])
{
    const bool finite = isfinite(rx) && isfinite(ry) && isfinite(rz) && isfinite(fe[0]) && isfinite(fe[1]) && isfinite(fe[2]);
    return (unsigned char)((capped ? 1 : 0) | (finite ? 0 : 2));
}

// Output records staged per lane in LDS.  A lane finishes a z step every
// ~100 force evaluations, and its column's records lie nz apart from its
// neighbours', so records written one at a time reach memory as partial
// lines (c1: 5x the output bytes).  Staged, the records of K consecutive
// steps of a column go out together: one contiguous chunk per output array.
// K = 1 stores directly.  Layout (T threads, index k*T + t): force float4,
// then x, y, z of the position (FP64), iteration count, status byte.
template <int K>
struct Stage {
    float4 *f;
    double *p;
    int *it;
    unsigned char *st;
    int T, t;
};

constexpr int kStageRecordBytes = 16 + 24 + 4 + 1;

template <int K>
__device__ __forceinline__ Stage<K> make_stage(unsigned char *lds, int T, int t)
{
    Stage<K> G;
    G.f = (float4 *)lds;
    G.p = (double *)(lds + 16 * K * T);
    G.it = (int *)(lds + 40 * K * T);
    G.st = lds + 44 * K * T;
    G.T = T;
    G.t = t;
    return G;
}

// The record of step s of a column (output offset o) into the lane's
// slot s % K; the slots go out when the chunk is full or s is the last step.
// Copy a lane's K staged values (slots J..K-1, T lanes apart in LDS) to
// dst[J..K-1]: every LDS read is issued before the first store, and the
// values live in registers (a local array here was kept in scratch memory,
// stored and read back at every flush).
template <typename V, int J, int K>
struct FlushChunk {
    static __device__ __forceinline__ void run(V *dst, const V *src, int T, int t)
    {
        const V v = src[J * T + t];
        FlushChunk<V, J + 1, K>::run(dst, src, T, t);
        dst[J] = v;
    }
};
template <typename V, int K>
struct FlushChunk<V, K, K> {
    static __device__ __forceinline__ void run(V *, const V *, int, int) {}
};

template <int K>
__device__ __forceinline__ void put_record(const DevArgs &A, Stage<K> &G, long o, int s, double rx, double ry, double rz,
                                           float4 f, int iters, unsigned char status)
{
    if (K == 1) {
        store_pos(A, o, rx, ry, rz);
        if (A.force) A.force[o] = f;
        if (A.iters) A.iters[o] = iters;
        if (A.status) A.status[o] = status;
        return;
    }
    const int k = s % K;
    const int KT = K * G.T;
    const int i = k * G.T + G.t;
    G.f[i] = f;
    G.p[i] = rx; G.p[KT + i] = ry; G.p[2 * KT + i] = rz;
    G.it[i] = iters;
    G.st[i] = status;
    if (K >= 8 && k == K - 1) {  // a full chunk: fixed trip counts, so the LDS reads are issued ahead of the
                                 // stores (K >= 8 only: the batch kernels, K = 4, would lose a wave per SIMD
                                 // to the registers this takes)
        const long o0 = o - (K - 1);
#pragma unroll
        for (int j = 0; j < K; j++) {
            const int q = j * G.T + G.t;
            store_pos(A, o0 + j, G.p[q], G.p[KT + q], G.p[2 * KT + q]);
        }
        if (A.force) FlushChunk<float4, 0, K>::run(A.force + o0, G.f, G.T, G.t);
        if (A.iters) FlushChunk<int, 0, K>::run(A.iters + o0, G.it, G.T, G.t);
        if (A.status) FlushChunk<unsigned char, 0, K>::run(A.status + o0, G.st, G.T, G.t);
    } else if (k == K - 1 || s == A.nz - 1) {  // a chunk of the batch kernels, or a column's last, partial one
        const long o0 = o - k;
        for (int j = 0; j <= k; j++) {
            const int q = j * G.T + G.t;
            store_pos(A, o0 + j, G.p[q], G.p[KT + q], G.p[2 * KT + q]);
            if (A.force) A.force[o0 + j] = G.f[q];
            if (A.iters) A.iters[o0 + j] = G.it[q];
            if (A.status) A.status[o0 + j] = G.st[q];
        }
    }
}

// grid force rotated into the tip frame (relax.cl:446-452), .w = energy
__device__ __forceinline__ float4 tip_frame_force(const DevArgs &A, const double fe[4])
{
    return make_float4((float)fma(fe[2], A.rot[2], fma(fe[1], A.rot[1], fe[0] * A.rot[0])),
                       (float)fma(fe[2], A.rot[5], fma(fe[1], A.rot[4], fe[0] * A.rot[3])),
                       (float)fma(fe[2], A.rot[8], fma(fe[1], A.rot[7], fe[0] * A.rot[6])), (float)fe[3]);
}

// Start of the stroke: tip step 0 and the probe at tip + rotated d0.
struct Stroke {
    double bx, by, bz;  // tip position of step 0
    long col0;
};

__device__ __forceinline__ Stroke stroke_at(const DevArgs &A, int i, int j)
{
    Stroke c;
    c.bx = A.p0[0] + i * A.da[0] + j * A.db[0];
    c.by = A.p0[1] + i * A.da[1] + j * A.db[1];
    c.bz = A.p0[2] + i * A.da[2] + j * A.db[2];
    c.col0 = ((long)i * A.ny + j) * A.nz;
    return c;
}

__device__ __forceinline__ void tip_pos(const DevArgs &A, const Stroke &c, int s, bool live, double &x, double &y,
                                        double &z)
{
    if (A.tips && live) {
        const double *p = A.tips + 3 * (c.col0 + s);
        x = p[0]; y = p[1]; z = p[2];
    } else {
        x = c.bx + s * A.dstep[0]; y = c.by + s * A.dstep[1]; z = c.bz + s * A.dstep[2];
    }
}

template <int MODE>
__device__ __forceinline__ void probe_offset(const DevArgs &A, double &ox, double &oy, double &oz)
{
    if (MODE == 0) {
        ox = A.d0[0]; oy = A.d0[1]; oz = A.d0[2];
    } else {
        ox = fma(A.rot[6], A.d0[2], fma(A.rot[3], A.d0[1], A.rot[0] * A.d0[0]));
        oy = fma(A.rot[7], A.d0[2], fma(A.rot[4], A.d0[1], A.rot[1] * A.d0[0]));
        oz = fma(A.rot[8], A.d0[2], fma(A.rot[5], A.d0[1], A.rot[2] * A.d0[0]));
    }
}

// Next stroke step: tip moves to step s, the probe follows (probeStart
// semantics of relaxTipStroke, ProbeParticle.cpp:644-656; mode 1 always
// continues, relax.cl:434-436) and the FIRE velocity resets.
template <int MODE>
__device__ __forceinline__ void next_step(const DevArgs &A, const Stroke &c, int s, bool live, double &tx, double &ty,
                                          double &tz, Fire &S)
{
    double nx_, ny_, nz_;
    tip_pos(A, c, s, live, nx_, ny_, nz_);
    if (MODE == 0 && A.probe_start == 0) {
        S.rx = nx_ + A.d0[0]; S.ry = ny_ + A.d0[1]; S.rz = nz_ + A.d0[2];
    } else if (MODE == 1 || A.probe_start == 1) {
        S.rx += nx_ - tx; S.ry += ny_ - ty; S.rz += nz_ - tz;
    }
    tx = nx_; ty = ny_; tz = nz_;
    S.vx = S.vy = S.vz = 0;
    if (MODE == 0) {
        S.dt = A.dtmax;
        S.ac = A.a0;
    }
}

}  // namespace

}  // namespace pp
