// FIRE relaxation, device path: the same semantics as relax.hip (see the
// citations there) with a leaner instruction stream, because each pixel
// column is a serial chain of ~10^4 dependent force evaluations and the
// per-iteration latency sets the kernel time:
//   * FMA contraction on (this path is not bit-matched to the CPU);
//   * 1/l and sqrt(vv/ff) from v_rsq_f64 plus two Newton steps instead of
//     IEEE divide / square root sequences;
//   * 32-bit grid indices, ProbeParticle or OpenCL semantics chosen at
//     compile time, per-lane tip base so the loop touches few uniforms.
// Lanes run independent state machines over (z step, iteration); an
// LDS-staged lock-step variant (relax_lds_kernel) is kept as an option.

#include <cstdlib>

#include "pp_common.hpp"

#pragma clang fp contract(fast)

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
    unsigned long long *counters;
};

__device__ __forceinline__ double rsq_nr(double x)
{
    double y = __builtin_amdgcn_rsq(x);
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const double h = 0.5 * x * y;
        y = fma(y, fma(-h, y, 0.5), y);
    }
    return y;
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

template <typename T4>
__device__ __forceinline__ void gather(const T4 *__restrict__ g, const DevArgs &A, double gx, double gy, double gz,
                                       double out[4], bool want_e)
{
    int ix, jx, iy, jy, iz, jz;
    double tx, ty, tz;
    axis(gx, A.n0, ix, jx, tx);
    axis(gy, A.n1, iy, jy, ty);
    axis(gz, A.n2, iz, jz, tz);
    const double mx = 1 - tx, my = 1 - ty, mz = 1 - tz;
    const unsigned c00 = ((unsigned)ix * A.n1 + iy) * A.n2, c01 = ((unsigned)ix * A.n1 + jy) * A.n2;
    const unsigned c10 = ((unsigned)jx * A.n1 + iy) * A.n2, c11 = ((unsigned)jx * A.n1 + jy) * A.n2;
    const T4 v0 = g[c00 + iz], v1 = g[c00 + jz], v2 = g[c01 + iz], v3 = g[c01 + jz];
    const T4 v4 = g[c10 + iz], v5 = g[c10 + jz], v6 = g[c11 + iz], v7 = g[c11 + jz];
    const double w00 = mx * my, w01 = mx * ty, w10 = tx * my, w11 = tx * ty;
    const double w[8] = {w00 * mz, w00 * tz, w01 * mz, w01 * tz, w10 * mz, w10 * tz, w11 * mz, w11 * tz};
    const T4 v[8] = {v0, v1, v2, v3, v4, v5, v6, v7};
    out[0] = out[1] = out[2] = out[3] = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        out[0] += (double)v[k].x * w[k];
        out[1] += (double)v[k].y * w[k];
        out[2] += (double)v[k].z * w[k];
        if (want_e) out[3] += (double)v[k].w * w[k];
    }
}

template <typename T4, int MODE>
__global__ __launch_bounds__(64) void relax_dev_kernel(DevArgs A, const T4 *__restrict__ grid)
{
    const int per_xcd = (A.ntiles + 7) / 8;
    const int b = blockIdx.x;
    const int tile = (b % 8) * per_xcd + b / 8;
    if (tile >= A.ntiles) return;
    const int ti = tile / A.tiles_y, tj = tile % A.tiles_y;
    const int lane = threadIdx.x;
    const int i = ti * 8 + lane / 8, j = tj * 8 + lane % 8;
    const bool live = i < A.nx && j < A.ny;
    const int nz = A.nz;
    const long col0 = ((long)i * A.ny + j) * nz;

    // tip position of step s: base + s*dstep (or the explicit array)
    double bx = A.p0[0] + i * A.da[0] + j * A.db[0];
    double by = A.p0[1] + i * A.da[1] + j * A.db[1];
    double bz = A.p0[2] + i * A.da[2] + j * A.db[2];
    auto tip_at = [&](int s, double &x, double &y, double &z) {
        if (A.tips) {
            const double *p = A.tips + 3 * (col0 + s);
            x = p[0]; y = p[1]; z = p[2];
        } else {
            x = bx + s * A.dstep[0]; y = by + s * A.dstep[1]; z = bz + s * A.dstep[2];
        }
    };

    double tx = 0, ty = 0, tz = 0;
    int s = live ? 0 : nz;
    if (live) tip_at(0, tx, ty, tz);
    double ox, oy, oz;  // equilibrium offset in the world frame
    if (MODE == 0) {
        ox = A.d0[0]; oy = A.d0[1]; oz = A.d0[2];
    } else {
        ox = A.rot[0] * A.d0[0] + A.rot[3] * A.d0[1] + A.rot[6] * A.d0[2];
        oy = A.rot[1] * A.d0[0] + A.rot[4] * A.d0[1] + A.rot[7] * A.d0[2];
        oz = A.rot[2] * A.d0[0] + A.rot[5] * A.d0[1] + A.rot[8] * A.d0[2];
    }
    double rx = tx + ox, ry = ty + oy, rz = tz + oz;
    double vx = 0, vy = 0, vz = 0;
    double dt = A.dtmax, ac = A.a0;
    int it = 0, itr = 0;
    bool sample = false;  // mode 0: next gather is the output sample
    unsigned long long nevals = 0, npos = 0;

    while (s < nz) {
        double fe[4];
        const double gx = rx * A.dinv[0] + ry * A.dinv[1] + rz * A.dinv[2];
        const double gy = rx * A.dinv[3] + ry * A.dinv[4] + rz * A.dinv[5];
        const double gz = rx * A.dinv[6] + ry * A.dinv[7] + rz * A.dinv[8];
        gather(grid, A, gx, gy, gz, fe, MODE == 1);
        nevals++;
        bool done = false;
        if (MODE == 0 && sample) {
            const long o = col0 + s;
            if (A.pos) {
                if (A.pos_f32) {
                    float *p = (float *)A.pos + 3 * o;
                    p[0] = (float)rx; p[1] = (float)ry; p[2] = (float)rz;
                } else {
                    double *p = (double *)A.pos + 3 * o;
                    p[0] = rx; p[1] = ry; p[2] = rz;
                }
            }
            if (A.force) A.force[o] = make_float4((float)fe[0], (float)fe[1], (float)fe[2], 0.f);
            if (A.iters) A.iters[o] = itr;
            done = true;
        } else {
            double fx = fe[0], fy = fe[1], fz = fe[2];
            const double dx = rx - tx, dy = ry - ty, dz = rz - tz;
            if (MODE == 0) {
                const double l2 = dx * dx + dy * dy + dz * dz;
                const double il = rsq_nr(l2);
                const double sr = A.krad * (l2 * il - A.lrad) * il;
                fx += dx * sr + (dx - A.d0[0]) * A.ks[0];
                fy += dy * sr + (dy - A.d0[1]) * A.ks[1];
                fz += dz * sr + (dz - A.d0[2]) * A.ks[2];
            } else {
                const double qx = dx * A.rot[0] + dy * A.rot[1] + dz * A.rot[2];
                const double qy = dx * A.rot[3] + dy * A.rot[4] + dz * A.rot[5];
                const double qz = dx * A.rot[6] + dy * A.rot[7] + dz * A.rot[8];
                const double l2 = qx * qx + qy * qy + qz * qz;
                const double il = rsq_nr(l2);
                const double sr = A.krad * (l2 * il - A.lrad) * il;
                const double ax = (qx - A.d0[0]) * A.ks[0] + qx * sr;
                const double ay = (qy - A.d0[1]) * A.ks[1] + qy * sr;
                const double az = (qz - A.d0[2]) * A.ks[2] + qz * sr + A.surf_ff;
                fx += A.rot[0] * ax + A.rot[3] * ay + A.rot[6] * az;
                fy += A.rot[1] * ax + A.rot[4] * ay + A.rot[7] * az;
                fz += A.rot[2] * ax + A.rot[5] * ay + A.rot[8] * az;
            }
            const double ff = fx * fx + fy * fy + fz * fz;
            const double vv = vx * vx + vy * vy + vz * vz;
            const double vf = vx * fx + vy * fy + vz * fz;
            if (vf < 0) {
                vx = vy = vz = 0;
                dt = MODE == 0 ? dt * A.fdec : fmax(A.dtmin, dt * A.fdec);
                ac = A.a0;
            } else {
                const double cf = vv > 0 ? ac * (vv * rsq_nr(vv)) * rsq_nr(MODE == 0 ? ff : ff + A.f2safe) : 0.0;
                const double cv = 1 - ac;
                vx = vx * cv + fx * cf;
                vy = vy * cv + fy * cf;
                vz = vz * cv + fz * cf;
                dt = fmin(dt * A.finc, A.dtmax);
                ac *= A.falpha;
            }
            vx += fx * dt; vy += fy * dt; vz += fz * dt;
            rx += vx * dt; ry += vy * dt; rz += vz * dt;
            it++;
            const bool conv = ff < A.conv_f2;
            if (conv || it >= A.max_iters) {
                if (MODE == 0) {
                    itr = conv ? it - 1 : A.max_iters;
                    sample = true;
                } else {
                    const long o = col0 + s;
                    if (A.force)
                        A.force[o] = make_float4((float)(fe[0] * A.rot[0] + fe[1] * A.rot[1] + fe[2] * A.rot[2]),
                                                 (float)(fe[0] * A.rot[3] + fe[1] * A.rot[4] + fe[2] * A.rot[5]),
                                                 (float)(fe[0] * A.rot[6] + fe[1] * A.rot[7] + fe[2] * A.rot[8]),
                                                 (float)fe[3]);
                    if (A.pos) {
                        if (A.pos_f32) {
                            float *p = (float *)A.pos + 3 * o;
                            p[0] = (float)rx; p[1] = (float)ry; p[2] = (float)rz;
                        } else {
                            double *p = (double *)A.pos + 3 * o;
                            p[0] = rx; p[1] = ry; p[2] = rz;
                        }
                    }
                    if (A.iters) A.iters[o] = it;
                    done = true;
                }
            }
        }
        if (done) {
            npos++;
            s++;
            if (s < nz) {
                double nx_, ny_, nz_;
                tip_at(s, nx_, ny_, nz_);
                if (MODE == 0 && A.probe_start == 0) {
                    rx = nx_ + A.d0[0]; ry = ny_ + A.d0[1]; rz = nz_ + A.d0[2];
                } else if (MODE == 1 || A.probe_start == 1) {
                    rx += nx_ - tx; ry += ny_ - ty; rz += nz_ - tz;
                }
                tx = nx_; ty = ny_; tz = nz_;
                vx = vy = vz = 0;
                if (MODE == 0) {
                    dt = A.dtmax;
                    ac = A.a0;
                }
                it = 0;
                sample = false;
            }
        }
    }

    if (A.counters) {
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

// ---------------------------------------------------------------------------
// LDS-staged variant.  The lanes of a wave relax z step s together; before
// the step the wave reduces its probes' bounding box (grid index space),
// copies that block of the grid (plus a margin) into LDS as three
// structure-of-arrays planes, and every gather whose 8 corners fall inside
// the block reads LDS; the others read global memory.  The per-lane
// arithmetic is that of relax_dev_kernel, so the results are identical.
constexpr int kMargin = 2;

template <typename T>
__device__ __forceinline__ T wave_min(T v)
{
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    return v;
}
template <typename T>
__device__ __forceinline__ T wave_max(T v)
{
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ int wrapi(int i, int n)
{
    i %= n;
    return i < 0 ? i + n : i;
}

template <typename T4, int MODE>
__global__ __launch_bounds__(64) void relax_lds_kernel(DevArgs A, const T4 *__restrict__ grid, int cap)
{
    using T = decltype(T4::x);
    extern __shared__ __align__(16) unsigned char smem[];
    T *LX = (T *)smem;
    T *LY = LX + cap;
    T *LZ = LY + cap;
    T *LW = LZ + cap;  // energy plane, mode 1 only (cap counts 4 planes then)

    const int per_xcd = (A.ntiles + 7) / 8;
    const int b = blockIdx.x;
    const int tile = (b % 8) * per_xcd + b / 8;
    if (tile >= A.ntiles) return;  // uniform per block
    const int ti = tile / A.tiles_y, tj = tile % A.tiles_y;
    const int lane = threadIdx.x;
    const int i = ti * 8 + lane / 8, j = tj * 8 + lane % 8;
    const bool live = i < A.nx && j < A.ny;
    const int nz = A.nz;
    const long col0 = ((long)i * A.ny + j) * nz;
    const double bx = A.p0[0] + i * A.da[0] + j * A.db[0];
    const double by = A.p0[1] + i * A.da[1] + j * A.db[1];
    const double bz = A.p0[2] + i * A.da[2] + j * A.db[2];
    auto tip_at = [&](int s, double &x, double &y, double &z) {
        if (A.tips && live) {
            const double *p = A.tips + 3 * (col0 + s);
            x = p[0]; y = p[1]; z = p[2];
        } else {
            x = bx + s * A.dstep[0]; y = by + s * A.dstep[1]; z = bz + s * A.dstep[2];
        }
    };
    double tx, ty, tz;
    tip_at(0, tx, ty, tz);
    double ox, oy, oz;
    if (MODE == 0) {
        ox = A.d0[0]; oy = A.d0[1]; oz = A.d0[2];
    } else {
        ox = A.rot[0] * A.d0[0] + A.rot[3] * A.d0[1] + A.rot[6] * A.d0[2];
        oy = A.rot[1] * A.d0[0] + A.rot[4] * A.d0[1] + A.rot[7] * A.d0[2];
        oz = A.rot[2] * A.d0[0] + A.rot[5] * A.d0[1] + A.rot[8] * A.d0[2];
    }
    double rx = tx + ox, ry = ty + oy, rz = tz + oz;
    double vx = 0, vy = 0, vz = 0, dt = A.dtmax, ac = A.a0;
    unsigned long long nevals = 0, npos = 0;

    for (int s = 0; s < nz; s++) {
        if (s > 0) {
            double nx_, ny_, nz_;
            tip_at(s, nx_, ny_, nz_);
            if (MODE == 0 && A.probe_start == 0) {
                rx = nx_ + A.d0[0]; ry = ny_ + A.d0[1]; rz = nz_ + A.d0[2];
            } else if (MODE == 1 || A.probe_start == 1) {
                rx += nx_ - tx; ry += ny_ - ty; rz += nz_ - tz;
            }
            tx = nx_; ty = ny_; tz = nz_;
            vx = vy = vz = 0;
            if (MODE == 0) { dt = A.dtmax; ac = A.a0; }
        }
        // ---- stage the block around the wave's probes
        const double g0x = rx * A.dinv[0] + ry * A.dinv[1] + rz * A.dinv[2];
        const double g0y = rx * A.dinv[3] + ry * A.dinv[4] + rz * A.dinv[5];
        const double g0z = rx * A.dinv[6] + ry * A.dinv[7] + rz * A.dinv[8];
        const float big = 3.0e38f;
        const float mnx = wave_min(live ? (float)g0x : big), mxx = wave_max(live ? (float)g0x : -big);
        const float mny = wave_min(live ? (float)g0y : big), mxy = wave_max(live ? (float)g0y : -big);
        const float mnz = wave_min(live ? (float)g0z : big), mxz = wave_max(live ? (float)g0z : -big);
        // block only when all bounds are finite and small (a diverging probe
        // sends its wave to global memory instead)
        const bool sane = mnx > -1.0e7f && mxx < 1.0e7f && mny > -1.0e7f && mxy < 1.0e7f && mnz > -1.0e7f && mxz < 1.0e7f &&
                          mxx - mnx < 512.f && mxy - mny < 512.f && mxz - mnz < 512.f;
        int lox = 0, loy = 0, loz = 0, ex = 0, ey = 0, ez = 0;
        if (sane) {
            lox = (int)floorf(mnx) - kMargin; loy = (int)floorf(mny) - kMargin; loz = (int)floorf(mnz) - kMargin;
            ex = (int)floorf(mxx) + 1 + kMargin - lox + 1;
            ey = (int)floorf(mxy) + 1 + kMargin - loy + 1;
            ez = (int)floorf(mxz) + 1 + kMargin - loz + 1;
        }
        const bool staged = sane && (long)ex * ey * ez <= cap;
        if (staged) {
            const int vol = ex * ey * ez;
            for (int q = lane; q < vol; q += 64) {
                const int c = q % ez, t = q / ez;
                const int bb = t % ey, a = t / ey;
                const unsigned gi = ((unsigned)wrapi(lox + a, A.n0) * A.n1 + wrapi(loy + bb, A.n1)) * A.n2 + wrapi(loz + c, A.n2);
                const T4 v = grid[gi];
                LX[q] = v.x; LY[q] = v.y; LZ[q] = v.z;
                if (MODE == 1) LW[q] = v.w;
            }
        }
        __syncthreads();

        // ---- relax this step, all lanes together
        int st = live ? 0 : 2;  // 0 = FIRE, 1 = output sample pending (mode 0), 2 = done
        int it = 0, itr = 0;
        double ofx = 0, ofy = 0, ofz = 0, ofe = 0;
        while (__ballot(st < 2) != 0) {
            if (st < 2) {
                const double gx = rx * A.dinv[0] + ry * A.dinv[1] + rz * A.dinv[2];
                const double gy = rx * A.dinv[3] + ry * A.dinv[4] + rz * A.dinv[5];
                const double gz = rx * A.dinv[6] + ry * A.dinv[7] + rz * A.dinv[8];
                const double flx = floor(gx), fly = floor(gy), flz = floor(gz);
                double fe[4];
                // inside test in floating point (false for NaN), before any int conversion
                if (staged && flx >= lox && flx + 1 < lox + ex && fly >= loy && fly + 1 < loy + ey && flz >= loz &&
                    flz + 1 < loz + ez) {
                    const int a = (int)flx - lox, bb = (int)fly - loy, c = (int)flz - loz;
                    const double wx = gx - flx, wy = gy - fly, wz = gz - flz;
                    const double mx = 1 - wx, my = 1 - wy, mz = 1 - wz;
                    const int q00 = (a * ey + bb) * ez + c, q01 = q00 + ez, q10 = q00 + ey * ez, q11 = q10 + ez;
                    const int q[8] = {q00, q00 + 1, q01, q01 + 1, q10, q10 + 1, q11, q11 + 1};
                    const double w00 = mx * my, w01 = mx * wy, w10 = wx * my, w11 = wx * wy;
                    const double w[8] = {w00 * mz, w00 * wz, w01 * mz, w01 * wz, w10 * mz, w10 * wz, w11 * mz, w11 * wz};
                    fe[0] = fe[1] = fe[2] = fe[3] = 0;
#pragma unroll
                    for (int k = 0; k < 8; k++) {
                        fe[0] += (double)LX[q[k]] * w[k];
                        fe[1] += (double)LY[q[k]] * w[k];
                        fe[2] += (double)LZ[q[k]] * w[k];
                        if (MODE == 1) fe[3] += (double)LW[q[k]] * w[k];
                    }
                } else {
                    gather(grid, A, gx, gy, gz, fe, MODE == 1);
                }
                nevals++;
                if (MODE == 0 && st == 1) {
                    ofx = fe[0]; ofy = fe[1]; ofz = fe[2];
                    st = 2;
                } else {
                    double fx = fe[0], fy = fe[1], fz = fe[2];
                    const double dx = rx - tx, dy = ry - ty, dz = rz - tz;
                    if (MODE == 0) {
                        const double l2 = dx * dx + dy * dy + dz * dz;
                        const double il = rsq_nr(l2);
                        const double sr = A.krad * (l2 * il - A.lrad) * il;
                        fx += dx * sr + (dx - A.d0[0]) * A.ks[0];
                        fy += dy * sr + (dy - A.d0[1]) * A.ks[1];
                        fz += dz * sr + (dz - A.d0[2]) * A.ks[2];
                    } else {
                        const double qx = dx * A.rot[0] + dy * A.rot[1] + dz * A.rot[2];
                        const double qy = dx * A.rot[3] + dy * A.rot[4] + dz * A.rot[5];
                        const double qz = dx * A.rot[6] + dy * A.rot[7] + dz * A.rot[8];
                        const double l2 = qx * qx + qy * qy + qz * qz;
                        const double il = rsq_nr(l2);
                        const double sr = A.krad * (l2 * il - A.lrad) * il;
                        const double ax = (qx - A.d0[0]) * A.ks[0] + qx * sr;
                        const double ay = (qy - A.d0[1]) * A.ks[1] + qy * sr;
                        const double az = (qz - A.d0[2]) * A.ks[2] + qz * sr + A.surf_ff;
                        fx += A.rot[0] * ax + A.rot[3] * ay + A.rot[6] * az;
                        fy += A.rot[1] * ax + A.rot[4] * ay + A.rot[7] * az;
                        fz += A.rot[2] * ax + A.rot[5] * ay + A.rot[8] * az;
                    }
                    const double ff = fx * fx + fy * fy + fz * fz;
                    const double vv = vx * vx + vy * vy + vz * vz;
                    const double vf = vx * fx + vy * fy + vz * fz;
                    if (vf < 0) {
                        vx = vy = vz = 0;
                        dt = MODE == 0 ? dt * A.fdec : fmax(A.dtmin, dt * A.fdec);
                        ac = A.a0;
                    } else {
                        const double cf = vv > 0 ? ac * (vv * rsq_nr(vv)) * rsq_nr(MODE == 0 ? ff : ff + A.f2safe) : 0.0;
                        const double cv = 1 - ac;
                        vx = vx * cv + fx * cf;
                        vy = vy * cv + fy * cf;
                        vz = vz * cv + fz * cf;
                        dt = fmin(dt * A.finc, A.dtmax);
                        ac *= A.falpha;
                    }
                    vx += fx * dt; vy += fy * dt; vz += fz * dt;
                    rx += vx * dt; ry += vy * dt; rz += vz * dt;
                    it++;
                    const bool conv = ff < A.conv_f2;
                    if (conv || it >= A.max_iters) {
                        if (MODE == 0) {
                            itr = conv ? it - 1 : A.max_iters;
                            st = 1;
                        } else {
                            ofx = fe[0] * A.rot[0] + fe[1] * A.rot[1] + fe[2] * A.rot[2];
                            ofy = fe[0] * A.rot[3] + fe[1] * A.rot[4] + fe[2] * A.rot[5];
                            ofz = fe[0] * A.rot[6] + fe[1] * A.rot[7] + fe[2] * A.rot[8];
                            ofe = fe[3];
                            itr = it;
                            st = 2;
                        }
                    }
                }
            }
        }
        if (live) {
            const long o = col0 + s;
            if (A.pos) {
                if (A.pos_f32) {
                    float *p = (float *)A.pos + 3 * o;
                    p[0] = (float)rx; p[1] = (float)ry; p[2] = (float)rz;
                } else {
                    double *p = (double *)A.pos + 3 * o;
                    p[0] = rx; p[1] = ry; p[2] = rz;
                }
            }
            if (A.force) A.force[o] = make_float4((float)ofx, (float)ofy, (float)ofz, (float)ofe);
            if (A.iters) A.iters[o] = itr;
            npos++;
        }
        __syncthreads();
    }

    if (A.counters) {
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

}  // namespace

// 0 on launch, 1 when the parameters need the generic kernel (leap-frog,
// max_iters < 1), negative on error.
int launch_relax_fast(const RelaxLaunch &L, const void *grid, bool f64, hipStream_t st)
{
    if (L.rp.alg != 1 || L.rp.max_iters < 1) return 1;
    PP_REQUIRE(L.rp.mode == 0 || L.rp.mode == 1, "relax mode must be 0 or 1");
    PP_REQUIRE(L.rp.probe_start >= -1 && L.rp.probe_start <= 1, "probeStart must be -1, 0 or 1");
    PP_REQUIRE((long)L.g.n[0] * L.g.n[1] * L.g.n[2] < (1L << 31), "grid too large for 32-bit indices");
    PP_REQUIRE(L.scan.nx >= 0 && L.scan.ny >= 0 && L.scan.nz >= 0, "scan sizes must be >= 0");
    DevArgs A;
    A.n0 = L.g.n[0]; A.n1 = L.g.n[1]; A.n2 = L.g.n[2];
    A.nx = L.scan.nx; A.ny = L.scan.ny; A.nz = L.scan.nz;
    A.tiles_y = (A.ny + 7) / 8;
    A.ntiles = ((A.nx + 7) / 8) * A.tiles_y;
    A.max_iters = L.rp.max_iters;
    A.probe_start = L.rp.probe_start;
    for (int k = 0; k < 9; k++) { A.dinv[k] = L.g.dinv[k]; A.rot[k] = L.tip.rot[k]; }
    A.conv_f2 = L.rp.conv_f2; A.dtmax = L.rp.dt; A.a0 = L.rp.damping;
    A.finc = L.rp.finc; A.fdec = L.rp.fdec; A.falpha = L.rp.falpha;
    A.dtmin = L.rp.dtmin_frac * L.rp.dt; A.f2safe = L.rp.f2safe;
    for (int k = 0; k < 3; k++) {
        A.ks[k] = L.tip.kspring[k]; A.d0[k] = L.tip.d0[k];
        A.p0[k] = L.scan.p0[k]; A.da[k] = L.scan.da[k]; A.db[k] = L.scan.db[k]; A.dstep[k] = L.scan.dstep[k];
    }
    A.krad = L.tip.krad; A.lrad = L.tip.lrad; A.surf_ff = L.tip.surf_ff;
    A.tips = L.tips;
    A.pos = L.out.pos; A.pos_f32 = L.out.pos_f32; A.force = (float4 *)L.out.force; A.iters = L.out.iters;
    A.counters = L.out.counters;
    if (A.ntiles == 0 || A.nz == 0) return 0;
    const long nblk = ((A.ntiles + 7L) / 8) * 8;
    PP_REQUIRE(nblk < (1L << 31), "scan too large");
    dim3 grd((unsigned)nblk), blk(64);
    // LDS-staged lock-step kernel: off unless PPAFM_RELAX_LDS_KB > 0 (38 KB
    // lets 4 waves share a CU's 160 KB).  Lock-step makes a wave wait for its
    // slowest lane at every z step; on the 120-atom scan that lengthens the
    // slowest wave's chain from 10,411 to 23,650 evaluations, more than the
    // ~30% per-evaluation gain of LDS gathers buys back.
    static const char *lds_env = getenv("PPAFM_RELAX_LDS_KB");
    const int lds_kb = lds_env ? atoi(lds_env) : 0;
    if (lds_kb > 0) {
        const int esz = f64 ? 8 : 4;
        const int planes = L.rp.mode == 1 ? 4 : 3;
        const int cap = (lds_kb * 1024) / (planes * esz);
        const size_t bytes = (size_t)cap * planes * esz;
        if (f64) {
            if (L.rp.mode == 0) hipLaunchKernelGGL((relax_lds_kernel<double4, 0>), grd, blk, bytes, st, A, (const double4 *)grid, cap);
            else hipLaunchKernelGGL((relax_lds_kernel<double4, 1>), grd, blk, bytes, st, A, (const double4 *)grid, cap);
        } else {
            if (L.rp.mode == 0) hipLaunchKernelGGL((relax_lds_kernel<float4, 0>), grd, blk, bytes, st, A, (const float4 *)grid, cap);
            else hipLaunchKernelGGL((relax_lds_kernel<float4, 1>), grd, blk, bytes, st, A, (const float4 *)grid, cap);
        }
        return 0;
    }
    if (f64) {
        if (L.rp.mode == 0) hipLaunchKernelGGL((relax_dev_kernel<double4, 0>), grd, blk, 0, st, A, (const double4 *)grid);
        else hipLaunchKernelGGL((relax_dev_kernel<double4, 1>), grd, blk, 0, st, A, (const double4 *)grid);
    } else {
        if (L.rp.mode == 0) hipLaunchKernelGGL((relax_dev_kernel<float4, 0>), grd, blk, 0, st, A, (const float4 *)grid);
        else hipLaunchKernelGGL((relax_dev_kernel<float4, 1>), grd, blk, 0, st, A, (const float4 *)grid);
    }
    return 0;
}

}  // namespace pp
