// FIRE relaxation, device path: the same semantics as relax.hip (see the
// citations there) with a leaner instruction stream, because each pixel
// column is a serial chain of ~10^4 dependent force evaluations and the
// per-iteration latency sets the kernel time:
//   * 1/l and sqrt(vv/ff) from v_rsq_f64 plus two Newton steps instead of
//     IEEE divide / square root sequences;
//   * fused multiply-adds written out (contraction is off, so every kernel
//     that shares these helpers rounds identically: a batch reproduces the
//     single-scan results bit for bit; this path is not bit-matched to the
//     CPU reference);
//   * 32-bit grid indices, ProbeParticle or OpenCL semantics chosen at
//     compile time, per-lane tip base so the loop touches few uniforms.
// Lanes run independent state machines over (z step, iteration); an
// LDS-staged lock-step variant (relax_lds_kernel) is kept as an option.

#include <cstdlib>

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

// grid index coordinates of a position (rows of dinv)
__device__ __forceinline__ void to_grid(const DevArgs &A, double rx, double ry, double rz, double &gx, double &gy,
                                        double &gz)
{
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

template <typename T4>
__device__ __forceinline__ void gather(const T4 *__restrict__ g, const DevArgs &A, double gx, double gy, double gz,
                                       double out[4], bool want_e)
{
    int ix, jx, iy, jy, iz, jz;
    double tx, ty, tz;
    axis(gx, A.n0, ix, jx, tx);
    axis(gy, A.n1, iy, jy, ty);
    axis(gz, A.n2, iz, jz, tz);
    const unsigned c00 = ((unsigned)ix * A.n1 + iy) * A.n2, c01 = ((unsigned)ix * A.n1 + jy) * A.n2;
    const unsigned c10 = ((unsigned)jx * A.n1 + iy) * A.n2, c11 = ((unsigned)jx * A.n1 + jy) * A.n2;
    const T4 v[8] = {g[c00 + iz], g[c00 + jz], g[c01 + iz], g[c01 + jz],
                     g[c10 + iz], g[c10 + jz], g[c11 + iz], g[c11 + jz]};
    interp8(v, tx, ty, tz, out, want_e);
}

// FIRE state of one probe
struct Fire {
    double rx, ry, rz, vx, vy, vz, dt, ac;
};

// Total force (grid force fe + tip springs, ProbeParticle.cpp:282-294 /
// relax.cl:37-41 in the tip frame) and one FIRE move (ProbeParticle.cpp:
// 142-174 / relax.cl:105-126).  Returns |F|^2 of the force it moved with.
template <int MODE>
__device__ __forceinline__ double fire_move(const DevArgs &A, const double fe[4], double tx, double ty, double tz,
                                            Fire &S)
{
    double fx = fe[0], fy = fe[1], fz = fe[2];
    const double dx = S.rx - tx, dy = S.ry - ty, dz = S.rz - tz;
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
    const double ff = fma(fx, fx, fma(fy, fy, fz * fz));
    const double vv = fma(S.vx, S.vx, fma(S.vy, S.vy, S.vz * S.vz));
    const double vf = fma(S.vx, fx, fma(S.vy, fy, S.vz * fz));
    if (vf < 0) {
        S.vx = S.vy = S.vz = 0;
        S.dt = MODE == 0 ? S.dt * A.fdec : fmax(A.dtmin, S.dt * A.fdec);
        S.ac = A.a0;
    } else {
        const double cf = vv > 0 ? S.ac * (vv * rsq_nr(vv)) * rsq_nr(MODE == 0 ? ff : ff + A.f2safe) : 0.0;
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

// One pixel column (i, j): every z step of the stroke, FIRE to convergence,
// outputs in stroke order.  Counts force evaluations and finished positions.
template <typename T4, int MODE>
__device__ __forceinline__ void relax_column(const DevArgs &A, const T4 *__restrict__ grid, int i, int j, bool live,
                                             unsigned long long &nevals, unsigned long long &npos)
{
    const int nz = A.nz;
    const Stroke c = stroke_at(A, i, j);
    double tx = 0, ty = 0, tz = 0;
    int s = live ? 0 : nz;
    if (live) tip_pos(A, c, 0, live, tx, ty, tz);
    double ox, oy, oz;  // equilibrium offset in the world frame
    probe_offset<MODE>(A, ox, oy, oz);
    Fire S{tx + ox, ty + oy, tz + oz, 0, 0, 0, A.dtmax, A.a0};
    int it = 0, itr = 0;
    bool sample = false;  // mode 0: next gather is the output sample

    while (s < nz) {
        double fe[4], gx, gy, gz;
        to_grid(A, S.rx, S.ry, S.rz, gx, gy, gz);
        gather(grid, A, gx, gy, gz, fe, MODE == 1);
        nevals++;
        bool done = false;
        const long o = c.col0 + s;
        if (MODE == 0 && sample) {
            store_pos(A, o, S.rx, S.ry, S.rz);
            if (A.force) A.force[o] = make_float4((float)fe[0], (float)fe[1], (float)fe[2], 0.f);
            if (A.iters) A.iters[o] = itr;
            done = true;
        } else {
            const double ff = fire_move<MODE>(A, fe, tx, ty, tz, S);
            it++;
            const bool conv = ff < A.conv_f2;
            if (conv || it >= A.max_iters) {
                if (MODE == 0) {
                    itr = conv ? it - 1 : A.max_iters;
                    sample = true;
                } else {
                    if (A.force) A.force[o] = tip_frame_force(A, fe);
                    store_pos(A, o, S.rx, S.ry, S.rz);
                    if (A.iters) A.iters[o] = it;
                    done = true;
                }
            }
        }
        if (done) {
            npos++;
            s++;
            if (s < nz) {
                next_step<MODE>(A, c, s, live, tx, ty, tz, S);
                it = 0;
                sample = false;
            }
        }
    }
}

// Wave sum of the per-lane counters, one atomic per wave.
__device__ __forceinline__ void flush_counters(unsigned long long *counters, unsigned long long nevals,
                                               unsigned long long npos, int lane)
{
    if (!counters) return;
    for (int off = 32; off > 0; off >>= 1) {
        nevals += __shfl_xor(nevals, off, 64);
        npos += __shfl_xor(npos, off, 64);
    }
    if (lane == 0) {
        atomicAdd(&counters[0], nevals);
        atomicAdd(&counters[1], npos);
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
    unsigned long long nevals = 0, npos = 0;
    relax_column<T4, MODE>(A, grid, i, j, i < A.nx && j < A.ny, nevals, npos);
    flush_counters(A.counters, nevals, npos, lane);
}

// ---------------------------------------------------------------------------
// Batches: a persistent kernel.  Each wave takes pixel tiles (64 columns of
// one molecule, TY columns along y) from a work queue until none are left.
// The tile range is split into 8 contiguous parts, one per XCD (workgroups
// are dispatched to XCDs round robin, so blockIdx % 8 names the XCD): the
// molecules of one part share that XCD's L2.  A wave that finds its part
// empty moves on to the next part, which balances the tail.  queue[0..7]
// are the part heads and must be zero at launch.
struct BatchArgs {
    int nitems;
    long long ntiles;
    unsigned *queue;
    float *pos;  // (positions, 3) FP32
    float4 *force;
    int *iters;
};

__device__ __forceinline__ int find_batch_item(const long long *__restrict__ prefix, int nitems, long long x)
{
    int lo = 0, hi = nitems;  // prefix[lo] <= x < prefix[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <typename T4, int MODE, int TY>
__global__ __launch_bounds__(256) void relax_batch_kernel(DevArgs A0, BatchArgs B, const BatchRelaxItem *__restrict__ items,
                                                          const long long *__restrict__ prefix,
                                                          const T4 *__restrict__ grids)
{
    constexpr int TX = 64 / TY;
    const int lane = threadIdx.x & 63;
    const int xcd = blockIdx.x & 7;
    const long long per = (B.ntiles + 7) / 8;
    unsigned long long nevals = 0, npos = 0;
    for (int q = 0; q < 8; q++) {
        const int part = (xcd + q) & 7;
        const long long lo = part * per;
        const long long hi = min(B.ntiles, lo + per);
        if (lo >= hi) continue;
        while (true) {
            unsigned t = 0;
            if (lane == 0) t = atomicAdd(&B.queue[part], 1u);
            t = __builtin_amdgcn_readfirstlane(__shfl(t, 0, 64));
            const long long tile = lo + (long long)t;
            if (tile >= hi) break;
            const int m = find_batch_item(prefix, B.nitems, tile);
            const BatchRelaxItem &it = items[m];
            DevArgs A = A0;
            A.n0 = it.n[0]; A.n1 = it.n[1]; A.n2 = it.n[2];
            A.nx = it.nx; A.ny = it.ny; A.nz = it.nz;
            for (int k = 0; k < 9; k++) A.dinv[k] = it.dinv[k];
            for (int k = 0; k < 3; k++) {
                A.p0[k] = it.p0[k]; A.da[k] = it.da[k]; A.db[k] = it.db[k]; A.dstep[k] = it.dstep[k];
            }
            A.tips = nullptr;
            A.pos = B.pos ? (void *)(B.pos + 3 * it.out_off) : nullptr;
            A.pos_f32 = 1;
            A.force = B.force ? B.force + it.out_off : nullptr;
            A.iters = B.iters ? B.iters + it.out_off : nullptr;
            const int lt = (int)(tile - prefix[m]);
            const int ti = lt / it.tiles_y, tj = lt % it.tiles_y;
            const int i = ti * TX + lane / TY, j = tj * TY + lane % TY;
            relax_column<T4, MODE>(A, grids + it.grid_off, i, j, i < A.nx && j < A.ny, nevals, npos);
        }
    }
    flush_counters(A0.counters, nevals, npos, lane);
}

// ---------------------------------------------------------------------------
// LDS-staged variant.  The lanes of a wave relax z step s together; before
// the step the wave reduces its probes' bounding box (grid index space),
// copies that block of the grid (plus a margin) into LDS as structure-of-
// arrays planes, and every gather whose 8 corners fall inside the block
// reads LDS; the others read global memory.  The per-lane arithmetic is the
// shared helpers', so the results equal relax_dev_kernel's.
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
    const Stroke c = stroke_at(A, i, j);
    double tx, ty, tz;
    tip_pos(A, c, 0, live, tx, ty, tz);
    double ox, oy, oz;
    probe_offset<MODE>(A, ox, oy, oz);
    Fire S{tx + ox, ty + oy, tz + oz, 0, 0, 0, A.dtmax, A.a0};
    unsigned long long nevals = 0, npos = 0;

    for (int s = 0; s < nz; s++) {
        if (s > 0) next_step<MODE>(A, c, s, live, tx, ty, tz, S);
        // ---- stage the block around the wave's probes
        double g0x, g0y, g0z;
        to_grid(A, S.rx, S.ry, S.rz, g0x, g0y, g0z);
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
                const int cz = q % ez, t = q / ez;
                const int bb = t % ey, a = t / ey;
                const unsigned gi = ((unsigned)wrapi(lox + a, A.n0) * A.n1 + wrapi(loy + bb, A.n1)) * A.n2 + wrapi(loz + cz, A.n2);
                const T4 v = grid[gi];
                LX[q] = v.x; LY[q] = v.y; LZ[q] = v.z;
                if (MODE == 1) LW[q] = v.w;
            }
        }
        __syncthreads();

        // ---- relax this step, all lanes together
        int st = live ? 0 : 2;  // 0 = FIRE, 1 = output sample pending (mode 0), 2 = done
        int it = 0, itr = 0;
        float4 outf = make_float4(0.f, 0.f, 0.f, 0.f);
        while (__ballot(st < 2) != 0) {
            if (st < 2) {
                double gx, gy, gz;
                to_grid(A, S.rx, S.ry, S.rz, gx, gy, gz);
                const double flx = floor(gx), fly = floor(gy), flz = floor(gz);
                double fe[4];
                // inside test in floating point (false for NaN), before any int conversion
                if (staged && flx >= lox && flx + 1 < lox + ex && fly >= loy && fly + 1 < loy + ey && flz >= loz &&
                    flz + 1 < loz + ez) {
                    const int a = (int)flx - lox, bb = (int)fly - loy, cz = (int)flz - loz;
                    const int q00 = (a * ey + bb) * ez + cz, q01 = q00 + ez, q10 = q00 + ey * ez, q11 = q10 + ez;
                    const int q[8] = {q00, q00 + 1, q01, q01 + 1, q10, q10 + 1, q11, q11 + 1};
                    T4 v[8];
#pragma unroll
                    for (int k = 0; k < 8; k++) {
                        v[k].x = LX[q[k]]; v[k].y = LY[q[k]]; v[k].z = LZ[q[k]];
                        v[k].w = MODE == 1 ? LW[q[k]] : (T)0;
                    }
                    interp8(v, gx - flx, gy - fly, gz - flz, fe, MODE == 1);
                } else {
                    gather(grid, A, gx, gy, gz, fe, MODE == 1);
                }
                nevals++;
                if (MODE == 0 && st == 1) {
                    outf = make_float4((float)fe[0], (float)fe[1], (float)fe[2], 0.f);
                    st = 2;
                } else {
                    const double ff = fire_move<MODE>(A, fe, tx, ty, tz, S);
                    it++;
                    const bool conv = ff < A.conv_f2;
                    if (conv || it >= A.max_iters) {
                        if (MODE == 0) {
                            itr = conv ? it - 1 : A.max_iters;
                            st = 1;
                        } else {
                            outf = tip_frame_force(A, fe);
                            itr = it;
                            st = 2;
                        }
                    }
                }
            }
        }
        if (live) {
            const long o = c.col0 + s;
            store_pos(A, o, S.rx, S.ry, S.rz);
            if (A.force) A.force[o] = outf;
            if (A.iters) A.iters[o] = itr;
            npos++;
        }
        __syncthreads();
    }
    flush_counters(A.counters, nevals, npos, lane);
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

static void fill_shared_args(DevArgs &A, const ppafm_relax_params &rp, const ppafm_tip &tip)
{
    A.max_iters = rp.max_iters;
    A.probe_start = rp.probe_start;
    for (int k = 0; k < 9; k++) A.rot[k] = tip.rot[k];
    A.conv_f2 = rp.conv_f2; A.dtmax = rp.dt; A.a0 = rp.damping;
    A.finc = rp.finc; A.fdec = rp.fdec; A.falpha = rp.falpha;
    A.dtmin = rp.dtmin_frac * rp.dt; A.f2safe = rp.f2safe;
    for (int k = 0; k < 3; k++) { A.ks[k] = tip.kspring[k]; A.d0[k] = tip.d0[k]; }
    A.krad = tip.krad; A.lrad = tip.lrad; A.surf_ff = tip.surf_ff;
}

template <typename T4, int MODE>
static int launch_batch_ty(int ty, dim3 grd, dim3 blk, hipStream_t st, const DevArgs &A, const BatchArgs &B,
                           const BatchRelaxItem *items, const long long *prefix, const void *grids)
{
    switch (ty) {
    case 1: hipLaunchKernelGGL((relax_batch_kernel<T4, MODE, 1>), grd, blk, 0, st, A, B, items, prefix, (const T4 *)grids); break;
    case 2: hipLaunchKernelGGL((relax_batch_kernel<T4, MODE, 2>), grd, blk, 0, st, A, B, items, prefix, (const T4 *)grids); break;
    case 4: hipLaunchKernelGGL((relax_batch_kernel<T4, MODE, 4>), grd, blk, 0, st, A, B, items, prefix, (const T4 *)grids); break;
    case 8: hipLaunchKernelGGL((relax_batch_kernel<T4, MODE, 8>), grd, blk, 0, st, A, B, items, prefix, (const T4 *)grids); break;
    case 16: hipLaunchKernelGGL((relax_batch_kernel<T4, MODE, 16>), grd, blk, 0, st, A, B, items, prefix, (const T4 *)grids); break;
    default: set_error("tile_y must be 1, 2, 4, 8 or 16"); return -2;
    }
    return 0;
}

int relax_batch_tiles(int nx, int ny, int tile_y, int &tiles_y)
{
    const int tx = 64 / tile_y;
    tiles_y = (ny + tile_y - 1) / tile_y;
    return ((nx + tx - 1) / tx) * tiles_y;
}

int launch_relax_batch(const RelaxBatchLaunch &L, hipStream_t st)
{
    PP_REQUIRE(L.rp.alg == 1 && L.rp.max_iters >= 1, "batched relaxation needs FIRE (alg 1) and max_iters >= 1");
    PP_REQUIRE(L.rp.mode == 0 || L.rp.mode == 1, "relax mode must be 0 or 1");
    PP_REQUIRE(L.rp.probe_start >= -1 && L.rp.probe_start <= 1, "probeStart must be -1, 0 or 1");
    PP_REQUIRE(L.waves_per_block >= 1 && L.waves_per_block <= 4, "waves_per_block must be 1..4");
    PP_REQUIRE(L.nblocks >= 1, "nblocks must be >= 1");
    DevArgs A = {};
    fill_shared_args(A, L.rp, L.tip);
    A.counters = L.counters;
    BatchArgs B{L.nitems, L.ntiles, L.queue, L.pos, L.force, L.iters};
    if (L.ntiles == 0) return 0;
    dim3 grd((unsigned)L.nblocks), blk(64 * L.waves_per_block);
    if (L.f64) {
        if (L.rp.mode == 0) return launch_batch_ty<double4, 0>(L.tile_y, grd, blk, st, A, B, L.items, L.prefix, L.grids);
        return launch_batch_ty<double4, 1>(L.tile_y, grd, blk, st, A, B, L.items, L.prefix, L.grids);
    }
    if (L.rp.mode == 0) return launch_batch_ty<float4, 0>(L.tile_y, grd, blk, st, A, B, L.items, L.prefix, L.grids);
    return launch_batch_ty<float4, 1>(L.tile_y, grd, blk, st, A, B, L.items, L.prefix, L.grids);
}

}  // namespace pp
