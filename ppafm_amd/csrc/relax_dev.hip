// FIRE relaxation, device path: the same semantics as relax.hip (see the
// citations there) with a leaner instruction stream, because each pixel
// column is a serial chain of ~10^4 dependent force evaluations and the
// per-iteration latency sets the kernel time:
//   * 1/l and |v|/|f| from v_rsq_f64 plus one Newton step (one reciprocal
//     square root per FIRE move) instead of IEEE divide / square root
//     sequences;
//   * fused multiply-adds written out (contraction is off, so every kernel
//     that shares these helpers rounds identically: a batch reproduces the
//     single-scan results bit for bit; this path is not bit-matched to the
//     CPU reference);
//   * 32-bit grid indices, ProbeParticle or OpenCL semantics chosen at
//     compile time, per-lane tip base so the loop touches few uniforms.
// Lanes run independent state machines over (z step, iteration) and stage
// their output records in LDS (Stage, relax_dev.hpp); an LDS-staged
// lock-step variant (relax_lds_kernel) is kept as an option.

#include <cstdlib>

#include "pp_common.hpp"

#pragma clang fp contract(off)

#include "relax_dev.hpp"

namespace pp {

namespace {

// The state of one pixel column's relaxation, advanced one force
// evaluation per step: the stroke's z step s, the FIRE state and, with warm
// starts, the Newton phase.  A column is done when s reaches nz.  The plain
// and warm kernels, per tile or with lane refill, all run these functions,
// so a column's results do not depend on the lane, wave or launch shape.
template <typename T4>
struct Column {
    Stroke c;
    long obase;  // output offset of the column's item (0: pre-offset outputs)
    double tx, ty, tz;
    Fire S;
    int s, it, itr;
    bool sample;  // mode 0: next gather is the output sample
    CellCache<T4> cache;
    // warm starts: Newton iterate q, iterations taken and allowed, FIRE
    // iteration of the next Newton try; `opening` = the step's first Newton
    // run (its failure is a fallback); `check` = S.r is a Newton zero not yet
    // evaluated by FIRE
    double qx, qy, qz;
    int next_try, nk, kmax;
    double ux1, uy1, uz1;  // probe offset from the tip at the previous step (predictor)
    bool newton, opening, check;
    bool gate;  // step 0's first Newton evaluation is at the network's estimate: test its residual
};

// Step 0 of column (i, j) (nothing to do when !live).  pred: the column's
// step-0 estimate (3 floats) or nullptr.
template <typename T4, int MODE, bool WARM>
__device__ __forceinline__ void column_init(const DevArgs &A, Column<T4> &C, int i, int j, bool live, long obase,
                                            const float *pred)
{
    C.c = stroke_at(A, i, j);
    C.obase = obase;
    C.tx = C.ty = C.tz = 0;
    C.s = live ? 0 : A.nz;
    if (live) tip_pos(A, C.c, 0, live, C.tx, C.ty, C.tz);
    double ox, oy, oz;  // equilibrium offset in the world frame
    probe_offset<MODE>(A, ox, oy, oz);
    C.S = Fire{C.tx + ox, C.ty + oy, C.tz + oz, 0, 0, 0, A.dtmax, A.a0};
    C.it = 0;
    C.itr = 0;
    C.sample = false;
    C.cache.ix = C.cache.iy = C.cache.iz = -1;
    if (WARM) {
        C.next_try = A.newton_retry;
        C.qx = C.S.rx; C.qy = C.S.ry; C.qz = C.S.rz;
        bool est = false;
        if (pred && live && isfinite(pred[0]) && isfinite(pred[1]) && isfinite(pred[2])) {  // else the plain start
            C.qx = (double)pred[0]; C.qy = (double)pred[1]; C.qz = (double)pred[2];
            est = true;
        }
        C.nk = 0;
        C.ux1 = C.uy1 = C.uz1 = 0;
        C.kmax = A.newton_max0;
        C.newton = live && C.kmax > 0;
        C.opening = C.newton;
        C.check = false;
        C.gate = C.newton && est && A.ws_res2 > 0;
        if (live && !C.newton && pred) { C.S.rx = C.qx; C.S.ry = C.qy; C.S.rz = C.qz; }  // the estimate as it is
    }
}

// One force evaluation of a plain column: FIRE to convergence at each z
// step, outputs in stroke order.  Counts force evaluations and finished
// positions.
// BUF: grid loads through a buffer resource (orthogonal kernels whose
// lanes share one grid); the lanes of a refill wave may each read another
// grid, and a buffer resource per lane would be a loop over the lanes'
// distinct grids at every load, so they use plain global loads.
template <typename T4, int MODE, int K, bool ORTHO = false, bool BUF = ORTHO>
__device__ __forceinline__ void column_step(const DevArgs &A, const T4 *__restrict__ grid, Column<T4> &C, Stage<K> &G,
                                            unsigned long long &nevals, unsigned long long &npos)
{
    double fe[4], gx, gy, gz;
    to_grid<ORTHO>(A, C.S.rx, C.S.ry, C.S.rz, gx, gy, gz);
    // the orthogonal kernels also read the grid through a buffer resource
    // (the host picks them for grids under 4 GiB only)
    gather_cached<T4, BUF, MODE == 1>(grid, A, gx, gy, gz, fe, C.cache);
    nevals++;
    bool done = false;
    const long o = C.obase + C.c.col0 + C.s;
    if (MODE == 0 && C.sample) {
        put_record<K>(A, G, o, C.s, C.S.rx, C.S.ry, C.S.rz, make_float4((float)fe[0], (float)fe[1], (float)fe[2], 0.f),
                      C.itr, A.status ? status_byte(C.itr >= A.max_iters, C.S.rx, C.S.ry, C.S.rz, fe) : 0);
        done = true;
    } else {
        const double ff = fire_move<MODE>(A, fe, C.tx, C.ty, C.tz, C.S);
        C.it++;
        const bool conv = ff < A.conv_f2;
        if (conv || C.it >= A.max_iters) {
            if (MODE == 0) {
                C.itr = conv ? C.it - 1 : A.max_iters;
                C.sample = true;
            } else {
                put_record<K>(A, G, o, C.s, C.S.rx, C.S.ry, C.S.rz, tip_frame_force(A, fe), C.it,
                              A.status ? status_byte(!conv, C.S.rx, C.S.ry, C.S.rz, fe) : 0);
                done = true;
            }
        }
    }
    if (done) {
        npos++;
        C.s++;
        if (C.s < A.nz) {
            next_step<MODE>(A, C.c, C.s, true, C.tx, C.ty, C.tz, C.S);
            C.it = 0;
            C.sample = false;
        }
    }
}

// One force evaluation of a warm-started column.  Every step begins with
// Newton iterations on the interpolated field (newton_from) from the
// continuation start the reference uses (step 0: from the column's estimate,
// or the plain start without one; with newton_extrap2 > 0, steps >= 2 from
// the continuation moved by the last change of the probe's offset from the
// tip, u(s-1) - u(s-2), when that change is at most sqrt(newton_extrap2)
// long), each step cut to at most newton_cap.
// When they reach a stable zero (a full step shorter than sqrt(newton_tol2)
// with a stable Jacobian at every iterate, at most newton_max of them,
// newton_max0 at step 0), FIRE runs from there; otherwise (an unstable
// Jacobian, a non-finite step, no convergence) FIRE runs from the
// continuation start itself, exactly as without warm starts (counted in
// nfb).  With ws_res2 > 0 the first evaluation at the network's step-0
// estimate also tests the estimate's residual: where |F|^2 there exceeds
// ws_res2 (or is not finite), step 0 is plain FIRE from the plain start
// (counted in nfb).  While FIRE runs, every newton_retry of its iterations one Newton
// step is tried from its current position with the corners FIRE just
// loaded; where the Jacobian is stable Newton iterations take over as above,
// and on convergence FIRE restarts there (velocity zero; mode 0 also resets
// dt and alpha), its iteration count running on toward max_iters.  When a
// Newton iterate (Jacobian stable) or FIRE's first evaluation at a Newton
// zero already meets the stop rule, the step ends at that point without a
// move (output force: the grid force there).  Either way FIRE ends the step by the reference's stop rule, and
// the next step continues from FIRE's result.  Newton iterates and FIRE
// moves share the cell cache of the corners (a load only when the cell
// changes); Newton evaluations count in nevals.
template <typename T4, int MODE, int K, bool ORTHO = false, bool BUF = ORTHO>
__device__ __forceinline__ void column_step_warm(const DevArgs &A, const T4 *__restrict__ grid, Column<T4> &C,
                                                 Stage<K> &G, unsigned long long &nevals, unsigned long long &npos,
                                                 unsigned long long &nfb)
{
    double gx, gy, gz, ux, uy, uz;
    to_grid<ORTHO>(A, C.newton ? C.qx : C.S.rx, C.newton ? C.qy : C.S.ry, C.newton ? C.qz : C.S.rz, gx, gy, gz);
    // the cell cache: Newton iterates near convergence and FIRE's moves mostly stay in one cell
    corners_cached<T4, BUF, MODE == 1>(grid, A, gx, gy, gz, C.cache, ux, uy, uz);
    const T4 *v = C.cache.v;
    nevals++;
    bool done = false;
    const long o = C.obase + C.c.col0 + C.s;
    const bool retry = !C.newton && !(MODE == 0 && C.sample) && A.newton_retry > 0 && C.it >= C.next_try;
    bool fire = !C.newton;
    double fe[4];
    interp8(v, ux, uy, uz, fe, MODE == 1);
    if (C.newton || retry) {
        double px = C.newton ? C.qx : C.S.rx, py = C.newton ? C.qy : C.S.ry, pz = C.newton ? C.qz : C.S.rz;
        bool stable;
        double st2, f2;
        const bool ok =
            newton_from<T4, MODE, ORTHO>(A, v, ux, uy, uz, fe, C.tx, C.ty, C.tz, px, py, pz, A.newton_cap, stable, st2, f2) &&
            stable;
        if (retry) C.next_try = C.it + A.newton_retry;
        const bool gated = C.gate && !(f2 <= A.ws_res2);
        C.gate = false;
        if (gated) {  // the estimate's residual |F| exceeds the threshold: plain FIRE from the plain start (C.S)
            nfb++;
            C.newton = false;
            C.opening = false;
        } else if (C.newton && stable && f2 < A.conv_f2) {  // the iterate meets the stop rule: the step ends here
            fire = false;
            C.S.rx = C.qx; C.S.ry = C.qy; C.S.rz = C.qz;
            put_record<K>(A, G, o, C.s, C.qx, C.qy, C.qz,
                          MODE == 0 ? make_float4((float)fe[0], (float)fe[1], (float)fe[2], 0.f) : tip_frame_force(A, fe),
                          C.it, A.status ? status_byte(false, C.qx, C.qy, C.qz, fe) : 0);
            C.newton = false;
            C.opening = false;
            done = true;
        } else if (ok || C.newton) {
            fire = false;
            C.nk = C.newton ? C.nk + 1 : 1;
            if (!C.newton) C.kmax = A.newton_max;
            C.qx = px; C.qy = py; C.qz = pz;
            if (ok && st2 < A.newton_tol2) {  // converged: FIRE from here
                C.S.rx = C.qx; C.S.ry = C.qy; C.S.rz = C.qz;
                C.S.vx = C.S.vy = C.S.vz = 0;
                if (MODE == 0) { C.S.dt = A.dtmax; C.S.ac = A.a0; }
                C.newton = false;
                C.opening = false;
                C.check = true;
            } else if (!ok || C.nk >= C.kmax) {  // FIRE from where it was
                if (C.opening) nfb++;
                C.newton = false;
                C.opening = false;
            } else {
                C.newton = true;
            }
        }
    }
    if (fire) {
        if (MODE == 0 && C.sample) {
            put_record<K>(A, G, o, C.s, C.S.rx, C.S.ry, C.S.rz, make_float4((float)fe[0], (float)fe[1], (float)fe[2], 0.f),
                          C.itr, A.status ? status_byte(C.itr >= A.max_iters, C.S.rx, C.S.ry, C.S.rz, fe) : 0);
            done = true;
        } else {
            double fx, fy, fz;
            total_force<MODE>(A, fe, C.tx, C.ty, C.tz, C.S.rx, C.S.ry, C.S.rz, fx, fy, fz);
            const double f2 = fma(fx, fx, fma(fy, fy, fz * fz));
            const bool settled = C.check && f2 < A.conv_f2;
            C.check = false;
            if (settled) {  // Newton's zero meets the stop rule: the step ends here, no move
                put_record<K>(A, G, o, C.s, C.S.rx, C.S.ry, C.S.rz,
                              MODE == 0 ? make_float4((float)fe[0], (float)fe[1], (float)fe[2], 0.f) : tip_frame_force(A, fe),
                              C.it, A.status ? status_byte(false, C.S.rx, C.S.ry, C.S.rz, fe) : 0);
                done = true;
            } else {
                const double ff = fire_update<MODE>(A, fx, fy, fz, C.S);
                C.it++;
                const bool conv = ff < A.conv_f2;
                if (conv || C.it >= A.max_iters) {
                    if (MODE == 0) {
                        C.itr = conv ? C.it - 1 : A.max_iters;
                        C.sample = true;
                    } else {
                        put_record<K>(A, G, o, C.s, C.S.rx, C.S.ry, C.S.rz, tip_frame_force(A, fe), C.it,
                                      A.status ? status_byte(!conv, C.S.rx, C.S.ry, C.S.rz, fe) : 0);
                        done = true;
                    }
                }
            }
        }
    }
    if (done) {
        npos++;
        C.s++;
        if (C.s < A.nz) {
            // the probe's offset from the tip at the step just ended
            const double ux = C.S.rx - C.tx, uy = C.S.ry - C.ty, uz = C.S.rz - C.tz;
            next_step<MODE>(A, C.c, C.s, true, C.tx, C.ty, C.tz, C.S);
            C.it = 0;
            C.next_try = A.newton_retry;
            C.sample = false;
            C.qx = C.S.rx; C.qy = C.S.ry; C.qz = C.S.rz;
            if (A.newton_extrap2 > 0 && C.s >= 2) {  // second-order predictor: the offset keeps its last change
                const double ex = ux - C.ux1, ey = uy - C.uy1, ez = uz - C.uz1;
                if (fma(ex, ex, fma(ey, ey, ez * ez)) <= A.newton_extrap2) {
                    C.qx += ex; C.qy += ey; C.qz += ez;
                }
            }
            C.ux1 = ux; C.uy1 = uy; C.uz1 = uz;
            C.nk = 0;
            C.kmax = A.newton_max;
            C.newton = C.opening = C.kmax > 0;
            C.check = false;
        }
    }
}

// One pixel column (i, j) from start to end.
template <typename T4, int MODE, int K, bool ORTHO = false>
__device__ __forceinline__ void relax_column(const DevArgs &A, const T4 *__restrict__ grid, int i, int j, bool live,
                                             Stage<K> &G, unsigned long long &nevals, unsigned long long &npos)
{
    Column<T4> C;
    column_init<T4, MODE, false>(A, C, i, j, live, 0, nullptr);
    while (C.s < A.nz) column_step<T4, MODE, K, ORTHO>(A, grid, C, G, nevals, npos);
}

template <typename T4, int MODE, int K, bool ORTHO = false>
__device__ __forceinline__ void relax_column_warm(const DevArgs &A, const T4 *__restrict__ grid, int i, int j,
                                                  bool live, Stage<K> &G, unsigned long long &nevals,
                                                  unsigned long long &npos, unsigned long long &nfb)
{
    Column<T4> C;
    column_init<T4, MODE, true>(A, C, i, j, live, 0, A.pred ? A.pred + 3 * ((long)i * A.ny + j) : nullptr);
    while (C.s < A.nz) column_step_warm<T4, MODE, K, ORTHO>(A, grid, C, G, nevals, npos, nfb);
}

// Wave sum of the per-lane counters, one atomic per wave.  `passes` (the
// same in every lane) is the wave's loop passes; 64 lane slots each go to
// counters[3] when with_slots.
__device__ __forceinline__ void flush_counters(unsigned long long *counters, unsigned long long nevals,
                                               unsigned long long npos, int lane, unsigned long long nfb = 0,
                                               bool with_fb = false, unsigned long long passes = 0,
                                               bool with_slots = false)
{
    if (!counters) return;
    for (int off = 32; off > 0; off >>= 1) {
        nevals += __shfl_xor(nevals, off, 64);
        npos += __shfl_xor(npos, off, 64);
        nfb += __shfl_xor(nfb, off, 64);
    }
    if (lane == 0) {
        atomicAdd(&counters[0], nevals);
        atomicAdd(&counters[1], npos);
        if (with_fb) atomicAdd(&counters[2], nfb);
        if (with_slots) atomicAdd(&counters[3], 64ull * passes);
    }
}

// records of 8 steps per lane: 23 KB of LDS per wave
constexpr int kStageSingle = 8;
// warm-started batches: 4 steps per lane, 11.5 KB per wave, so LDS does not
// limit their occupancy; plain-FIRE batches store directly (staged, they ran
// 3-7 % slower for the same output)
template <bool WARM>
constexpr int stage_batch()
{
    return WARM ? 4 : 1;
}

// WPB waves (tiles) per workgroup.  Block b runs WPB consecutive tiles of
// part b % 8 (one contiguous part of the tiles per XCD).
template <typename T4, int MODE, bool ORTHO, int WPB>
__global__ __launch_bounds__(64 * WPB) void relax_dev_kernel(DevArgs A, const T4 *__restrict__ grid)
{
    __shared__ __align__(16) unsigned char lds[kStageSingle * 64 * WPB * kStageRecordBytes];
    Stage<kStageSingle> G = make_stage<kStageSingle>(lds, 64 * WPB, threadIdx.x);
    const int per_xcd = ((A.ntiles + WPB - 1) / WPB + 7) / 8;
    const int b = blockIdx.x;
    int tile, lane;
    if constexpr (WPB == 1) {
        tile = (b % 8) * per_xcd + b / 8;
        lane = threadIdx.x;
    } else {
        tile = ((b % 8) * per_xcd + b / 8) * WPB + (int)(threadIdx.x / 64);
        lane = threadIdx.x % 64;
    }
    if (tile >= A.ntiles) return;
    const int ti = tile / A.tiles_y, tj = tile % A.tiles_y;
    const int i = ti * 8 + lane / 8, j = tj * 8 + lane % 8;
    unsigned long long nevals = 0, npos = 0;
    relax_column<T4, MODE, kStageSingle, ORTHO>(A, grid, i, j, i < A.nx && j < A.ny, G, nevals, npos);
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
    const float *pred;  // step-0 estimates (pixel columns, 3), WARM kernels only
    unsigned char *status;
    // relax_refill_kernel: 0, lanes take columns one by one from the XCD's
    // queue; > 0, a wave takes this many consecutive columns at once into a
    // pool of its own and its lanes take the pool's columns in order
    int chunk;
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

template <typename T4, int MODE, int TY, bool WARM, bool ORTHO>
__global__ __launch_bounds__(256) void relax_batch_kernel(DevArgs A0, BatchArgs B, const BatchRelaxItem *__restrict__ items,
                                                          const long long *__restrict__ prefix,
                                                          const T4 *__restrict__ grids)
{
    constexpr int TX = 64 / TY;
    constexpr int KS = stage_batch<WARM>();
    extern __shared__ __align__(16) unsigned char dyn_lds[];  // blockDim.x * KS records (none when KS = 1)
    Stage<KS> G = make_stage<KS>(dyn_lds, blockDim.x, threadIdx.x);
    const int lane = threadIdx.x & 63;
    const int xcd = blockIdx.x & 7;
    const long long per = (B.ntiles + 7) / 8;
    unsigned long long nevals = 0, npos = 0, nfb = 0, passes = 0;
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
            A.status = B.status ? B.status + it.out_off : nullptr;
            if (WARM) A.pred = B.pred ? B.pred + 3 * (long long)it.col_off : nullptr;
            const int lt = (int)(tile - prefix[m]);
            const int ti = lt / it.tiles_y, tj = lt % it.tiles_y;
            const int i = ti * TX + lane / TY, j = tj * TY + lane % TY;
            const unsigned long long before = nevals;
            if (WARM)
                relax_column_warm<T4, MODE, KS, ORTHO>(A, grids + it.grid_off, i, j, i < A.nx && j < A.ny, G, nevals, npos,
                                                       nfb);
            else relax_column<T4, MODE, KS, ORTHO>(A, grids + it.grid_off, i, j, i < A.nx && j < A.ny, G, nevals, npos);
            // the tile's loop passes: every pass evaluates once in each lane still running
            unsigned long long d = nevals - before;
            for (int off = 32; off > 0; off >>= 1) d = max(d, (unsigned long long)__shfl_xor(d, off, 64));
            passes += d;
        }
    }
    flush_counters(A0.counters, nevals, npos, lane, nfb, WARM, passes, true);
}

// ---------------------------------------------------------------------------
// Batches whose items share one geometry (grid shape, cell, scan; only the
// grid, output and estimate offsets differ), with lane refill: a lane takes
// the next pixel column as soon as its own is done, so a wave never waits
// for its slowest column (a warm-started column that falls back to FIRE
// costs ~20x an average one).  Columns are numbered item by item in tile
// order (64 per tile, TX x TY pixels), so the 64 columns a wave starts with
// form one pixel tile and later ones stay close to it.  The column range is
// split into 8 contiguous parts, one per XCD (blockIdx % 8), each with its
// own counter queue[part] (zero at launch); the lanes of a wave that need
// work take consecutive columns with one atomic per wave, and the wave moves
// on to the next part when its part is empty.  Every lane leaves the loop
// once all parts are empty and its column is done, so the grid drains.
template <typename T4, int MODE, int TY, bool WARM, bool ORTHO>
__global__ __launch_bounds__(256) void relax_refill_kernel(DevArgs A, BatchArgs B, const BatchRelaxItem *__restrict__ items,
                                                           const T4 *__restrict__ grids, long long cols_item)
{
    constexpr int TX = 64 / TY;
    constexpr int KS = stage_batch<WARM>();
    extern __shared__ __align__(16) unsigned char dyn_lds[];  // blockDim.x * KS records (none when KS = 1)
    Stage<KS> G = make_stage<KS>(dyn_lds, blockDim.x, threadIdx.x);
    const int lane = threadIdx.x & 63;
    const int xcd = blockIdx.x & 7;
    const long long ncols = (long long)B.nitems * cols_item;
    const long long per = (ncols + 7) / 8;
    unsigned long long nevals = 0, npos = 0, nfb = 0, passes = 0;
    Column<T4> C;
    C.s = A.nz;  // no column yet
    const T4 *grid = grids;
    int q = 0;         // parts of the column range this wave found empty (uniform)
    bool more = true;  // this lane may still find a column
    long long pool_lo = 0, pool_hi = 0;  // the wave's pool (B.chunk > 0): columns not yet taken (uniform)
    while (true) {
        bool need = more && C.s >= A.nz;
        unsigned long long mask = __ballot(need);
        while (mask != 0 && q < 8) {
            const int part = (xcd + q) & 7;
            const long long lo = part * per;
            const long long hi = min(ncols, lo + per);
            long long col0;  // this pass hands out columns col0, col0 + 1, ... by lane rank
            long long colend = hi;
            if (B.chunk > 0) {
                if (pool_lo >= pool_hi) {
                    unsigned base = 0;
                    if (lane == 0) base = atomicAdd(&B.queue[part], (unsigned)B.chunk);
                    base = __shfl(base, 0, 64);
                    if (lo + (long long)base >= hi) {  // this part is empty
                        q++;
                        continue;
                    }
                    pool_lo = lo + (long long)base;
                    pool_hi = min(hi, pool_lo + B.chunk);
                }
                col0 = pool_lo;
                colend = pool_hi;
                pool_lo = min(pool_hi, pool_lo + (long long)__popcll(mask));
            } else {
                const int leader = __ffsll((long long)mask) - 1;
                unsigned base = 0;
                if (lane == leader) base = atomicAdd(&B.queue[part], (unsigned)__popcll(mask));
                base = __shfl(base, leader, 64);
                col0 = lo + (long long)base;
            }
            if (need) {
                const long long col = col0 + __popcll(mask & ((1ull << lane) - 1));
                if (col < colend) {
                    const long long m = col / cols_item;
                    const long long lc = col - m * cols_item;
                    const int t = (int)(lc >> 6), l = (int)(lc & 63);
                    const int ti = t / A.tiles_y, tj = t - ti * A.tiles_y;
                    const int i = ti * TX + l / TY, j = tj * TY + l % TY;
                    const bool live = i < A.nx && j < A.ny;
                    const BatchRelaxItem &it = items[m];
                    grid = grids + it.grid_off;
                    const float *pred =
                        WARM && B.pred ? B.pred + 3 * ((long long)it.col_off + (long long)i * A.ny + j) : nullptr;
                    column_init<T4, MODE, WARM>(A, C, i, j, live, it.out_off, pred);
                    need = false;  // a column outside the scan (live false) is done at once
                }
            }
            mask = __ballot(need);
            if (mask != 0 && B.chunk == 0) q++;  // this part is empty
        }
        if (need) more = false;  // every part is empty
        if (__ballot(C.s < A.nz) == 0) {
            if (__ballot(more) == 0) break;
            continue;
        }
        if (C.s < A.nz) {
            if (WARM) column_step_warm<T4, MODE, KS, ORTHO, false>(A, grid, C, G, nevals, npos, nfb);
            else column_step<T4, MODE, KS, ORTHO, false>(A, grid, C, G, nevals, npos);
        }
        passes++;
    }
    flush_counters(A.counters, nevals, npos, lane, nfb, WARM, passes, true);
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
        bool capped = false;
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
                        capped = !conv;
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
            const double fo[4] = {outf.x, outf.y, outf.z, 0.0};
            store_status(A, o, capped, S.rx, S.ry, S.rz, fo);
            npos++;
        }
        __syncthreads();
    }
    flush_counters(A.counters, nevals, npos, lane);
}

}  // namespace

// 0 on launch, 1 when the parameters need the generic kernel (leap-frog,
// max_iters < 1), negative on error.
// SIMDs of the current device (4 per CU), 0 if unknown
static int simd_count()
{
    static int cache[64];
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return 0;
    if (dev < 64 && cache[dev] > 0) return cache[dev];
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (dev < 64) cache[dev] = 4 * cus;
    return 4 * cus;
}

// Single-scan launch.  With one-wave workgroups the dispatcher has been
// seen to place two waves of this launch on one SIMD while another SIMD of
// the CU idles, depending on what ran before it (c1, 1024 tiles, right after
// its grid build: 8.9-9.3 ms, against 7.1-7.4 ms after another FIRE launch).
// A scan with more than 3 tiles per CU and at most one per SIMD (c1)
// therefore uses four-wave workgroups, one per CU with its waves on the
// CU's four SIMDs.  Smaller scans keep one-wave workgroups spread over the
// CUs (four-wave workgroups there leave CUs idle: a 256-tile scan 3.80 ->
// 4.07 ms; two-wave workgroups on 512 tiles ran 1 % slower), and so do
// scans beyond one tile per SIMD, which LDS lets 6 to a CU.
template <typename T4, int MODE, bool ORTHO, int WPB>
static void launch_single_w(const DevArgs &A, const void *grid, hipStream_t st)
{
    const long nb = (((A.ntiles + WPB - 1L) / WPB + 7) / 8) * 8;
    hipLaunchKernelGGL((relax_dev_kernel<T4, MODE, ORTHO, WPB>), dim3((unsigned)nb), dim3(64 * WPB), 0, st, A,
                       (const T4 *)grid);
}

template <typename T4, int MODE, bool ORTHO>
static void launch_single(const DevArgs &A, const void *grid, hipStream_t st)
{
    const int simds = simd_count();
    const int cus = simds / 4;
    if (cus > 0 && A.ntiles > 3L * cus && A.ntiles <= simds) {
        launch_single_w<T4, MODE, ORTHO, 4>(A, grid, st);
    } else {
        launch_single_w<T4, MODE, ORTHO, 1>(A, grid, st);
    }
}

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
    A.status = L.out.status;
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
    // orthogonal cell axes (every configuration here): the diagonal of dinv
    // alone, and grid reads through a buffer resource (32-bit offsets: the
    // grid under 4 GiB)
    const unsigned long long gbytes = (unsigned long long)A.n0 * A.n1 * A.n2 * (f64 ? 32 : 16);
    const bool ortho = A.dinv[1] == 0 && A.dinv[2] == 0 && A.dinv[3] == 0 && A.dinv[5] == 0 && A.dinv[6] == 0 &&
                       A.dinv[7] == 0 && gbytes <= 0xFFFFFFFFull;
    if (f64) {
        if (L.rp.mode == 0) {
            if (ortho) launch_single<double4, 0, true>(A, grid, st);
            else launch_single<double4, 0, false>(A, grid, st);
        } else {
            if (ortho) launch_single<double4, 1, true>(A, grid, st);
            else launch_single<double4, 1, false>(A, grid, st);
        }
    } else {
        if (L.rp.mode == 0) {
            if (ortho) launch_single<float4, 0, true>(A, grid, st);
            else launch_single<float4, 0, false>(A, grid, st);
        } else {
            if (ortho) launch_single<float4, 1, true>(A, grid, st);
            else launch_single<float4, 1, false>(A, grid, st);
        }
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

template <typename T4, int MODE, bool WARM, bool ORTHO>
static int launch_refill_ty(int ty, dim3 grd, dim3 blk, hipStream_t st, const DevArgs &A, const BatchArgs &B,
                            const BatchRelaxItem *items, const void *grids, long long cols_item)
{
    const T4 *g = (const T4 *)grids;
    const size_t lds = stage_batch<WARM>() > 1 ? (size_t)blk.x * stage_batch<WARM>() * kStageRecordBytes : 0;
    switch (ty) {
    case 1: hipLaunchKernelGGL((relax_refill_kernel<T4, MODE, 1, WARM, ORTHO>), grd, blk, lds, st, A, B, items, g, cols_item); break;
    case 2: hipLaunchKernelGGL((relax_refill_kernel<T4, MODE, 2, WARM, ORTHO>), grd, blk, lds, st, A, B, items, g, cols_item); break;
    case 4: hipLaunchKernelGGL((relax_refill_kernel<T4, MODE, 4, WARM, ORTHO>), grd, blk, lds, st, A, B, items, g, cols_item); break;
    case 8: hipLaunchKernelGGL((relax_refill_kernel<T4, MODE, 8, WARM, ORTHO>), grd, blk, lds, st, A, B, items, g, cols_item); break;
    case 16: hipLaunchKernelGGL((relax_refill_kernel<T4, MODE, 16, WARM, ORTHO>), grd, blk, lds, st, A, B, items, g, cols_item); break;
    default: set_error("tile_y must be 1, 2, 4, 8 or 16"); return -2;
    }
    return 0;
}

template <typename T4, int MODE, bool WARM, bool ORTHO>
static int launch_batch_ty(int ty, dim3 grd, dim3 blk, hipStream_t st, const DevArgs &A, const BatchArgs &B,
                           const BatchRelaxItem *items, const long long *prefix, const void *grids)
{
    const T4 *g = (const T4 *)grids;
    const size_t lds = stage_batch<WARM>() > 1 ? (size_t)blk.x * stage_batch<WARM>() * kStageRecordBytes : 0;
    switch (ty) {
    case 1: hipLaunchKernelGGL((relax_batch_kernel<T4, MODE, 1, WARM, ORTHO>), grd, blk, lds, st, A, B, items, prefix, g); break;
    case 2: hipLaunchKernelGGL((relax_batch_kernel<T4, MODE, 2, WARM, ORTHO>), grd, blk, lds, st, A, B, items, prefix, g); break;
    case 4: hipLaunchKernelGGL((relax_batch_kernel<T4, MODE, 4, WARM, ORTHO>), grd, blk, lds, st, A, B, items, prefix, g); break;
    case 8: hipLaunchKernelGGL((relax_batch_kernel<T4, MODE, 8, WARM, ORTHO>), grd, blk, lds, st, A, B, items, prefix, g); break;
    case 16: hipLaunchKernelGGL((relax_batch_kernel<T4, MODE, 16, WARM, ORTHO>), grd, blk, lds, st, A, B, items, prefix, g); break;
    default: set_error("tile_y must be 1, 2, 4, 8 or 16"); return -2;
    }
    return 0;
}

template <typename T4, int MODE>
static int launch_batch_w(const RelaxBatchLaunch &L, dim3 grd, dim3 blk, hipStream_t st, const DevArgs &A, const BatchArgs &B)
{
    if (L.refill) {
        // one geometry: the uniforms of item 0, columns counted in tiles
        DevArgs G = A;
        const BatchRelaxItem &it = L.item0;
        G.n0 = it.n[0]; G.n1 = it.n[1]; G.n2 = it.n[2];
        G.nx = it.nx; G.ny = it.ny; G.nz = it.nz;
        int tiles_y = 1;
        const long long cols_item = 64LL * relax_batch_tiles(it.nx, it.ny, L.tile_y, tiles_y);
        G.tiles_y = tiles_y;
        for (int k = 0; k < 9; k++) G.dinv[k] = it.dinv[k];
        for (int k = 0; k < 3; k++) {
            G.p0[k] = it.p0[k]; G.da[k] = it.da[k]; G.db[k] = it.db[k]; G.dstep[k] = it.dstep[k];
        }
        G.tips = nullptr;
        G.pos = B.pos;
        G.pos_f32 = 1;
        G.force = B.force;
        G.iters = B.iters;
        G.status = B.status;
        G.pred = nullptr;
        PP_REQUIRE((long long)L.nitems * cols_item < (1LL << 35), "batch too large for the column queues");
        if (L.ortho) {
            if (L.warm) return launch_refill_ty<T4, MODE, true, true>(L.tile_y, grd, blk, st, G, B, L.items, L.grids, cols_item);
            return launch_refill_ty<T4, MODE, false, true>(L.tile_y, grd, blk, st, G, B, L.items, L.grids, cols_item);
        }
        if (L.warm) return launch_refill_ty<T4, MODE, true, false>(L.tile_y, grd, blk, st, G, B, L.items, L.grids, cols_item);
        return launch_refill_ty<T4, MODE, false, false>(L.tile_y, grd, blk, st, G, B, L.items, L.grids, cols_item);
    }
    // orthogonal cell axes in every item (every configuration here): the diagonal of dinv alone
    if (L.ortho) {
        if (L.warm) return launch_batch_ty<T4, MODE, true, true>(L.tile_y, grd, blk, st, A, B, L.items, L.prefix, L.grids);
        return launch_batch_ty<T4, MODE, false, true>(L.tile_y, grd, blk, st, A, B, L.items, L.prefix, L.grids);
    }
    if (L.warm) return launch_batch_ty<T4, MODE, true, false>(L.tile_y, grd, blk, st, A, B, L.items, L.prefix, L.grids);
    return launch_batch_ty<T4, MODE, false, false>(L.tile_y, grd, blk, st, A, B, L.items, L.prefix, L.grids);
}

// ---------------------------------------------------------------------------
// Independent positions: FIRE from given starts (a state reset per position),
// used to polish relaxed positions to a tight threshold.  Position g of the
// batch belongs to the item with pos_prefix[m] <= g < pos_prefix[m + 1]; its
// tip follows the item's scan.  One thread per position.
template <typename T4, int MODE>
__global__ __launch_bounds__(256) void fire_points_kernel(DevArgs A0, const BatchRelaxItem *__restrict__ items,
                                                          const long long *__restrict__ pos_prefix, int nitems,
                                                          long long npos, const T4 *__restrict__ grids,
                                                          const float *__restrict__ start, float *__restrict__ out,
                                                          int *__restrict__ iters)
{
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= npos) return;
    const int m = find_batch_item(pos_prefix, nitems, g);
    const BatchRelaxItem &it = items[m];
    DevArgs A = A0;
    A.n0 = it.n[0]; A.n1 = it.n[1]; A.n2 = it.n[2];
    A.nx = it.nx; A.ny = it.ny; A.nz = it.nz;
    for (int k = 0; k < 9; k++) A.dinv[k] = it.dinv[k];
    const long long o = g - pos_prefix[m];
    const int s = (int)(o % it.nz);
    const long long col = o / it.nz;
    const int i = (int)(col / it.ny), j = (int)(col % it.ny);
    const double tx = it.p0[0] + i * it.da[0] + j * it.db[0] + s * it.dstep[0];
    const double ty = it.p0[1] + i * it.da[1] + j * it.db[1] + s * it.dstep[1];
    const double tz = it.p0[2] + i * it.da[2] + j * it.db[2] + s * it.dstep[2];
    const T4 *grid = grids + it.grid_off;
    Fire S{(double)start[3 * g], (double)start[3 * g + 1], (double)start[3 * g + 2], 0, 0, 0, A.dtmax, A.a0};
    int k = 0;
    for (; k < A.max_iters; k++) {
        double fe[4], gx, gy, gz;
        to_grid(A, S.rx, S.ry, S.rz, gx, gy, gz);
        gather(grid, A, gx, gy, gz, fe, false);
        const double ff = fire_move<MODE>(A, fe, tx, ty, tz, S);
        if (ff < A.conv_f2) break;
    }
    out[3 * g] = (float)S.rx;
    out[3 * g + 1] = (float)S.ry;
    out[3 * g + 2] = (float)S.rz;
    if (iters) iters[g] = k;
}

int launch_fire_points(const RelaxBatchLaunch &L, const long long *pos_prefix, long long npos, const float *start,
                       float *out, int *iters, hipStream_t st)
{
    PP_REQUIRE(L.rp.mode == 0 || L.rp.mode == 1, "relax mode must be 0 or 1");
    PP_REQUIRE(L.rp.max_iters >= 1, "max_iters must be >= 1");
    DevArgs A = {};
    fill_shared_args(A, L.rp, L.tip);
    if (npos <= 0) return 0;
    const long long nblk = (npos + 255) / 256;
    PP_REQUIRE(nblk < (1LL << 31), "too many positions");
    dim3 grd((unsigned)nblk), blk(256);
    if (L.f64) {
        if (L.rp.mode == 0) hipLaunchKernelGGL((fire_points_kernel<double4, 0>), grd, blk, 0, st, A, L.items, pos_prefix, L.nitems, npos, (const double4 *)L.grids, start, out, iters);
        else hipLaunchKernelGGL((fire_points_kernel<double4, 1>), grd, blk, 0, st, A, L.items, pos_prefix, L.nitems, npos, (const double4 *)L.grids, start, out, iters);
    } else {
        if (L.rp.mode == 0) hipLaunchKernelGGL((fire_points_kernel<float4, 0>), grd, blk, 0, st, A, L.items, pos_prefix, L.nitems, npos, (const float4 *)L.grids, start, out, iters);
        else hipLaunchKernelGGL((fire_points_kernel<float4, 1>), grd, blk, 0, st, A, L.items, pos_prefix, L.nitems, npos, (const float4 *)L.grids, start, out, iters);
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
    const bool warm = L.warm;
    if (warm) {
        PP_REQUIRE(L.newton_max >= 0 && L.newton_max <= 64 && L.newton_max0 >= 0 && L.newton_max0 <= 64,
                   "Newton iteration limits must be 0..64");
        PP_REQUIRE(L.newton_tol > 0 && L.newton_cap > 0, "Newton tolerance and step cap must be positive");
        PP_REQUIRE(L.newton_retry >= 0, "newton_retry must be >= 0");
        A.newton_max = L.newton_max;
        A.newton_max0 = L.newton_max0;
        A.newton_tol2 = L.newton_tol * L.newton_tol;
        A.newton_cap = L.newton_cap;
        A.newton_retry = L.newton_retry;
        A.newton_extrap2 = L.newton_extrap * L.newton_extrap;
        A.ws_res2 = L.residual_max * L.residual_max;
    }
    BatchArgs B{L.nitems, L.ntiles, L.queue, L.pos, L.force, L.iters, warm ? L.pred : nullptr, L.status, L.refill_chunk};
    if (L.ntiles == 0) return 0;
    dim3 grd((unsigned)L.nblocks), blk(64 * L.waves_per_block);
    if (L.f64) {
        if (L.rp.mode == 0) return launch_batch_w<double4, 0>(L, grd, blk, st, A, B);
        return launch_batch_w<double4, 1>(L, grd, blk, st, A, B);
    }
    if (L.rp.mode == 0) return launch_batch_w<float4, 0>(L, grd, blk, st, A, B);
    return launch_batch_w<float4, 1>(L, grd, blk, st, A, B);
}

}  // namespace pp
