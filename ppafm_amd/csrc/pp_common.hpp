// Internal helpers shared by the kernels and the C ABI of libppafm_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/ppafm_hip.h"

namespace pp {

// Last error message of the calling thread (ppafm_last_error).
void set_error(const std::string &msg);
const char *get_error();

#define PP_TRY_HIP(expr)                                                                         \
    do {                                                                                         \
        hipError_t _e = (expr);                                                                  \
        if (_e != hipSuccess) {                                                                  \
            ::pp::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));                \
            return -1;                                                                           \
        }                                                                                        \
    } while (0)

#define PP_REQUIRE(cond, msg)                                                                    \
    do {                                                                                         \
        if (!(cond)) {                                                                           \
            ::pp::set_error(msg);                                                                \
            return -2;                                                                           \
        }                                                                                        \
    } while (0)

// Grid geometry in the form the kernels use: step vectors (rows of cell/n)
// and the map from a Cartesian offset to fractional grid indices, computed
// on the host in FP64 like GridShape::setCell (Grid.h:29-36).
struct GridGeom {
    int n[3];
    double dcell[9];  // rows: a/na, b/nb, c/nc
    double dinv[9];   // (dcell^-1)^T: row r dotted with a position gives index coordinate r
};

GridGeom make_geom(const int n[3], const double cell[9]);

// Step vectors (rows of d) along the axes only: the FP64 grid kernels then
// share dx, dy between the z points of a lane.
inline bool ortho_cell(const double d[9])
{
    return d[1] == 0 && d[2] == 0 && d[3] == 0 && d[5] == 0 && d[6] == 0 && d[7] == 0;
}

// Atom record for the FP32 grid kernel: coordinates relative to the brick
// centre are formed in FP64 on the fly and split hi + lo, so that
// d = (p - hi) - lo keeps FP32 relative accuracy in the pair distance.
using AtomRec = ppafm_atom;  // x, y, z (grid frame), q * coulomb constant, c6, c12

hipStream_t as_stream(void *s);

// One molecule of a batched grid build: its grid and its slice of the packed
// atom table.  Grids of a batch lie back to back in one allocation.
struct BatchGridItem {
    int n[3];
    int natoms;
    long long atom_off;  // first atom in the packed table
    long long grid_off;  // first grid point (element index)
    double dcell[9];
};

// One molecule of a batched relaxation: the grid it samples and its scan.
struct BatchRelaxItem {
    int n[3];
    int nx, ny, nz;
    int tiles_y;
    int col_off;         // first pixel column in the batch
    long long grid_off;  // first grid point (element index)
    long long out_off;   // first output position (stroke order, nx*ny*nz block)
    double dinv[9];
    double p0[3], da[3], db[3], dstep[3];
    double dcell[9];  // grid step vectors (rows)
};

}  // namespace pp

// Launchers implemented in the .hip files.
namespace pp {

// Device grid (z fastest): float4 (f64 = false) or double4 (f64 = true).
int launch_ff_build(const GridGeom &g, const ppafm_ff_model &m, int natoms, const AtomRec *atoms_dev, void *out,
                    bool f64, hipStream_t st);
// Bricks of one grid (launch blocks of the grid kernels).
long long ff_bricks(const int n[3]);
// Every grid of a batch in one launch; prefix[k] = bricks of items < k.
int launch_ff_build_batch(const ppafm_ff_model &m, int nitems, const BatchGridItem *items_dev,
                          const long long *prefix_dev, long long nbricks, const AtomRec *atoms_dev, void *grids,
                          bool f64, bool qskip, bool ortho, hipStream_t st);
// Pair models of the reference C ABI (ff_models.hip).  Coefficients per
// atom: LJ (c6, c12); Coulomb (kqq); vdW (c6), parameter ADamp; vdW with a
// damping function (R, E0), parameter ADamp; D3 (c6, c8, R0^6, R0^8); Morse
// (R0, E0), parameter alpha; densities (amplitude, width); LJ in (R0, E0);
// bare r^-6 (c6); the OpenCL vdW variants as their CPU counterparts.  The
// three densities write energy only.
enum FFModelF64 {
    kModelLJ = 0,
    kModelCoulS,
    kModelCoulPz,
    kModelCoulDz2,
    kModelVdW,
    kModelVdWR2,
    kModelVdWR4,
    kModelVdWInvR4,
    kModelVdWInvR8,
    kModelD3,
    kModelMorse,
    kModelGauss,
    kModelSlater,
    kModelR4,
    kModelLJRE,
    kModelInvR6,
    kModelVdWNoDampOcl,
    kModelVdWInvR4Ocl,
    kModelVdWInvR8Ocl,
    kModelCount
};
// FP64 accumulate (+=) into the (npts,3) grid F and (npts) E, either may be
// NULL; pos (natoms,3) relative to the grid origin, coefs (natoms,ncoef).
int launch_ff_models(const GridGeom &g, int model, double par, int natoms, const double *pos, const double *coefs,
                     int ncoef, double *F, double *E, hipStream_t st);
// The same with F and E `fstride` and `estride` doubles apart per point
// (4 and 4 for the force and energy channels of a double4 grid).
int launch_ff_models_strided(const GridGeom &g, int model, double par, int natoms, const double *pos,
                             const double *coefs, int ncoef, double *F, int fstride, double *E, int estride,
                             hipStream_t st);
// LJ at npoints points pts (3 each) from atoms pos/cljs -> fes (npoints,4).
int launch_lj_points(long npoints, const double *pts, int natoms, const double *pos, const double *cljs, double *fes,
                     hipStream_t st);
// One pair model at separations (r,0,0): es = energy, fs = x force.
int launch_radial(long n, const double *rs, int model, double c0, double c1, double par, double *es, double *fs,
                  hipStream_t st);
// DFT-D3 coefficients (c6 s6, c8 s8, R0^6, R0^8) per atom.
int launch_d3_coeffs(int natoms, const double *pos, const int *elems, const double *r_cov, const double *ref_cn,
                     const double *ref_c6, const double *r4r2, const double k[3], const double params[4], int elem_pp,
                     double *out, hipStream_t st);
int launch_spline_uniform(double x0, double dx, int n, const double *ydys, long m, const double *xs, double *ys,
                          hipStream_t st);
int launch_spline_nonuniform(int n, const double *kx, const double *ydys, long m, const double *xs, double *ys,
                             hipStream_t st);
int launch_test_force(int type, int n, const double r0[3], const double dr[3], const double R[3], double krad,
                      double lrad, int sn, const double *sxs, const double *sydys, double *fs, hipStream_t st);

struct RelaxLaunch {
    GridGeom g;
    ppafm_relax_params rp;
    ppafm_tip tip;
    ppafm_scan scan;
    const double *tips;  // optional explicit tip positions (nx,ny,nz,3)
    ppafm_relax_out out;
    // mode-0 radial tip force from a spline (n knots, device xs and ydys)
    int spline_n = 0;
    const double *spline_xs = nullptr, *spline_ydys = nullptr;
};
int launch_relax_dev(const RelaxLaunch &L, const void *grid, bool f64, hipStream_t st);
int launch_relax_fast(const RelaxLaunch &L, const void *grid, bool f64, hipStream_t st);
int launch_relax_f64grid(const RelaxLaunch &L, const double *grid, double *pos, double *force, int *iters,
                         hipStream_t st);
// Dynamical matrix and its eigen-decomposition at L.scan's positions (pos,
// float32 or float64 (npos,3) in stroke order), mode-0 tip.  grid_kind:
// 0 float4, 1 double4, 2 double (npts,3) (the host ABI's grid).
int launch_stiffness(const RelaxLaunch &L, const void *grid, int grid_kind, const void *pos, int pos_f32,
                     double ddisp, int which, double *evals, double *ev1, double *ev2, double *ev3, hipStream_t st);
// test_eigen3x3: eigenvalues (reference order) and eigenvectors of one
// device 3x3 matrix -> out[12].
int launch_eigen3x3(const double *m, double *out, hipStream_t st);

// Batched FIRE relaxation (persistent kernel, see relax_dev.hip).
struct RelaxBatchLaunch {
    ppafm_relax_params rp;
    ppafm_tip tip;
    const BatchRelaxItem *items;  // device
    const long long *prefix;      // device, tiles of items < k
    int nitems;
    long long ntiles;
    const void *grids;  // float4 or double4, back to back
    bool f64;
    unsigned *queue;  // 8 part heads, zero at launch
    float *pos;
    float4 *force;
    int *iters;
    unsigned long long *counters;
    unsigned char *status = nullptr;  // per-position status, optional
    int tile_y;           // pixel tile: (64 / tile_y) x tile_y columns
    int waves_per_block;  // 1..4
    int nblocks;          // persistent workgroups
    // lane refill (every item of item0's geometry): lanes take pixel
    // columns one by one in tile order from per-XCD column queues, or, with
    // refill_chunk > 0, from a pool of that many consecutive columns that
    // their wave takes at once
    bool refill = false;
    int refill_chunk = 0;
    BatchRelaxItem item0{};
    // warm starts (relax_column_warm): on when `warm`; pred = step-0
    // estimates per pixel column (3 floats) or nullptr
    bool warm = false;
    const float *pred = nullptr;
    int newton_max = 0, newton_max0 = 0, newton_retry = 0;
    double newton_tol = 0, newton_cap = 0;
    double newton_extrap = 0;  // second-order predictor for steps >= 2 up to this step (A), 0 = off
    double residual_max = 0;   // eV/A, 0 = off (step-0 estimate residual gate)
    bool ortho = false;        // every item's dinv is diagonal (orthogonal cell axes) and its grid under 4 GiB
};
// Tiles of one scan for a given tile_y (also returns tiles along y).
int relax_batch_tiles(int nx, int ny, int tile_y, int &tiles_y);
int launch_relax_batch(const RelaxBatchLaunch &L, hipStream_t st);
// FIRE from the given starts, one independent position each (items, grids,
// rp and tip of L; pos_prefix[k] = positions of items < k).
int launch_fire_points(const RelaxBatchLaunch &L, const long long *pos_prefix, long long npos, const float *start,
                       float *out, int *iters, hipStream_t st);

// Learned warm start (warmstart.hip).
struct WsNetView {
    const float *theta;
    const unsigned short *W1b, *W2b;
};
WsNetView ws_view(const ppafm_ws_model *m);
// Step-0 estimates of every pixel column of a batch: pred (columns, 3),
// item k's columns from items[k].col_off; wave_prefix[k] = 8x8-column tiles
// of items < k (nwaves in all); reps network passes.
int launch_ws_predict(const WsNetView &net, const RelaxBatchLaunch &L, const long long *wave_prefix, long long nwaves,
                      float *pred, int reps, hipStream_t st);
// Training pairs at the selected positions of a relaxed batch: patch around
// the continuation start (from pos) moved by a pseudo-random offset of up to
// jitter A (half along z), target = truth - patch node (in 0.1 A).
int launch_ws_samples(const RelaxBatchLaunch &L, const long long *pos_prefix, const float *pos, const float *truth,
                      long long nsel, const long long *sel, unsigned short *patches, float *targets, float jitter,
                      unsigned seed, hipStream_t st);

int launch_limit_vec_field(double *F, long npts, double fmax, hipStream_t st);
int launch_copy(const void *src, void *dst, long n16, hipStream_t st);
int launch_fill_u32(void *p, long n, unsigned v, hipStream_t st);
int launch_convolve_z(long ncol, int nz, int in_stride, int in_step, int in_offset, const void *in, int nw,
                      const double *w, double scale, void *out, int f64, hipStream_t st);

}  // namespace pp
