/*
 * ppafm_hip.h -- C ABI of libppafm_hip.so, the MI355X (gfx950) implementation
 * of the probe-particle hot path: LJ + point-charge Coulomb force-field grid,
 * FIRE relaxation of the probe through the interpolated grid, and the z
 * convolution that turns the force stack into a frequency shift.
 *
 * Four groups of entry points.
 *
 * (1) Reference-compatible host ABI.  Same names, argument lists and
 *     semantics as the ctypes bindings of ppafm/core.py:33-324 over
 *     ppafm/cpp/ProbeParticle.cpp.  Arrays are caller-owned, C-contiguous
 *     float64 host arrays with the reference layouts (grids (nx,ny,nz,3) with
 *     z fastest).  Like the reference these keep process-wide state set by the
 *     set* calls and are not reentrant.  Work runs on the GPU in FP64.
 *
 * (2) Device ABI (ppafm_* prefix).  Plain device pointers and sizes, explicit
 *     descriptors, an explicit HIP stream (NULL = default stream), an int
 *     status (0 = success; ppafm_last_error() describes a failure).  The grid
 *     is float4 or double4 (Fx,Fy,Fz,E) with z fastest; relaxation arithmetic
 *     is FP64.  These have no counterpart in the reference; they are what the
 *     Python mirror of ppafm.ocl.AFMulator and the batched path call.
 *
 * (3) Grid files: the XSF writer and number reader of ppafm/io.py and
 *     GridUtils.cpp, native and multithreaded (host code).
 *
 * (4) Grids of the Hartree and full-density-based force fields (the
 *     ppafm/ocl/field.py operations), device ABI conventions as in (2).
 */
#ifndef PPAFM_HIP_H
#define PPAFM_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------
 * (1) Reference-compatible host ABI
 * ---------------------------------------------------------------------- */

/* Tip radial spline (TIP::SplineParams, ProbeParticle.cpp:58-62): rff_n
 * knots rff_xs and (rff_n,2) values and slopes rff_ydys of the radial force;
 * NULL selects the harmonic radial spring of setTip.  The relaxation and
 * stiffness calls evaluate it as forceRSpline (ProbeParticle.cpp:211-224),
 * extrapolating the last interval beyond the last knot. */
typedef struct {
    int rff_n;
    double *rff_xs;
    double *rff_ydys;
} ppafm_spline_params;

/* ProbeParticle.cpp:358  setGridN(int* n): grid points along a, b, c. */
void setGridN(int *n);
/* ProbeParticle.cpp:365  setGridCell(double* cell): 3x3 lattice, rows a,b,c. */
void setGridCell(double *cell);
/* ProbeParticle.cpp:338  setFF_Fpointer: (nx,ny,nz,3) float64 force grid. */
void setFF_Fpointer(double *gridF);
/* ProbeParticle.cpp:343  setFF_Epointer: (nx,ny,nz) float64 energy grid. */
void setFF_Epointer(double *gridE);
/* ProbeParticle.cpp:348 / 353 */
void deleteFF_Fpointer(void);
void deleteFF_Epointer(void);
/* ProbeParticle.cpp:371  setTip(lRad, kRad, rPP0[3], kSpring[3]); rPP0.z is
 * moved onto the sphere of radius lRad when |kRad| > 1e-8 (makeConsistent). */
void setTip(double lRad, double kRad, double *rPP0, double *kSpring);
/* ProbeParticle.cpp:322  setRelax(maxIters, convF2, dt, damping). */
void setRelax(int maxIters, double convF2, double dt, double damping);
/* ProbeParticle.cpp:331  setFIRE(finc, fdec, falpha). */
void setFIRE(double finc, double fdec, double falpha);
/* ProbeParticle.cpp:508  F (and E if set) += LJ grid of natoms atoms;
 * Rs (natoms,3) relative to the grid origin, cLJs (natoms,2) = (c6, c12). */
void getLennardJonesFF(int natoms, double *Rs, double *cLJs);
/* ProbeParticle.cpp:540  F (and E) += Coulomb grid for a unit probe charge;
 * kQQs (natoms) = Q_i * 14.399645; kind 0 = s, 1 = pz, 2 = dz2. */
void getCoulombFF(int natoms, double *Rs, double *kQQs, int kind);
/* ProbeParticle.cpp:633  relax one stroke of nstep tip positions; rs, fs
 * (nstep,3) out; returns the sum of FIRE iteration indices. */
int relaxTipStroke(int probeStart, int relaxAlg, int nstep, double *rTips, double *rs, double *fs,
                   ppafm_spline_params *splineParams);
/* ProbeParticle.cpp:682  relax nx*ny strokes; rTips, rs, fs (nx,ny,nz,3);
 * returns 0 like the reference (ppafm_last_itrsum() gives the sum). */
int relaxTipStrokes_omp(int nx, int ny, int probeStart, int relaxAlg, int nstep, double *rTips, double *rs,
                        double *fs, ppafm_spline_params *splineParams);
/* ProbeParticle.cpp:703  dynamical matrix of the probe at n relaxed positions
 * rPPs for tip positions rTips (n,3): rows -(F(r+h e_k) - F(r-h e_k))/2h with
 * h = ddisp, made symmetric; eigenvals (n,3) ascending; evec1/2/3 (n,3) the
 * unit eigenvectors of the smallest/middle/largest eigenvalue, written for
 * which >= 1/2/3 (may be NULL otherwise).  A NaN eigenvalue is reported
 * (stderr, ppafm_last_error) where the reference exits the process. */
void stiffnessMatrix(double ddisp, int which, int n, double *rTips, double *rPPs, double *eigenvals, double *evec1,
                     double *evec2, double *evec3, ppafm_spline_params *splineParams);

/* Further force-field grids on the process-wide grid, += like
 * getLennardJonesFF (coefficients per atom, positions relative to the grid
 * origin):
 *   ProbeParticle.cpp:515  vdW -6 c6 / (r^8 + ADamp c6) r, ADamp 180 (cLJs: c6, c12).
 *                          Its energy is 0, as in the reference. */
void getVdWFF(int natoms, double *Rs, double *cLJs);
/* ProbeParticle.cpp:577  vdW with damping function kind 1..4 (R2, R4,
 * invR4, invR8) of (r/R)^2, REs (R, E0); ADamp > 0 replaces the kind's
 * constant (0.5, 0.5, 0.03, 0.01) for this and later calls; other kinds add
 * nothing. */
void getVdWFF_RE(int natoms, double *Rs, double *REs, int kind, double ADamp);
/* ProbeParticle.cpp:523  DFT-D3 dispersion, coefficients (c6, c8, R0^6,
 * R0^8) per atom from computeD3Coeffs; pairs beyond 20 A add nothing. */
void getDFTD3FF(int natoms, double *Rs, double *d3_coeffs);
/* ProbeParticle.cpp:531  Morse E0 (e^{2a(r-R0)} - 2 e^{a(r-R0)}), REs (R0, E0). */
void getMorseFF(int natoms, double *Rs, double *REs, double alpha);
/* ProbeParticle.cpp:601, 610, 619  densities into the E grid only (the F
 * grid is left alone): amp exp(-(r^2/s)^2/2), amp exp(-r/s),
 * amp (1 - r^2/R^2)^2 inside R; cRAs (amp, width). */
void getGaussDensity(int natoms, double *Rs, double *cRAs);
void getSlaterDensity(int natoms, double *Rs, double *cRAs);
void getDensityR4spline(int natoms, double *Rs, double *cRAs);
/* ProbeParticle.cpp:379  LJ force and energy at npoints points (npoints,3)
 * -> FEs (npoints,4), separation taken atom minus point as there. */
void getInPoints_LJ(int npoints, double *points, double *FEs, int natoms, double *Rs, double *cLJs);
/* ProbeParticle.cpp:556  one pair term at separations (r,0,0) -> Es, Fs (x
 * force); kind -3 LJ in (R, E), -2 LJ (c6, c12), -1 bare r^-6, 0 vdW with
 * constant damping, 1..4 damping functions as getVdWFF_RE (ADamp > 0
 * replaces the kind's constant). */
void evalRadialFF(int n, double *rs, double *coefs, double *Es, double *Fs, int kind, double ADamp);
/* ProbeParticle.cpp:397  DFT-D3 coefficients per atom from the caller's
 * tables (r_cov[94], ref_cn[94*5], ref_c6[94*94*5*5], r4r2[94], k[3],
 * params {s6, s8, a1, a2}); r_cut is accepted and unused, as there.
 * Reference coordination slots < 0 of the probe get weight 0 (the reference
 * leaves them uninitialised). */
void computeD3Coeffs(int natoms, const double *rs, const int *elems, const double *r_cov, const double *r_cut,
                     const double *ref_cn, const double *ref_c6, const double *r4r2, const double *k,
                     const double *params, int elem_pp, double *d3_coeffs);
/* ProbeParticle.cpp:752, 765  cubic Hermite spline (n knots, ydys (n,2)
 * values and slopes) sampled at m points; uniform knots x0 + i dx, or knots
 * xs with the reference's carried-over interval search.  Samples beyond the
 * last interval extrapolate it (the reference reads past the table). */
void subsample_uniform_spline(double x0, double dx, int n, double *ydys, int m, double *xs, double *ys);
void subsample_nonuniform_spline(int n, double *xs, double *ydys, int m, double *xs_, double *ys_);
/* ProbeParticle.cpp:786  radial tip force at r0 + k dr (k < n) about R:
 * type 1 the spline, type 2 the harmonic spring of setTip; fs (n,3). */
void test_force(int type, int n, double *r0, double *dr, double *R, double *fs, ppafm_spline_params *splineParams);
/* ProbeParticle.cpp:805  eigenvalues of a 3x3 matrix in Mat3::eigenvals'
 * order (largest, smallest, middle for a positive largest element) and an
 * eigenvector of each: evs (4,3). */
void test_eigen3x3(double *mat, double *evs);

/* ------------------------------------------------------------------------
 * (2) Device ABI
 * ---------------------------------------------------------------------- */

/* Regular grid: n points along the rows of cell (a, b, c); point (i,j,k) sits
 * at i*cell[0..2]/n0 + j*cell[3..5]/n1 + k*cell[6..8]/n2 relative to the grid
 * origin; all positions below are relative to the grid origin.  The device
 * grid holds (Fx, Fy, Fz, E) per point, z fastest, as float4 (f64 = 0) or
 * double4 (f64 = 1; every pair summed in FP64). */
typedef struct {
    int n[3];
    double cell[9];
    int f64;
} ppafm_grid;

/* Force-field model of one grid build.
 *  mode 0 (ProbeParticle semantics): F = clamp(F_LJ) + probe_q * clamp(F_El)
 *         with R2SAFE = (double)1e-8f, clamp = io.limit_vec_field to fmax
 *         (fmax <= 0: no clamp); F_El is the unit-charge field of the
 *         stencil below (kind 0: one point, weight 1).
 *  mode 1 (FF.cl evalLJC_QZs_noPos semantics): F = F_LJ + F_El, R2SAFE 1e-6,
 *         no clamp; the stencil carries the tip charges Qs (weights) at
 *         QZs (z offsets).
 * Coulomb charges enter as atoms[:,3] * coulomb_const. */
typedef struct {
    int mode;
    int with_lj;
    int n_stencil;          /* 0..4 Coulomb evaluation points */
    int with_energy;        /* write E into the .w channel (else 0) */
    double stencil_w[4];
    double stencil_dz[4];
    double coulomb_const;
    double probe_q;
    double fmax;
    double fp64_radius;     /* float4 grids only.  > 0: atom-point pairs closer
                               than this [A] are evaluated in FP64 (at least 0.2 A
                               beyond the largest stencil offset), the rest in FP32
                               with FP64 accumulation.  0: every pair in FP32 with
                               FP64 accumulation; only points where an FP32 pair
                               overflows (an atom within ~0.005 A) are summed in
                               FP64 */
} ppafm_ff_model;

/* One atom as the grid kernels read it: position relative to the grid
 * origin, charge already multiplied by the Coulomb constant, LJ c6/c12. */
typedef struct {
    double x, y, z, q, c6, c12;
} ppafm_atom;

/* Pack host arrays atoms_xyzq (natoms,4) and cLJs (natoms,2, may be NULL)
 * into natoms ppafm_atom records at out_host, q scaled by coulomb_const. */
int ppafm_pack_atoms(int natoms, const double *atoms_xyzq, const double *cLJs, double coulomb_const,
                     ppafm_atom *out_host);

/* Build the float4 / double4 (grid->f64) grid from device-resident atoms (no
 * host work, no synchronisation: capturable in a HIP graph). */
int ppafm_ff_build_dev(const ppafm_grid *grid, const ppafm_ff_model *model, int natoms, const ppafm_atom *atoms_dev,
                       void *out_dev, void *stream);

/* Build the grid out[(ix*ny+iy)*nz+iz] (float4 or double4) on the device.
 * atoms_xyzq (natoms,4) and cLJs (natoms,2) are HOST float64 arrays
 * (positions relative to the grid origin).  Waits for earlier work on
 * the stream before it stages the atoms. */
int ppafm_ff_build(const ppafm_grid *grid, const ppafm_ff_model *model, int natoms, const double *atoms_xyzq,
                   const double *cLJs, void *out_dev, void *stream);

/* ppafm_ff_build without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (serving loops,
 * harnesses that forbid hipMalloc).  The packed atoms are staged in a fixed
 * array of 65536 records (3 MiB) in device memory of each GPU the library
 * runs on, held until the process ends; natoms above 65536 is an error.
 * Calls on one GPU share that array and keep the last call's atoms: calls
 * on one stream are ordered, calls that overlap on one GPU race.  Waits for
 * earlier work on the stream before it stages the atoms (a stream
 * synchronisation, so the call cannot be recorded into a HIP graph). */
int ppafm_ff_build_noalloc(const ppafm_grid *grid, const ppafm_ff_model *model, int natoms, const double *atoms_xyzq,
                           const double *cLJs, void *out_dev, void *stream);

/* Relaxation semantics.
 *  mode 0 (ProbeParticle.cpp): FIRE state (dt, alpha) reset for every tip
 *         position; output force = grid force sampled at the final position.
 *  mode 1 (relax.cl relaxStrokesTilted): dt and alpha persist along the
 *         stroke, dt >= dtmin_frac*dt0, F2SAFE added under the FIRE sqrt;
 *         output force = last evaluated grid force rotated into the tip
 *         frame, .w = energy. */
typedef struct {
    int mode;
    int alg;                /* 1 = FIRE, 0 = damped leap-frog (mode 0 only) */
    int max_iters;
    int probe_start;        /* mode 0: -1, 0, 1 as relaxTipStroke */
    double conv_f2, dt, damping, finc, fdec, falpha, dtmin_frac, f2safe;
} ppafm_relax_params;

/* Tip: rot rows A,B,C (identity for mode 0), d0 = equilibrium offset in the
 * tip frame, kspring (3), krad, lrad, surf_ff (mode 1 constant force along C). */
typedef struct {
    double rot[9];
    double d0[3];
    double kspring[3];
    double krad, lrad, surf_ff;
} ppafm_tip;

/* Scan: nx*ny columns of nz tip positions.  If tips_dev is non-NULL it holds
 * explicit float64 positions (nx,ny,nz,3) in stroke order; otherwise the
 * position of step s of column (i,j) is p0 + i*da + j*db + s*dstep. */
typedef struct {
    int nx, ny, nz;
    double p0[3], da[3], db[3], dstep[3];
} ppafm_scan;

/* Outputs (device, any may be NULL): pos (nx,ny,nz,3) float64 or float32
 * (pos_f32 != 0), force (nx,ny,nz,4) float32, iters (nx,ny,nz) int32,
 * counters[2] uint64 accumulated: total force evaluations, total positions,
 * status (nx,ny,nz) uint8 per position: 0 converged by the stop rule, bit 0
 * set = the iteration cap ended the step without meeting it, bit 1 set = a
 * non-finite position or grid force (no counterpart in the reference, whose
 * relaxTipStroke test itr > maxIters, ProbeParticle.cpp:658-661, never
 * fires).  Layout index ((i*ny + j)*nz + s) (column-major over the scan
 * lattice matches (nx,ny,nz) C order). */
typedef struct {
    void *pos;
    int pos_f32;
    void *force;
    int *iters;
    unsigned long long *counters;
    unsigned char *status;
} ppafm_relax_out;

int ppafm_relax(const ppafm_grid *grid, const void *grid_dev, const ppafm_relax_params *rp, const ppafm_tip *tip,
                const ppafm_scan *scan, const double *tips_dev, const ppafm_relax_out *out, void *stream);

/* stiffnessMatrix (ProbeParticle.cpp:703-750) on a device grid for the
 * positions of a scan: pos_dev (nx,ny,nz,3) float32 (pos_f32) or float64 in
 * stroke order -- the relaxation's pos output -- with tip positions from
 * `scan` or tips_dev as for ppafm_relax, mode-0 tip (tip->rot unused).
 * evals_dev (npos,3) float64 ascending; evec1/2/3_dev (npos,3) float64 for
 * which >= 1/2/3.  NaN positions give NaN outputs.  No synchronisation. */
int ppafm_stiffness(const ppafm_grid *grid, const void *grid_dev, const ppafm_tip *tip, const ppafm_scan *scan,
                    const double *tips_dev, const void *pos_dev, int pos_f32, double ddisp, int which,
                    double *evals_dev, double *evec1_dev, double *evec2_dev, double *evec3_dev, void *stream);

/* Valid-mode correlation along z for ncol columns (convolveZ, relax.cl:571-608;
 * with reversed weights it is np.convolve(..., 'valid') of common.py:429):
 *   out[c*nzout + j] = scale * sum_{k<nw} in[c*in_stride + (j+k)*in_step + in_offset] * w[k]
 * with nzout = nz - nw + 1.  in/out are float32 (f64 = 0) or float64
 * (f64 = 1) device arrays; w is a host array of nw <= 64 weights. */
int ppafm_convolve_z(long ncol, int nz, int in_stride, int in_step, int in_offset, const void *in_dev, int nw,
                     const double *w, double scale, void *out_dev, int f64, void *stream);

/* io.limit_vec_field (io.py:766-778) in place on a float64 (npts,3) device
 * array: vectors longer than fmax are scaled to length fmax. */
int ppafm_limit_vec_field(double *F_dev, long npts, double fmax, void *stream);

/* Device-to-device copy by a plain kernel (bytes a multiple of 16), the
 * known-size reference for HBM byte counters. */
int ppafm_copy_kernel(const void *src_dev, void *dst_dev, unsigned long long bytes, void *stream);

/* ---------------------------------------------------------------- batches
 * No counterpart in the reference, which runs one molecule per call
 * (AFMulator.__call__, AFMulator.py:142-179; relaxTipStrokes_omp,
 * ProbeParticle.cpp:682-701).  A batch plan holds many molecules of uneven
 * atom counts, grid sizes and scans and runs grid build -> FIRE relaxation
 * -> df for all of them as one captured HIP graph: one grid-build launch
 * over every brick of every grid, one persistent relaxation launch fed by
 * per-XCD work queues, and one z-correlation launch per run of equal nz.
 *
 * Device layout: grids back to back (float4 or double4, z fastest), outputs
 * back to back per molecule in stroke order: pos float32 (.., 3), force
 * float32 (.., 4), iters int32, df float32 (ncol, nz - nw + 1). */
typedef struct {
    int n[3];         /* grid points along a, b, c */
    double cell[9];   /* grid cell rows a, b, c; atoms and scan in the grid frame */
    int natoms;       /* atoms of this molecule in the packed table */
    ppafm_scan scan;  /* apex scan of this molecule */
} ppafm_batch_item;

/* Launch shape of the persistent relaxation: pixel tiles of
 * (64 / tile_y) x tile_y columns per wave (tile_y in 1, 2, 4, 8, 16),
 * waves_per_block in 1..4, blocks_per_cu persistent workgroups per CU.
 * refill 1: each lane takes the next pixel column (in tile order) as soon as
 * its own is done, instead of the wave taking the next tile when all 64 are;
 * refill 2 and 3: the wave takes 64 (2: one tile) or 128 (3) consecutive
 * columns at a time into a pool of its own, and a lane whose column is done
 * takes the pool's next one, so a wave's lanes stay on neighbouring columns.
 * Refill is used when every item shares one grid shape, cell and scan
 * (otherwise the tile queue runs).  Results do not depend on the launch
 * shape. */
typedef struct {
    int tile_y;
    int waves_per_block;
    int blocks_per_cu;
    int refill;
} ppafm_batch_tuning;

typedef struct ppafm_batch ppafm_batch;
typedef struct ppafm_ws_model ppafm_ws_model;

/* model / rp / tip as for ppafm_ff_build_dev and ppafm_relax (rp->alg must
 * be 1).  df: nw (<= 64) host weights correlated with force channel 2 of
 * each column, times df_scale; nw = 0 skips df.  keep_iters: allocate the
 * per-position iteration and status outputs (status as in ppafm_relax_out).  tuning NULL = defaults.  stream NULL = the
 * plan creates its own.  Returns NULL on error (ppafm_last_error). */
ppafm_batch *ppafm_batch_create(int nitems, const ppafm_batch_item *items, const ppafm_ff_model *model,
                                const ppafm_relax_params *rp, const ppafm_tip *tip, int grid_f64, int nw,
                                const double *df_w, double df_scale, int keep_iters,
                                const ppafm_batch_tuning *tuning, void *stream);
/* Upload the packed atoms of all molecules (ppafm_pack_atoms records, item
 * order, natoms of each item).  Synchronous. */
int ppafm_batch_set_atoms(ppafm_batch *b, const ppafm_atom *atoms);
/* The same with new atom counts per item (natoms[nitems]): the molecules of
 * a plan can change between runs while grids and scans stay, as in a
 * training-data generator.  Grows the atom buffer (and captures the graph
 * again) when the new total exceeds it.  Synchronous. */
int ppafm_batch_set_atoms_n(ppafm_batch *b, const int *natoms, const ppafm_atom *atoms);
/* Enqueue the whole batch (graph launch) on the plan's stream. */
int ppafm_batch_run(ppafm_batch *b);
/* Enqueue single stages without the graph: bit 0 grid, bit 1 relaxation
 * (with the step-0 estimates first when warm starts with a model are on),
 * bit 2 df, bit 3 estimates only, bit 4 relaxation only (for timing stages
 * separately). */
int ppafm_batch_run_stages(ppafm_batch *b, int stages);
/* Autotuner: time every launch shape of the persistent relaxation as the
 * plan runs it (warm starts included) on the first nsample molecules (grids
 * and step-0 estimates computed once), keep the fastest, re-capture the
 * graph and return it in best (and its time in best_ms, may be NULL). */
int ppafm_batch_tune(ppafm_batch *b, int nsample, int reps, ppafm_batch_tuning *best, double *best_ms);
int ppafm_batch_set_tuning(ppafm_batch *b, const ppafm_batch_tuning *tuning);
int ppafm_batch_get_tuning(const ppafm_batch *b, ppafm_batch_tuning *tuning);
/* Device buffers: 0 grids, 1 pos, 2 force, 3 iters, 4 df, 5 counters
 * (uint64[3]: force evaluations, positions, warm-start fallbacks), 6 packed
 * atoms, 7 step-0 estimates of the network (pixel columns, 3) float32,
 * 8 per-position status uint8 (with keep_iters). */
void *ppafm_batch_buffer(const ppafm_batch *b, int which);
/* Per-item element offsets (nitems + 1 entries each, NULL to skip): grid
 * points, positions, df values. */
int ppafm_batch_offsets(const ppafm_batch *b, long long *grid_off, long long *pos_off, long long *df_off);
void *ppafm_batch_stream(const ppafm_batch *b);
void ppafm_batch_destroy(ppafm_batch *b);
/* Warm starts for the batch's relaxation.  No counterpart in the reference:
 * relaxTipStroke (ProbeParticle.cpp:644-656) starts FIRE at every step from
 * the continuation (the previous step's result moved with the tip).  With
 * warm starts each step first runs Newton iterations on the interpolated
 * field from that same continuation (step 0: from the network's estimate
 * when a model is given, else from the plain start tip + probe offset):
 * J dr = -F with the exact Jacobian of the trilinear interpolant plus the
 * springs', each step at most newton_cap long and taken only where the
 * Jacobian is stable (all eigenvalues with negative real part, so the
 * iterates head for a minimum, not a saddle).  A step shorter than
 * newton_tol ends them: FIRE starts there.  Otherwise (an unstable Jacobian,
 * a step above the cap, newton_max iterations, newton_max0 at step 0) FIRE
 * starts from the continuation, as without warm starts, and counters[2]
 * counts the step.  While FIRE runs, every newton_retry of its iterations a
 * Newton step is tried from its position; where the Jacobian is stable,
 * Newton iterations take over as above and FIRE restarts at their zero
 * (velocity zero), its iteration count running on toward max_iters.  FIRE
 * ends every step by the plan's own stop rule (conv_f2, max_iters), and the
 * next step continues from FIRE's result.
 * Newton evaluations count in counters[0].  With a model, stage bit 3 (8)
 * computes the step-0 estimates (buffer 7): net_passes network passes, each
 * on the patch centred at the last estimate; step 0's Newton run starts
 * there.
 * The model must outlive the plan's runs; the graph is re-captured. */
typedef struct ppafm_warm_params {
    int enable;        /* 0: plain FIRE, the other fields are ignored */
    int net_passes;    /* network passes at step 0 (model given), 1..8 */
    int newton_max;    /* Newton iterations per step > 0, 0..64 (0: plain start) */
    int newton_max0;   /* Newton iterations at step 0, 0..64 (0: the estimate or plain start as is) */
    int newton_retry;  /* FIRE iterations between Newton tries from FIRE's position, 0..1000 (0: none) */
    double newton_tol; /* A: a shorter Newton step means converged */
    double newton_cap; /* A: longest Newton step */
    double extrap_max; /* A, 0 = off: steps >= 2 start their Newton run from the continuation
                          moved by the change of the probe's offset from the tip over the last
                          two steps, du = u(s-1) - u(s-2) (a second-order predictor), where
                          |du| <= extrap_max; FIRE fallbacks still start from the continuation */
    double residual_max; /* eV/A, 0 = off: with a model and newton_max0 > 0, step 0 evaluates the
                            total force at the network's estimate first; where |F| there exceeds
                            residual_max (or is not finite) the column's step 0 is plain FIRE from
                            the reference's start (counted as a fallback step) */
} ppafm_warm_params;
int ppafm_batch_set_warmstart(ppafm_batch *b, const ppafm_warm_params *p, ppafm_ws_model *m);
/* FIRE from the plan's relaxed positions (buffer 1) to |F|^2 < conv_f2
 * (at most max_iters steps, state reset per position) into out_dev
 * (positions, 3) float32; iters_dev optional.  Enqueued on the plan's stream. */
int ppafm_batch_polish(ppafm_batch *b, double conv_f2, int max_iters, float *out_dev, int *iters_dev);
/* Training pairs for the warm-start network at nsel selected positions
 * (global position indices, device int64): patch rows (nsel, 232) bf16 around
 * the continuation start built from the relaxed positions of the last run,
 * moved by a pseudo-random offset of up to jitter A laterally (half that
 * along z; seed selects the offsets), targets (nsel, 3) = (truth - patch
 * centre node) / 0.1 A.  Enqueued on the plan's stream. */
int ppafm_batch_ws_samples(ppafm_batch *b, const float *truth_dev, long long nsel, const long long *sel_dev,
                           unsigned short *patches_dev, float *targets_dev, double jitter, unsigned seed);

/* ------------------------------------------------------- learned warm start
 * No counterpart in the reference.  A 3D-conv network (3 -> 32 -> 32 -> 3
 * channels, kernels 3x3x3, 3x3x1, 1x1x1, SiLU) maps the RMS-normalised
 * total-force patch (5x5x3 grid nodes x 3 components, bf16 rows of 232
 * values: index ((x*5 + y)*3 + z)*3 + c) around a probe's start to its
 * relaxed displacement (in units of 0.1 A from the patch centre node); each
 * force component of the patch is scaled by its own RMS.
 * Matrix products on bf16 MFMA with FP32 accumulation; FP32 master weights
 * (11,971 values: W1[32][81] b1[32] W2[32][288] b2[32] W3[3][32] b3[3]);
 * Adam.  The model keeps scratch for minibatches of up to max_batch samples. */
ppafm_ws_model *ppafm_ws_create(int max_batch);
void ppafm_ws_destroy(ppafm_ws_model *m);
int ppafm_ws_nparams(void);
int ppafm_ws_set_params(ppafm_ws_model *m, const float *theta_host, int n);
int ppafm_ws_get_params(const ppafm_ws_model *m, float *theta_host, int n);
/* Training state for checkpoints: parameters, Adam first and second moments
 * (ppafm_ws_nparams() floats each, host) and the Adam step count.  set_state
 * restores all of them (set_params resets the moments and the step). */
int ppafm_ws_get_state(const ppafm_ws_model *m, float *theta_host, float *m1_host, float *m2_host, long long *step);
int ppafm_ws_set_state(ppafm_ws_model *m, const float *theta_host, const float *m1_host, const float *m2_host,
                       long long step);
/* out_dev (n, 3) float32 = network(patches_dev (n, 232) bf16). */
int ppafm_ws_forward(ppafm_ws_model *m, int n, const unsigned short *patches_dev, float *out_dev, void *stream);
/* Mean squared error over n <= max_batch samples and its gradient (host,
 * 11,971 values), without updating the model.  Synchronous. */
int ppafm_ws_loss_grad(ppafm_ws_model *m, int n, const unsigned short *patches_dev, const float *targets_dev,
                       float *grad_host, float *loss_host);
/* Adam over consecutive minibatches of `batch` samples (a multiple of 64)
 * of the n samples (cycling), `steps` minibatches (<= 0: one pass), cosine
 * learning rate from lr0 to lr0 * lr_end_frac; per-step losses to
 * losses_host (steps values, may be NULL).  Synchronous. */
int ppafm_ws_train(ppafm_ws_model *m, long n, const unsigned short *patches_dev, const float *targets_dev, int batch,
                   long steps, float lr0, float lr_end_frac, float *losses_host, void *stream);

/* Memory helpers so Python callers need no other GPU runtime. */
int ppafm_malloc(void **ptr, unsigned long long bytes);
int ppafm_free(void *ptr);
int ppafm_memcpy_h2d(void *dst, const void *src, unsigned long long bytes, void *stream);
int ppafm_memcpy_d2h(void *dst, const void *src, unsigned long long bytes, void *stream);
int ppafm_memcpy_d2d(void *dst, const void *src, unsigned long long bytes, void *stream);
int ppafm_memset(void *dst, int value, unsigned long long bytes, void *stream);
int ppafm_synchronize(void *stream);
int ppafm_device_count(void);
/* Name (gcnArchName) and compute-unit count of the current device. */
int ppafm_device_info(char *name, int len, int *ncu);
/* Measured peak rate of the current device in TFLOP/s, for the rooflines:
 * kind 0 FP32 vector FMA (unpacked, as every kernel here is built), 1 FP64
 * vector FMA, 2 bf16 MFMA (v_mfma_f32_16x16x32_bf16); iters sets the run
 * length.  No counterpart in the reference. */
int ppafm_measure_peak(int kind, int iters, double *tflops);
/* HIP events for timing (elapsed waits for ev1) and non-blocking streams of
 * default priority. */
int ppafm_event_create(void **ev);
int ppafm_event_record(void *ev, void *stream);
int ppafm_event_elapsed_ms(void *ev0, void *ev1, float *ms);
int ppafm_event_destroy(void *ev);
int ppafm_stream_create(void **stream);
int ppafm_stream_destroy(void *stream);
const char *ppafm_last_error(void);
/* Clear the calling thread's last error (the reference-compatible void
 * entry points report failures only through ppafm_last_error and stderr). */
void ppafm_clear_error(void);
/* ------------------------------------------------------------------------
 * (3) Grid files (host code, multithreaded; nthreads <= 0 = all cores)
 * ---------------------------------------------------------------------- */

/* The data block of an XSF grid as ppafm/io.py:575-599 (saveXSFData) writes
 * it, appended to (append != 0) or replacing fname: the (nx,ny,nz) grid
 * with one periodic plane repeated along each axis, x fastest, one "%10.5e"
 * value per line, then "   END_DATAGRID_3D" and "END_BLOCK_DATAGRID_3D".
 * Element (ix,iy,iz) is data[ix*sx + iy*sy + iz*sz] (host float64). */
int ppafm_xsf_write_data(const char *fname, int append, const double *data, int nx, int ny, int nz, long long sx,
                         long long sy, long long sz, int nthreads);
/* The same from channel `channel` (0..3) of a device grid (nx,ny,nz) of
 * float4 (f64 = 0) or double4 (f64 = 1) values, copied to the host first. */
int ppafm_xsf_write_data_dev(const char *fname, int append, const void *grid_dev, int f64, int channel, int nx, int ny,
                             int nz, int nthreads);
/* GridUtils.cpp:32 ReadNumsUpTo_C: skip noline lines (as fgets with a
 * 5000-byte buffer skips them), then read count whitespace-separated
 * numbers into out; fewer numbers or a non-number field is an error. */
int ppafm_read_numbers(const char *fname, int noline, long long count, double *out, int nthreads);

/* Sum of iteration indices of the last relaxTipStrokes_omp call. */
long long ppafm_last_itrsum(void);

/* ------------------------------------------------------------------------
 * (4) Grids of the Hartree and full-density-based force fields
 *     (ppafm/ocl/field.py DataGrid family, FFTCrossCorrelation,
 *     ForceField_LJC.calc_force_hartree / calc_force_fdbm; kernels of
 *     ppafm/ocl/cl/FF.cl).  Device ABI conventions as in (2).
 *
 * A grid is n[0]*n[1]*n[2] float64 values, C order (z fastest), with a
 * lattice lvec[12] = 4 rows (origin, a, b, c); its step vectors are a/n[0],
 * b/n[1], c/n[2].  The reference keeps these grids in FP32; here every
 * value and every coordinate is FP64.  Grids hold fewer than 2^31 points;
 * input and output grids of one call are different arrays unless stated. 
 * ---------------------------------------------------------------------- */

/* DataGrid.interp_at (field.py:528-575), kernel interp_at (FF.cl:1010-1046):
 * out[i,j,k] = periodic trilinear interpolation of `in` at the point
 * rot_center + rot (p - rot_center), p = out origin + i a' + j b' + k c'
 * (a', b', c' the out step vectors; rot row-major 3x3). */
int ppafm_grid_interp_at(const double in_lvec[12], const int in_n[3], const double *in_dev, const double out_lvec[12],
                         const int out_n[3], double *out_dev, const double rot[9], const double rot_center[3],
                         void *stream);
/* TipDensity.interp_at (field.py:607-660), kernel interp_tip_at
 * (FF.cl:1049-1093): the tip sits on the origin of both grids (their
 * origins are ignored); out indices past half way wrap to negative
 * offsets; points whose coordinates inv(A) p (A = rows a, b, c of in_lvec)
 * exceed 1/2 in magnitude are 0, others interpolate `in` periodically. */
int ppafm_grid_interp_tip_at(const double in_lvec[12], const int in_n[3], const double *in_dev,
                             const double out_lvec[12], const int out_n[3], double *out_dev, void *stream);
/* DataGrid.grad (field.py:471-526), kernel grad (FF.cl:920-961): periodic
 * centred differences along the three axes with steps step[3],
 * out[p] = scale * (dV/dx, dV/dy, dV/dz, V) as double4; out in C order
 * (order_c != 0) or Fortran order (x fastest); accumulate != 0 adds to out
 * instead of overwriting it (the force field of calc_force_hartree is
 * this gradient plus the Lennard-Jones grid). */
int ppafm_grid_grad(const int n[3], const double step[3], const double *in_dev, const double scale[4], int order_c,
                    int accumulate, double *out_dev, void *stream);
/* ForceField_LJC.run_gradPotentialGrid (field.py:1225-1317), kernel
 * gradPotentialGrid (FF.cl:964-1007): the field -grad V on the grid
 * (lvec, n) by centred differences of step h between trilinear
 * interpolations of the potential grid pot at p -/+ h along x, y, z;
 * field_dev (n, double4) = (Ex, Ey, Ez, V(p)), C order. */
int ppafm_grid_grad_potential(const double pot_lvec[12], const int pot_n[3], const double *pot_dev,
                              const double lvec[12], const int n[3], double h, double *field_dev, void *stream);
/* The electrostatic part of evalLJC_Hartree (FF.cl:860-917): for every
 * point p of the grid (lvec, n), grid4 (n, double4) += sum over q with
 * Qs[q] != 0 of Qs[q] * (periodic trilinear interpolation of field_dev
 * (n, double4, same grid) at p + (0, 0, QZs[q])). */
int ppafm_grid_add_tip_charges(const double lvec[12], const int n[3], const double *field_dev, const double Qs[4],
                               const double QZs[4], double *grid4_dev, void *stream);
/* normalizeSumReduce + sumSingleGroup (FF.cl:1112-1176) as used by
 * DataGrid._get_normalization_factor (field.py:447-469): sums[0] = sum of
 * all n values, sums[1] = sum of the positive ones (host array; the call
 * waits for the stream).  The reduction order is fixed for a given n. */
int ppafm_grid_sums(long long n, const double *in_dev, double sums[2], void *stream);
/* Kernel power (FF.cl:1096-1107): out = scale * max(0, in)^p. */
int ppafm_grid_power(long long n, const double *in_dev, double *out_dev, double p, double scale, void *stream);
/* Kernel addMult (FF.cl:1213-1230): out = in1 + A * in2. */
int ppafm_grid_add_mult(long long n, const double *in1_dev, const double *in2_dev, double A, double *out_dev,
                        void *stream);
/* Kernels clamp_hard / clamp_soft (FF.cl:1179-1211): hard clips to
 * [lo, hi]; soft (soft != 0) applies the sigmoid transitions of width
 * `width` at each finite bound. */
int ppafm_grid_clamp(long long n, const double *in_dev, double *out_dev, double lo, double hi, int soft, double width,
                     void *stream);
/* MultipoleTipDensity._make_tip_density (field.py:671-690) with
 * getProbeDensity (fieldFFT.py:99-118) on the device: coordinates
 * X = cx[i], Y = cy[j], Z = cz[k] (host arrays, the wrapped linspace
 * coordinates along each lattice vector), r = mat (X, Y, Z) with mat
 * row-major 3x3 (the transposed normalised lattice), Gaussian
 * exp(-|r|^2 / (2 sigma^2)) normalised by its grid sum times
 * norm[0] * norm[1] * norm[2] * norm[3] (|det mat| and the three steps,
 * multiplied in that order), times sum_k coefs[k] Y_k(r / sigma) with the
 * harmonic of kinds[k] (0 s, 1 px, 2 py, 3 pz, 4 dz2, 5 dx2, 6 dy2, 7 dxy,
 * 8 dxz, 9 dyz, other values contribute 0) tilted by `tilt` in the zx
 * plane; nkinds < 0 keeps the normalised Gaussian.  nkinds <= 16. */
int ppafm_tip_density_multipole(const int n[3], const double *cx, const double *cy, const double *cz,
                                const double mat[9], double sigma, double tilt, int nkinds, const int *kinds,
                                const double *coefs, const double norm[4], double *out_dev, void *stream);

/* FFTCrossCorrelation (field.py:707-843): circular cross-correlation of
 * grids of shape n with a fixed tip density, by real-to-complex FP64 FFTs:
 * E = scale * cell_vol * IFFT(conj(FFT(rho)) FFT(in)) with the normalised
 * inverse, i.e. E[r] = scale * cell_vol * sum_s rho[s] in[s + r]
 * (indices periodic).  create makes the plans and buffers; set_rho
 * transforms rho (_set_rho, field.py:800-802; cell_vol is its grid cell
 * volume); correlate leaves in_dev unchanged and may write E_dev = in_dev.
 * Calls on one object are serialised by the caller. */
typedef struct ppafm_fft_corr ppafm_fft_corr;
int ppafm_fft_corr_create(const int n[3], ppafm_fft_corr **out);
int ppafm_fft_corr_set_rho(ppafm_fft_corr *c, const double *rho_dev, double cell_vol, void *stream);
int ppafm_fft_corr_correlate(ppafm_fft_corr *c, const double *in_dev, double scale, double *E_dev, void *stream);
int ppafm_fft_corr_destroy(ppafm_fft_corr *c);

/* addvdW (FF.cl:548-665) onto a double4 grid (C order, points of `grid`
 * relative to its origin; grid->f64 ignored): damp_method -1 bare
 * -C6/(r^2 + 1e-6)^3, 0 constant damping (force only), 1..4 the R2, R4,
 * invR4, invR8 damping functions of u^2 = r^2/R^2.  pos_dev (natoms,3)
 * relative to the grid origin; coefs_dev (natoms,2): (c6, c12) for
 * damp_method <= 0, (R, E0) otherwise. */
int ppafm_grid_add_vdw(const ppafm_grid *grid, int damp_method, int natoms, const double *pos_dev,
                       const double *coefs_dev, double *grid4_dev, void *stream);
/* addDFTD3_BJ (FF.cl:790-857) onto a double4 grid: coefs_dev (natoms,4) =
 * (c6 s6, c8 s8, R0^6, R0^8) as computeD3Coeffs returns them; pairs beyond
 * r^2 > 400 are skipped. */
int ppafm_grid_add_d3(const ppafm_grid *grid, int natoms, const double *pos_dev, const double *coefs_dev,
                      double *grid4_dev, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* PPAFM_HIP_H */
