/*
 * pp_oracle.c -- FP64 CPU oracle for the probe-particle hot path.
 *
 * TEST INFRASTRUCTURE.  This file is the checker, never the product: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * the library built from it.  The product path (ppafm_amd) runs on the GPU
 * and fails loudly when its HIP library is missing; it never calls in here.
 *
 * What it restates (semantics, written from scratch; file:line into the
 * reference ppafm v0.4.0 tree):
 *   - grid geometry   GridShape::setCell, Mat3::invert_T_to   Grid.h:29-36, Mat3.h:275-286
 *   - LJ pair term    addAtomLJ                               Forces.h:120-132
 *   - Coulomb pair    addAtomCoulomb (+ pz / dz2 stencils)    Forces.h:237-244, ProbeParticle.cpp:178-209
 *   - grid iteration  iterateGrid3D_omp + evalCell            Grid.h:142-170, ProbeParticle.cpp:262-278
 *   - force clamp     io.limit_vec_field                      io.py:766-778
 *   - interpolation   SET_GRID_IND + interpolate3DvecWrap     Grid.h:66-111
 *   - tip force       getPPforce + forceRSpring               ProbeParticle.cpp:282-294, Forces.h:85-89
 *   - FIRE            FIREstate::move, relaxProbe             ProbeParticle.cpp:142-174, 297-313
 *   - strokes         relaxTipStroke / relaxTipStrokes_omp    ProbeParticle.cpp:633-701
 *   - OpenCL-mode grid  evalLJC_QZs_noPos (FP64, x-fastest)   ocl/cl/FF.cl:115-130, 366-423
 *   - OpenCL-mode FIRE  relaxStrokesTilted, update_FIRE       ocl/cl/relax.cl:37-41, 90-126, 391-464
 *   - z convolution     convolveZ (valid-mode correlation)    ocl/cl/relax.cl:571-608
 *   - stiffness       stiffnessMatrix, Mat3::eigenvals/vec   ProbeParticle.cpp:703-750, Mat3.h:459-500
 *
 * The operation order of every floating-point expression on the CPU-mode
 * path follows the reference so that, built with gcc -O2 (no FMA
 * contraction on x86-64 baseline), results agree with the compiled
 * reference to the last bit.  Unlike the reference there is no global
 * state: every call takes its grid, tip and relaxation parameters.
 *
 * Parity pins (tests/test_oracle_cpu.py): compiled reference C++ (oracle/_ref),
 * committed golden fixtures, closed forms (single atom, affine fields,
 * harmonic tip minimum, 1-D force balance by bisection), F = -grad E.
 * The OpenCL-mode functions are pinned only by the mathematics (the
 * reference OpenCL path cannot run here).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Forces.h:12 writes R2SAFE as the float literal 1.0e-8f; promoted to double
 * it is 9.99999993922529e-09, which is what the reference adds to r^2. */
#define ORC_R2SAFE ((double)1.0e-8f)

typedef struct { double x, y, z; } v3;

static inline double v3_dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline double v3_norm2(v3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }

/* ---------------------------------------------------------------- geometry */

/* dcell[r] = cell[r]/n[r]; dinv = (dcell^-1)^T with the cofactor order of
 * Mat3::invert_T_to (Mat3.h:254-286). Both 3x3 row-major, rows a,b,c. */
void orc_grid_shape(const int n[3], const double cell[9], double dcell[9], double dinv[9])
{
    for (int r = 0; r < 3; r++) {
        double s = 1.0 / n[r];
        for (int c = 0; c < 3; c++) dcell[3 * r + c] = cell[3 * r + c] * s;
    }
    const double xx = dcell[0], xy = dcell[1], xz = dcell[2];
    const double yx = dcell[3], yy = dcell[4], yz = dcell[5];
    const double zx = dcell[6], zy = dcell[7], zz = dcell[8];
    double c_xx = yy * zz - yz * zy;
    double c_yx = yz * zx - yx * zz;
    double c_zx = yx * zy - yy * zx;
    double idet = 1 / (xx * c_xx + xy * c_yx + xz * c_zx);
    dinv[0] = (yy * zz - yz * zy) * idet;  /* xx */
    dinv[3] = (xz * zy - xy * zz) * idet;  /* yx */
    dinv[6] = (xy * yz - xz * yy) * idet;  /* zx */
    dinv[1] = (yz * zx - yx * zz) * idet;  /* xy */
    dinv[4] = (xx * zz - xz * zx) * idet;  /* yy */
    dinv[7] = (xz * yx - xx * yz) * idet;  /* zy */
    dinv[2] = (yx * zy - yy * zx) * idet;  /* xz */
    dinv[5] = (xy * zx - xx * zy) * idet;  /* yz */
    dinv[8] = (xx * yy - xy * yx) * idet;  /* zz */
}

/* ---------------------------------------------------------- pair functions */

static inline double pair_lj(v3 d, v3 *f, double c6, double c12)
{
    double ir2 = 1.0 / (v3_norm2(d) + ORC_R2SAFE);
    double ir6 = ir2 * ir2 * ir2;
    double e6 = c6 * ir6;
    double e12 = c12 * ir6 * ir6;
    double s = (12 * e12 - 6 * e6) * ir2;
    f->x += d.x * s; f->y += d.y * s; f->z += d.z * s;
    return e12 - e6;
}

static inline double pair_coul(v3 d, v3 *f, double kqq)
{
    double ir2 = 1.0 / (v3_norm2(d) + ORC_R2SAFE);
    double ir = sqrt(ir2);
    double e = ir * kqq;
    double s = e * ir2;
    f->x += d.x * s; f->y += d.y * s; f->z += d.z * s;
    return e;
}

/* Finite-difference multipole probes, dstep 0.1 A (ProbeParticle.cpp:178-209). */
static inline double pair_coul_kind(v3 d, v3 *fout, double kqq, int kind)
{
    if (kind == 0) return pair_coul(d, fout, kqq);
    v3 f = {0, 0, 0};
    double e = 0;
    if (kind == 1) {
        d.z -= 0.1;     e += pair_coul(d, &f, -kqq);
        d.z += 2 * 0.1; e += pair_coul(d, &f, +kqq);
        fout->x += f.x * 10.0; fout->y += f.y * 10.0; fout->z += f.z * 10.0;
        return e * 10.0;
    }
    e += pair_coul(d, &f, -2 * kqq);
    d.z -= 0.1;     e += pair_coul(d, &f, kqq);
    d.z += 2 * 0.1; e += pair_coul(d, &f, kqq);
    fout->x += f.x * 100.0; fout->y += f.y * 100.0; fout->z += f.z * 100.0;
    return e * 100.0;
}

/* ------------------------------------------------------------- grid build */

/* model 0 = LJ (coefs c6,c12 per atom), 1/2/3 = Coulomb kind 0/1/2 (coef kqq).
 * Accumulates (+=) into F[(ix*ny+iy)*nz+iz][3] and E[...] where non-NULL. */
static void grid_eval(int model, const int n[3], const double cell[9], int natoms,
                      const double *Rs, const double *coefs, double *F, double *E)
{
    double dcell[9], dinv[9];
    orc_grid_shape(n, cell, dcell, dinv);
    const int nx = n[0], ny = n[1], nz = n[2];
    #pragma omp parallel for collapse(3) schedule(static)
    for (int ia = 0; ia < nx; ia++)
        for (int ib = 0; ib < ny; ib++)
            for (int ic = 0; ic < nz; ic++) {
                v3 p;
                p.x = dcell[0] * ia + dcell[3] * ib + dcell[6] * ic;
                p.y = dcell[1] * ia + dcell[4] * ib + dcell[7] * ic;
                p.z = dcell[2] * ia + dcell[5] * ib + dcell[8] * ic;
                v3 f = {0, 0, 0};
                double e = 0;
                for (int i = 0; i < natoms; i++) {
                    v3 d = {p.x - Rs[3 * i], p.y - Rs[3 * i + 1], p.z - Rs[3 * i + 2]};
                    if (model == 0) e += pair_lj(d, &f, coefs[2 * i], coefs[2 * i + 1]);
                    else            e += pair_coul_kind(d, &f, coefs[i], model - 1);
                }
                long ig = ((long)ia * ny + ib) * nz + ic;
                if (F) { F[3 * ig] += f.x; F[3 * ig + 1] += f.y; F[3 * ig + 2] += f.z; }
                if (E) E[ig] += e;
            }
}

/* ProbeParticle.cpp:508-513; Rs already shifted by -lvec[0] (HighLevel.py:277). */
void orc_lj_ff(const int n[3], const double cell[9], int natoms, const double *Rs,
               const double *cLJs, double *F, double *E)
{
    grid_eval(0, n, cell, natoms, Rs, cLJs, F, E);
}

/* ProbeParticle.cpp:540-554; kQQs = Qs * 14.399645 for a unit probe charge. */
void orc_coulomb_ff(const int n[3], const double cell[9], int natoms, const double *Rs,
                    const double *kQQs, int kind, double *F, double *E)
{
    if (kind < 0 || kind > 2) return;
    grid_eval(1 + kind, n, cell, natoms, Rs, kQQs, F, E);
}

/* io.py:766-778: scale vectors longer than Fmax to length Fmax, in place. */
void orc_limit_vec_field(long npts, double *F, double fmax)
{
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < npts; i++) {
        double fr = sqrt(F[3 * i] * F[3 * i] + F[3 * i + 1] * F[3 * i + 1] + F[3 * i + 2] * F[3 * i + 2]);
        if (fr > fmax) {
            double s = fmax / fr;
            F[3 * i] *= s; F[3 * i + 1] *= s; F[3 * i + 2] *= s;
        }
    }
}

/* ---------------------------------------------------------- interpolation */

/* One axis of SET_GRID_IND (Grid.h:66-69): lower index lo, upper index hi,
 * weights wlo (for lo) and whi (for hi), correct for negative coordinates. */
static inline void wrap_axis(double r, int n, int *lo, int *hi, double *wlo, double *whi)
{
    int i;
    if (r >= 0) {
        i = (int)r;
        *whi = r - i;
        *wlo = 1 - *whi;
        i %= n;
    } else {
        int t = (int)r;
        *wlo = t - r;
        *whi = 1 - *wlo;
        i = t % n + n - 1;
    }
    *lo = i;
    *hi = (i + 1) % n;
}

/* interpolate3DvecWrap (Grid.h:93-111), same corner order. g = grid coords. */
static inline v3 interp_wrap(const double *F, const int n[3], v3 g)
{
    int ix, jx, iy, jy, iz, jz;
    double mx, tx, my, ty, mz, tz;
    wrap_axis(g.x, n[0], &ix, &jx, &mx, &tx);
    wrap_axis(g.y, n[1], &iy, &jy, &my, &ty);
    wrap_axis(g.z, n[2], &iz, &jz, &mz, &tz);
    const long ny = n[1], nz = n[2];
#define G3(a, b, c) (F + 3 * (((long)(a) * ny + (b)) * nz + (c)))
    double mxmy = mx * my, mxty = mx * ty, txmy = tx * my, txty = tx * ty;
    const double *c[8] = {G3(ix, iy, iz), G3(ix, iy, jz), G3(ix, jy, iz), G3(ix, jy, jz),
                          G3(jx, jy, iz), G3(jx, jy, jz), G3(jx, iy, iz), G3(jx, iy, jz)};
    const double w[8] = {mxmy * mz, mxmy * tz, mxty * mz, mxty * tz,
                         txty * mz, txty * tz, txmy * mz, txmy * tz};
#undef G3
    v3 o = {c[0][0] * w[0], c[0][1] * w[0], c[0][2] * w[0]};
    for (int k = 1; k < 8; k++) { o.x += c[k][0] * w[k]; o.y += c[k][1] * w[k]; o.z += c[k][2] * w[k]; }
    return o;
}

/* Periodic trilinear sample at Cartesian r (grid origin at 0). */
void orc_interp(const int n[3], const double cell[9], const double *F, int npts,
                const double *rs, double *out)
{
    double dcell[9], dinv[9];
    orc_grid_shape(n, cell, dcell, dinv);
    for (int i = 0; i < npts; i++) {
        v3 r = {rs[3 * i], rs[3 * i + 1], rs[3 * i + 2]};
        v3 g = {v3_dot(r, (v3){dinv[0], dinv[1], dinv[2]}),
                v3_dot(r, (v3){dinv[3], dinv[4], dinv[5]}),
                v3_dot(r, (v3){dinv[6], dinv[7], dinv[8]})};
        v3 f = interp_wrap(F, n, g);
        out[3 * i] = f.x; out[3 * i + 1] = f.y; out[3 * i + 2] = f.z;
    }
}

/* ------------------------------------------------------------ relaxation */

/* Tip: lRadial, kRadial, rPP0 (already on the sphere, TIP::makeConsistent,
 * ProbeParticle.cpp:64-69), kSpring; all in eV/A units. */
typedef struct {
    double lrad, krad;
    double rpp0[3];
    double kspring[3];
    /* radial force spline (TIP::SplineParams); sn = 0: harmonic spring */
    int sn;
    const double *sxs, *sydys;
} orc_tip;

/* Spline_Hermite::find_index / ::val (spline_hermite.h:45-51, 159-168) and
 * forceRSpline (ProbeParticle.cpp:211-224): radial force / distance. */
static inline int find_index(int i, int di, double x, const double *xs)
{
    while (di > 1) {
        di >>= 1;
        int j = i + di;
        if (xs[j] < x) i = j;
    }
    return i;
}

static inline double hermite(double x, double y0, double y1, double dy0, double dy1)
{
    double y01 = y0 - y1;
    return y0 + x * (dy0 + x * (-3 * y01 - 2 * dy0 - dy1 + x * (2 * y01 + dy0 + dy1)));
}

static inline double spline_scale(const orc_tip *t, double x)
{
    int i = find_index(0, t->sn, x, t->sxs);
    if (i > t->sn - 2) i = t->sn - 2;  /* beyond the last knot: extrapolate */
    double x0 = t->sxs[i], x1 = t->sxs[i + 1];
    double dx = x1 - x0, denom = 1 / dx, u = (x - x0) * denom;
    i = i << 1;
    double fr = hermite(u, t->sydys[i], t->sydys[i + 2], t->sydys[i + 1] * dx, t->sydys[i + 3] * dx);
    return fr / x;
}

/* RELAX + FIRE namespaces (ProbeParticle.cpp:74-114). */
typedef struct {
    int max_iters;
    double conv_f2, dt, damping;
    double finc, fdec, falpha;
} orc_relax;

typedef struct {
    const double *F;
    int n[3];
    double dinv[9];
} orc_field;

/* getPPforce (ProbeParticle.cpp:282-294), harmonic radial spring. */
static inline v3 pp_force(const orc_field *ff, const orc_tip *tip, v3 rtip, v3 r)
{
    v3 g = {v3_dot(r, (v3){ff->dinv[0], ff->dinv[1], ff->dinv[2]}),
            v3_dot(r, (v3){ff->dinv[3], ff->dinv[4], ff->dinv[5]}),
            v3_dot(r, (v3){ff->dinv[6], ff->dinv[7], ff->dinv[8]})};
    v3 d = {r.x - rtip.x, r.y - rtip.y, r.z - rtip.z};
    v3 f = interp_wrap(ff->F, ff->n, g);
    double l = sqrt(v3_norm2(d));
    double s = tip->sn ? spline_scale(tip, l) : tip->krad * (l - tip->lrad) / l;
    f.x += d.x * s; f.y += d.y * s; f.z += d.z * s;
    d.x -= tip->rpp0[0]; d.y -= tip->rpp0[1]; d.z -= tip->rpp0[2];
    f.x += d.x * tip->kspring[0]; f.y += d.y * tip->kspring[1]; f.z += d.z * tip->kspring[2];
    return f;
}

/* relaxProbe (ProbeParticle.cpp:297-313) with FIREstate::move (142-174) for
 * alg 1, damped leap-frog RELAX::move (84-88) for alg 0.  Returns the
 * iteration index at which it stopped, like the reference; *nevals gets the
 * number of force evaluations made. */
static int relax_probe(const orc_field *ff, const orc_tip *tip, const orc_relax *rp, int alg,
                       v3 rtip, v3 *r, int *nevals)
{
    v3 v = {0, 0, 0};
    const double dtmax = rp->dt, a0 = rp->damping;
    double dt = dtmax, acoef = a0;
    int it;
    *nevals = 0;
    for (it = 0; it < rp->max_iters; it++) {
        v3 f = pp_force(ff, tip, rtip, *r);
        (*nevals)++;
        if (alg == 1) {
            double ff2 = v3_norm2(f), vv = v3_norm2(v), vf = v3_dot(f, v);
            if (vf < 0) {
                v.x = 0; v.y = 0; v.z = 0;
                dt = dt * rp->fdec;
                acoef = a0;
            } else {
                double cf = acoef * sqrt(vv / ff2);
                double cv = 1 - acoef;
                v.x *= cv; v.y *= cv; v.z *= cv;
                v.x += f.x * cf; v.y += f.y * cf; v.z += f.z * cf;
                dt = fmin(dt * rp->finc, dtmax);
                acoef = acoef * rp->falpha;
            }
            v.x += f.x * dt; v.y += f.y * dt; v.z += f.z * dt;
            r->x += v.x * dt; r->y += v.y * dt; r->z += v.z * dt;
        } else {
            double c = 1 - rp->damping;
            v.x *= c; v.y *= c; v.z *= c;
            v.x += f.x * rp->dt; v.y += f.y * rp->dt; v.z += f.z * rp->dt;
            r->x += v.x * rp->dt; r->y += v.y * rp->dt; r->z += v.z * rp->dt;
        }
        if (v3_norm2(f) < rp->conv_f2) break;
    }
    return it;
}

/* relaxTipStroke (ProbeParticle.cpp:633-676) for one column of nstep tip
 * positions.  iters[i] (optional) = iteration index returned per step,
 * nevals[i] (optional) = force evaluations of FIRE per step (the final
 * sample evaluation not included).  Returns the sum of iters. */
static int relax_stroke(const orc_field *ff, const orc_tip *tip, const orc_relax *rp,
                        int probe_start, int alg, int nstep, const double *rtips,
                        double *rs, double *fs, int *iters, int *nevals)
{
    v3 rtip = {rtips[0], rtips[1], rtips[2]};
    v3 rp_ = {rtip.x + tip->rpp0[0], rtip.y + tip->rpp0[1], rtip.z + tip->rpp0[2]};
    int itrsum = 0;
    for (int i = 0; i < nstep; i++) {
        v3 nt = {rtips[3 * i], rtips[3 * i + 1], rtips[3 * i + 2]};
        if (probe_start == -1) {
            rtip = nt;
        } else if (probe_start == 0) {
            rtip = nt;
            rp_.x = rtip.x + tip->rpp0[0]; rp_.y = rtip.y + tip->rpp0[1]; rp_.z = rtip.z + tip->rpp0[2];
        } else if (probe_start == 1) {
            v3 drp = {rp_.x - rtip.x, rp_.y - rtip.y, rp_.z - rtip.z};
            rtip = nt;
            rp_.x = rtip.x + drp.x; rp_.y = rtip.y + drp.y; rp_.z = rtip.z + drp.z;
        }
        int ne = 0;
        int itr = relax_probe(ff, tip, rp, alg, rtip, &rp_, &ne);
        v3 g = {v3_dot(rp_, (v3){ff->dinv[0], ff->dinv[1], ff->dinv[2]}),
                v3_dot(rp_, (v3){ff->dinv[3], ff->dinv[4], ff->dinv[5]}),
                v3_dot(rp_, (v3){ff->dinv[6], ff->dinv[7], ff->dinv[8]})};
        v3 fsamp = interp_wrap(ff->F, ff->n, g);
        rs[3 * i] = rp_.x; rs[3 * i + 1] = rp_.y; rs[3 * i + 2] = rp_.z;
        fs[3 * i] = fsamp.x; fs[3 * i + 1] = fsamp.y; fs[3 * i + 2] = fsamp.z;
        if (iters) iters[i] = itr;
        if (nevals) nevals[i] = ne;
        itrsum += itr;
    }
    return itrsum;
}

static void make_field(orc_field *ff, const int n[3], const double cell[9], const double *F)
{
    double dcell[9];
    ff->F = F;
    ff->n[0] = n[0]; ff->n[1] = n[1]; ff->n[2] = n[2];
    orc_grid_shape(n, cell, dcell, ff->dinv);
}

static void make_tip(orc_tip *t, const double tip[8])
{
    /* tip = {lRadial, kRadial, rPP0x, rPP0y, rPP0z, kSx, kSy, kSz}, rPP0 as given
     * to setTip; the z component is put on the sphere as TIP::makeConsistent does. */
    t->lrad = tip[0]; t->krad = tip[1];
    t->rpp0[0] = tip[2]; t->rpp0[1] = tip[3]; t->rpp0[2] = tip[4];
    t->kspring[0] = tip[5]; t->kspring[1] = tip[6]; t->kspring[2] = tip[7];
    t->sn = 0; t->sxs = 0; t->sydys = 0;
    if (fabs(t->krad) > 1e-8)
        t->rpp0[2] = -sqrt(t->lrad * t->lrad - t->rpp0[0] * t->rpp0[0] - t->rpp0[1] * t->rpp0[1]);
}

static void make_relax(orc_relax *r, const double relax[7])
{
    /* relax = {maxIters, convF2, dt, damping, finc, fdec, falpha} */
    r->max_iters = (int)relax[0];
    r->conv_f2 = relax[1]; r->dt = relax[2]; r->damping = relax[3];
    r->finc = relax[4]; r->fdec = relax[5]; r->falpha = relax[6];
}

/* One stroke (relaxTipStroke).  Returns sum of iteration indices. */
int orc_relax_stroke(const int n[3], const double cell[9], const double *F, const double tip[8],
                     const double relax[7], int probe_start, int alg, int nstep,
                     const double *rtips, double *rs, double *fs, int *iters, int *nevals)
{
    orc_field ff; orc_tip t; orc_relax r;
    make_field(&ff, n, cell, F); make_tip(&t, tip); make_relax(&r, relax);
    return relax_stroke(&ff, &t, &r, probe_start, alg, nstep, rtips, rs, fs, iters, nevals);
}

/* relaxTipStrokes_omp (ProbeParticle.cpp:682-701): rTips/rs/fs are
 * (nx,ny,nz,3) C order; stroke offset (ix + iy*nx)*nz as in the reference
 * (a bijection of the columns).  iters/nevals optional (nx*ny*nz). Returns
 * the total iteration sum (the reference returns 0; the sum is additive). */
long orc_relax_strokes(int nx, int ny, int nz, const int n[3], const double cell[9], const double *F,
                       const double tip[8], const double relax[7], int probe_start, int alg,
                       const double *rtips, double *rs, double *fs, int *iters, int *nevals)
{
    orc_field ff; orc_tip t; orc_relax r;
    make_field(&ff, n, cell, F); make_tip(&t, tip); make_relax(&r, relax);
    long total = 0;
    #pragma omp parallel for collapse(2) schedule(static) reduction(+ : total)
    for (int ix = 0; ix < nx; ix++)
        for (int iy = 0; iy < ny; iy++) {
            long off = ((long)ix + (long)iy * nx) * nz;
            total += relax_stroke(&ff, &t, &r, probe_start, alg, nz, rtips + 3 * off, rs + 3 * off,
                                  fs + 3 * off, iters ? iters + off : 0, nevals ? nevals + off : 0);
        }
    return total;
}

/* relaxTipStrokes_omp with a radial tip spline of sn knots (sxs, sydys). */
long orc_relax_strokes_spline(int nx, int ny, int nz, const int n[3], const double cell[9], const double *F,
                              const double tip[8], const double relax[7], int sn, const double *sxs,
                              const double *sydys, const double *rtips, double *rs, double *fs)
{
    orc_field ff; orc_tip t; orc_relax r;
    make_field(&ff, n, cell, F); make_tip(&t, tip); make_relax(&r, relax);
    t.sn = sn; t.sxs = sxs; t.sydys = sydys;
    long total = 0;
    #pragma omp parallel for collapse(2) schedule(static) reduction(+ : total)
    for (int ix = 0; ix < nx; ix++)
        for (int iy = 0; iy < ny; iy++) {
            long off = ((long)ix + (long)iy * nx) * nz;
            total += relax_stroke(&ff, &t, &r, 1, 1, nz, rtips + 3 * off, rs + 3 * off, fs + 3 * off, 0, 0);
        }
    return total;
}

/* Same strokes over a subset of columns (list of (ix,iy) pairs) -- used to
 * time a bounded sample of a large scan. */
long orc_relax_strokes_subset(int ncols, const int *cols, int nx, int nz, const int n[3],
                              const double cell[9], const double *F, const double tip[8],
                              const double relax[7], int probe_start, int alg,
                              const double *rtips, double *rs, double *fs, int *iters)
{
    orc_field ff; orc_tip t; orc_relax r;
    make_field(&ff, n, cell, F); make_tip(&t, tip); make_relax(&r, relax);
    long total = 0;
    #pragma omp parallel for schedule(dynamic, 4) reduction(+ : total)
    for (int c = 0; c < ncols; c++) {
        long off = ((long)cols[2 * c] + (long)cols[2 * c + 1] * nx) * nz;
        total += relax_stroke(&ff, &t, &r, probe_start, alg, nz, rtips + 3 * off, rs + 3 * off,
                              fs + 3 * off, iters ? iters + off : 0, 0);
    }
    return total;
}

/* --------------------------------------------- warm starts (Newton + FIRE)
 * No counterpart in the reference: the warm-start method of the GPU batch
 * path.  Every step first runs Newton iterations J dr = -F on the trilinear
 * field from the continuation start (step 0: from a given estimate or the
 * plain start), each step cut to at most `cap` and taken only where J is
 * stable; a full step shorter than `tol` means converged and FIRE
 * (relax_probe) starts there, otherwise it starts from the continuation
 * start. */

/* Force (as pp_force) and its Jacobian dF/dr at r: derivatives of the
 * trilinear weights of the cell holding r (wrap_axis conventions), mapped by
 * dinv, plus the radial spring's krad[(1 - L/l) I + (L/l^3) d d^T] and the
 * lateral springs' diag(kspring). */
static void force_jac(const orc_field *ff, const orc_tip *tip, v3 rtip, v3 r, double f[3], double J[3][3])
{
    const double g[3] = {v3_dot(r, (v3){ff->dinv[0], ff->dinv[1], ff->dinv[2]}),
                         v3_dot(r, (v3){ff->dinv[3], ff->dinv[4], ff->dinv[5]}),
                         v3_dot(r, (v3){ff->dinv[6], ff->dinv[7], ff->dinv[8]})};
    int lo[3], hi[3];
    double wl[3], wh[3];
    for (int a = 0; a < 3; a++) wrap_axis(g[a], ff->n[a], &lo[a], &hi[a], &wl[a], &wh[a]);
    const long ny = ff->n[1], nz = ff->n[2];
    double dFg[3][3] = {{0}};  /* dF_i / dg_a */
    for (int i = 0; i < 3; i++) f[i] = 0;
    for (int bx = 0; bx < 2; bx++)
        for (int by = 0; by < 2; by++)
            for (int bz = 0; bz < 2; bz++) {
                const int ix = bx ? hi[0] : lo[0], iy = by ? hi[1] : lo[1], iz = bz ? hi[2] : lo[2];
                const double *c = ff->F + 3 * (((long)ix * ny + iy) * nz + iz);
                const double w0 = bx ? wh[0] : wl[0], w1 = by ? wh[1] : wl[1], w2 = bz ? wh[2] : wl[2];
                const double s0 = bx ? 1 : -1, s1 = by ? 1 : -1, s2 = bz ? 1 : -1;
                for (int i = 0; i < 3; i++) {
                    f[i] += c[i] * w0 * w1 * w2;
                    dFg[i][0] += c[i] * s0 * w1 * w2;
                    dFg[i][1] += c[i] * w0 * s1 * w2;
                    dFg[i][2] += c[i] * w0 * w1 * s2;
                }
            }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            J[i][j] = dFg[i][0] * ff->dinv[j] + dFg[i][1] * ff->dinv[3 + j] + dFg[i][2] * ff->dinv[6 + j];
    const double d[3] = {r.x - rtip.x, r.y - rtip.y, r.z - rtip.z};
    const double l = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double a1 = tip->krad * (1 - tip->lrad / l), a2 = tip->krad * tip->lrad / (l * l * l);
    for (int i = 0; i < 3; i++) {
        f[i] += d[i] * a1 + (d[i] - tip->rpp0[i]) * tip->kspring[i];
        for (int j = 0; j < 3; j++) J[i][j] += a2 * d[i] * d[j] + (i == j ? a1 + tip->kspring[i] : 0);
    }
}

/* 1 when every eigenvalue of J has a negative real part (Routh-Hurwitz on
 * det(lambda I - J) = lambda^3 + c2 lambda^2 + c1 lambda + c0). */
int orc_stable3(const double *Jf)
{
    const double (*J)[3] = (const double (*)[3])Jf;
    const double tr = J[0][0] + J[1][1] + J[2][2];
    const double m2 = J[0][0] * J[1][1] - J[0][1] * J[1][0] + J[0][0] * J[2][2] - J[0][2] * J[2][0] +
                      J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    const double c2 = -tr, c1 = m2, c0 = -det;
    return c2 > 0 && c0 > 0 && c2 * c1 > c0;
}

/* Solve J x = b (Gaussian elimination, partial pivoting); 0 if singular. */
static int solve3(double J[3][3], const double b[3], double x[3])
{
    double A[3][4];
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) A[i][j] = J[i][j];
        A[i][3] = b[i];
    }
    for (int c = 0; c < 3; c++) {
        int p = c;
        for (int i = c + 1; i < 3; i++)
            if (fabs(A[i][c]) > fabs(A[p][c])) p = i;
        if (!(fabs(A[p][c]) > 1e-300)) return 0;
        for (int j = 0; j < 4; j++) { double t = A[c][j]; A[c][j] = A[p][j]; A[p][j] = t; }
        for (int i = c + 1; i < 3; i++) {
            const double m = A[i][c] / A[c][c];
            for (int j = c; j < 4; j++) A[i][j] -= m * A[c][j];
        }
    }
    for (int i = 2; i >= 0; i--) {
        double s = A[i][3];
        for (int j = i + 1; j < 3; j++) s -= A[i][j] * x[j];
        x[i] = s / A[i][i];
    }
    return 1;
}

/* One Newton iteration from *r, the step cut to at most cap (damped Newton):
 * returns 1 and moves r when the step is finite; *st2 = squared length of
 * the full step; *stable = stability of J at the old r; *f2 = squared total
 * force at the old r. */
static int newton_iter(const orc_field *ff, const orc_tip *tip, v3 rtip, v3 *r, double cap, int *stable,
                       double *st2, double *f2)
{
    double f[3], J[3][3], dr[3];
    force_jac(ff, tip, rtip, *r, f, J);
    *f2 = f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
    *stable = orc_stable3(&J[0][0]);
    const double b[3] = {-f[0], -f[1], -f[2]};
    if (!solve3(J, b, dr)) return 0;
    *st2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
    if (!(*st2 < 1e6)) return 0;
    const double sc = *st2 > cap * cap ? cap / sqrt(*st2) : 1.0;
    r->x += dr[0] * sc; r->y += dr[1] * sc; r->z += dr[2] * sc;
    return 1;
}

/* Force and Jacobian at npts points (tests: finite-difference pin). */
void orc_force_jac(const int n[3], const double cell[9], const double *F, const double tip[8], int npts,
                   const double *rtips, const double *rs, double *f_out, double *J_out)
{
    orc_field ff; orc_tip t;
    make_field(&ff, n, cell, F); make_tip(&t, tip);
    for (int i = 0; i < npts; i++) {
        double J[3][3];
        force_jac(&ff, &t, (v3){rtips[3 * i], rtips[3 * i + 1], rtips[3 * i + 2]},
                  (v3){rs[3 * i], rs[3 * i + 1], rs[3 * i + 2]}, f_out + 3 * i, J);
        for (int a = 0; a < 9; a++) J_out[9 * i + a] = J[a / 3][a % 3];
    }
}

/* Newton iterations from *q (at most kmax): 2 when an iterate with a
 * stable Jacobian already meets the stop rule (|F|^2 < conv_f2; *q is that
 * iterate, the step ends there), 1 when they converge (full step < tol,
 * every iterate stable; *q the converged point), else 0; counts
 * evaluations. */
static int newton_run(const orc_field *ff, const orc_tip *tip, v3 rtip, v3 *q, int kmax, double tol, double cap,
                      double conv_f2, int *ne)
{
    for (int k = 0; k < kmax; k++) {
        int stable = 0;
        double st2 = 0, f2 = 0;
        const v3 q0 = *q;
        const int moved = newton_iter(ff, tip, rtip, q, cap, &stable, &st2, &f2);
        (*ne)++;
        if (stable && f2 < conv_f2) {
            *q = q0;
            return 2;
        }
        if (moved && stable && st2 < tol * tol) return 1;
        if (!(moved && stable)) return 0;
    }
    return 0;
}

/* FIRE (relaxProbe, ProbeParticle.cpp:297-313, FIREstate::move 142-174)
 * from *r with a Newton try every `retry` iterations: one Newton step from
 * the current position; where it is stable and finite, Newton iterations
 * continue (newton_max in all, this step included) and on convergence FIRE
 * restarts at the zero with v = 0 and fresh dt, alpha, its iteration count
 * running on; when a later Newton iterate meets the stop rule the step ends
 * there.  `check`: *r is a Newton zero; when the first evaluation there
 * already meets the stop rule the step ends at *r without a move (returns
 * 0), as after any later convergence of Newton.  Returns the iteration index
 * at which FIRE stopped. */
static int fire_retry(const orc_field *ff, const orc_tip *tip, const orc_relax *rp, v3 rtip, v3 *r, int check,
                      int retry, int newton_max, double tol, double cap, int *ne)
{
    v3 v = {0, 0, 0};
    const double dtmax = rp->dt, a0 = rp->damping;
    double dt = dtmax, acoef = a0;
    int it = 0, next_try = retry;
    while (it < rp->max_iters) {
        v3 f = pp_force(ff, tip, rtip, *r);
        (*ne)++;
        if (check) {
            check = 0;
            if (v3_norm2(f) < rp->conv_f2) return it;
        }
        if (retry > 0 && it >= next_try) {
            next_try = it + retry;
            v3 q = *r;
            int stable = 0;
            double st2 = 0, f2 = 0;
            const int moved = newton_iter(ff, tip, rtip, &q, cap, &stable, &st2, &f2);
            if (moved && stable) {
                int conv = st2 < tol * tol;
                if (!conv) conv = newton_run(ff, tip, rtip, &q, newton_max - 1, tol, cap, rp->conv_f2, ne);
                if (conv == 2) {
                    *r = q;
                    return it;
                }
                if (conv) {
                    *r = q;
                    v = (v3){0, 0, 0};
                    dt = dtmax;
                    acoef = a0;
                    check = 1;
                }
                continue;
            }
        }
        double ff2 = v3_norm2(f), vv = v3_norm2(v), vf = v3_dot(f, v);
        if (vf < 0) {
            v.x = 0; v.y = 0; v.z = 0;
            dt = dt * rp->fdec;
            acoef = a0;
        } else {
            double cf = acoef * sqrt(vv / ff2);
            double cv = 1 - acoef;
            v.x *= cv; v.y *= cv; v.z *= cv;
            v.x += f.x * cf; v.y += f.y * cf; v.z += f.z * cf;
            dt = fmin(dt * rp->finc, dtmax);
            acoef = acoef * rp->falpha;
        }
        v.x += f.x * dt; v.y += f.y * dt; v.z += f.z * dt;
        r->x += v.x * dt; r->y += v.y * dt; r->z += v.z * dt;
        it++;
        if (ff2 < rp->conv_f2) return it - 1;
    }
    return rp->max_iters;
}

/* One stroke with warm starts (probeStart = 1 continuation, FIRE alg 1).
 * start0 (optional) = step-0 estimate.  nevals[i] = Newton + FIRE
 * evaluations of step i (the final sample not included); returns the
 * number of steps whose opening Newton run fell back to the continuation
 * start. */
static int warm_stroke(const orc_field *ff, const orc_tip *tip, const orc_relax *rp, int nstep,
                       const double *rtips, const double *start0, int newton_max, int newton_max0, int retry,
                       double tol, double cap, double extrap, double res_max, double *rs, double *fs, int *iters,
                       int *nevals)
{
    v3 rtip = {rtips[0], rtips[1], rtips[2]};
    v3 rp_ = {rtip.x + tip->rpp0[0], rtip.y + tip->rpp0[1], rtip.z + tip->rpp0[2]};
    v3 u1 = {0, 0, 0};  /* probe offset from the tip two steps back (predictor) */
    int nfb = 0;
    for (int i = 0; i < nstep; i++) {
        v3 nt = {rtips[3 * i], rtips[3 * i + 1], rtips[3 * i + 2]};
        v3 drp = {rp_.x - rtip.x, rp_.y - rtip.y, rp_.z - rtip.z};
        rtip = nt;
        rp_.x = rtip.x + drp.x; rp_.y = rtip.y + drp.y; rp_.z = rtip.z + drp.z;
        int ne = 0, check = 0, settled = 0;
        const int kmax = i == 0 ? newton_max0 : newton_max;
        v3 q = rp_;
        if (i == 0 && start0) q = (v3){start0[0], start0[1], start0[2]};
        /* second-order predictor: the offset drp of the last step keeps its
           last change drp - u1 when that is at most extrap long */
        if (extrap > 0 && i >= 2 && kmax > 0) {
            const v3 e = {drp.x - u1.x, drp.y - u1.y, drp.z - u1.z};
            if (e.x * e.x + e.y * e.y + e.z * e.z <= extrap * extrap) { q.x += e.x; q.y += e.y; q.z += e.z; }
        }
        if (i >= 1) u1 = drp;
        /* residual gate: where the total force at the step-0 estimate is
           larger than res_max (or not finite), step 0 is plain FIRE from the
           continuation start (one evaluation at the estimate, a fallback) */
        int gated = 0;
        if (i == 0 && start0 && res_max > 0 && kmax > 0) {
            const v3 f0 = pp_force(ff, tip, rtip, q);
            ne++;
            gated = !(v3_norm2(f0) <= res_max * res_max);
            if (!gated) ne--;  /* the Newton run's first evaluation is this one */
        }
        const int res = gated ? 0 : kmax == 0 ? -1 : newton_run(ff, tip, rtip, &q, kmax, tol, cap, rp->conv_f2, &ne);
        if (res == -1) {
            rp_ = q;
        } else if (res == 2) {
            rp_ = q;
            settled = 1;
        } else if (res == 1) {
            rp_ = q;
            check = 1;
        } else {
            nfb++;
        }
        const int itr = settled ? 0 : fire_retry(ff, tip, rp, rtip, &rp_, check, retry, newton_max, tol, cap, &ne);
        v3 g = {v3_dot(rp_, (v3){ff->dinv[0], ff->dinv[1], ff->dinv[2]}),
                v3_dot(rp_, (v3){ff->dinv[3], ff->dinv[4], ff->dinv[5]}),
                v3_dot(rp_, (v3){ff->dinv[6], ff->dinv[7], ff->dinv[8]})};
        v3 fsamp = interp_wrap(ff->F, ff->n, g);
        rs[3 * i] = rp_.x; rs[3 * i + 1] = rp_.y; rs[3 * i + 2] = rp_.z;
        fs[3 * i] = fsamp.x; fs[3 * i + 1] = fsamp.y; fs[3 * i + 2] = fsamp.z;
        if (iters) iters[i] = itr;
        if (nevals) nevals[i] = ne;
    }
    return nfb;
}

/* Warm-start strokes over (nx,ny,nz,3) C-order arrays, column offsets as
 * orc_relax_strokes; starts0 (optional) = (nx*ny, 3) step-0 estimates in the
 * same column order.  Returns the total number of fallback steps. */
long orc_relax_strokes_warm(int nx, int ny, int nz, const int n[3], const double cell[9], const double *F,
                            const double tip[8], const double relax[7], const double *rtips, const double *starts0,
                            int newton_max, int newton_max0, int retry, double tol, double cap, double extrap,
                            double res_max, double *rs, double *fs, int *iters, int *nevals)
{
    orc_field ff; orc_tip t; orc_relax r;
    make_field(&ff, n, cell, F); make_tip(&t, tip); make_relax(&r, relax);
    long total = 0;
    #pragma omp parallel for collapse(2) schedule(dynamic, 16) reduction(+ : total)
    for (int ix = 0; ix < nx; ix++)
        for (int iy = 0; iy < ny; iy++) {
            const long c = (long)ix + (long)iy * nx, off = c * nz;
            total += warm_stroke(&ff, &t, &r, nz, rtips + 3 * off, starts0 ? starts0 + 3 * c : 0, newton_max,
                                 newton_max0, retry, tol, cap, extrap, res_max, rs + 3 * off, fs + 3 * off,
                                 iters ? iters + off : 0, nevals ? nevals + off : 0);
        }
    return total;
}

/* ----------------------------------------------------- further pair models */

/* Pair terms of Forces.h:108-232 and ProbeParticle.cpp:227-256, each in the
 * reference's operation order.  model: 4 vdW with constant damping (c6;
 * par = ADamp), 5..8 vdW with damping function R2/R4/invR4/invR8 (R, E0;
 * par = ADamp), 9 DFT-D3 (c6, c8, R0^6, R0^8), 10 Morse (R0, E0; par =
 * alpha), 11/12/13 Gauss/Slater/R4 densities (amp, width; energy only),
 * 14 LJ in (R0, E0), 15 bare r^-6 (c6).  The vdW of the OpenCL force
 * field (addvdW, FF.cl:548-665) writes R2SAFE as 1e-6f: 16 no damping (c6),
 * 17/18 invR4/invR8 damping with that constant (models 5/6 serve its R2/R4). */
#define ORC_R2SAFE_OCL ((double)1.0e-6f)
static inline double damp_fn(int kind, double u2, double *dd, double r2safe)
{
    if (kind == 1) {
        if (u2 > 1) { *dd = 0; return 0; }
        *dd = -2;
        return 1 - u2;
    }
    if (kind == 2) {
        if (u2 > 1) { *dd = 0; return 0; }
        double e = 1 - u2;
        *dd = -4 * e;
        return e * e;
    }
    double ir2 = 1 / (u2 + r2safe);
    if (kind == 3) {
        double ir4 = ir2 * ir2;
        *dd = -4 * ir4 * ir2;
        return ir4;
    }
    double ir8 = ir2 * ir2;
    ir8 *= ir8;
    *dd = -8 * ir8 * ir2;
    return ir8;
}

static double pair_model(int model, v3 d, v3 *f, const double *c, double par)
{
    double r2 = v3_norm2(d), s, e;
    switch (model) {
    case 4:
        r2 *= r2;
        r2 *= r2;
        s = -6 * c[0] / (r2 + par * c[0]);
        f->x += d.x * s; f->y += d.y * s; f->z += d.z * s;
        return 0;
    case 5: case 6: case 7: case 8: case 17: case 18: {
        double ir2 = 1. / (c[0] * c[0]), u2 = r2 * ir2, u4 = u2 * u2, dd;
        double D = model < 17 ? damp_fn(model - 4, u2, &dd, ORC_R2SAFE) : damp_fn(model - 14, u2, &dd, ORC_R2SAFE_OCL);
        double ee = 1. / (u4 * u2 + D * par);
        e = -2 * c[1] * ee;
        s = e * ee * (6 * u4 + dd * par) * ir2;
        f->x += d.x * s; f->y += d.y * s; f->z += d.z * s;
        return e;
    }
    case 9: {
        if (r2 > 400.0) return 0;
        double r4 = r2 * r2, r6 = r4 * r2, r8 = r6 * r2;
        double d6 = 1.0 / (r6 + c[2]), d8 = 1.0 / (r8 + c[3]);
        double e6 = -c[0] * d6, e8 = -c[1] * d8;
        s = e6 * d6 * 6 * r4 + e8 * d8 * 8 * r6;
        f->x += d.x * s; f->y += d.y * s; f->z += d.z * s;
        return e6 + e8;
    }
    case 10: {
        double r = sqrt(r2 + ORC_R2SAFE);
        double ex = exp(par * (r - c[0]));
        e = c[1] * (ex * ex - 2 * ex);
        s = c[1] * 2 * par * (ex - ex * ex) / r;
        f->x += d.x * s; f->y += d.y * s; f->z += d.z * s;
        return e;
    }
    case 11: {
        double t = r2 * (1 / c[1]);
        return c[0] * exp(-0.5 * (t * t));
    }
    case 12:
        return c[0] * exp(-(1 / c[1]) * sqrt(r2));
    case 13: {
        double s2 = c[1] * c[1];
        if (r2 > s2) return 0;
        r2 /= s2;
        double rf = 1 - r2;
        return c[0] * rf * rf;
    }
    case 14: {
        double ir2 = 1 / (r2 + ORC_R2SAFE), u2 = c[0] * c[0] * ir2, u6 = u2 * u2 * u2;
        double e6 = c[1] * u6, e12 = e6 * u6;
        s = 12 * (e12 - e6) * ir2;
        f->x += d.x * s; f->y += d.y * s; f->z += d.z * s;
        return e12 - 2 * e6;
    }
    case 15: {
        double ir2 = 1 / r2, ir4 = ir2 * ir2;
        e = -c[0] * ir4 * ir2;
        s = e * 6 * ir2;
        f->x += d.x * s; f->y += d.y * s; f->z += d.z * s;
        return e;
    }
    case 16: {
        double ir2 = 1.0 / (r2 + ORC_R2SAFE_OCL), ir6 = ir2 * ir2 * ir2;
        e = -c[0] * ir6;
        s = 6.0 * e * ir2;
        f->x += d.x * s; f->y += d.y * s; f->z += d.z * s;
        return e;
    }
    }
    return 0;
}

/* evalCell<model> over the grid (ProbeParticle.cpp:262-278, 515-631):
 * += into F (unless a density) and E where non-NULL; coefs (natoms,ncoef). */
void orc_model_ff(int model, double par, const int n[3], const double cell[9], int natoms, const double *Rs,
                  const double *coefs, int ncoef, double *F, double *E)
{
    double dcell[9], dinv[9];
    orc_grid_shape(n, cell, dcell, dinv);
    const int nx = n[0], ny = n[1], nz = n[2];
    const int dens = model >= 11 && model <= 13;
    #pragma omp parallel for collapse(3) schedule(static)
    for (int ia = 0; ia < nx; ia++)
        for (int ib = 0; ib < ny; ib++)
            for (int ic = 0; ic < nz; ic++) {
                v3 p;
                p.x = dcell[0] * ia + dcell[3] * ib + dcell[6] * ic;
                p.y = dcell[1] * ia + dcell[4] * ib + dcell[7] * ic;
                p.z = dcell[2] * ia + dcell[5] * ib + dcell[8] * ic;
                v3 f = {0, 0, 0};
                double e = 0;
                for (int i = 0; i < natoms; i++) {
                    v3 d = {p.x - Rs[3 * i], p.y - Rs[3 * i + 1], p.z - Rs[3 * i + 2]};
                    e += pair_model(model, d, &f, coefs + (long)i * ncoef, par);
                }
                long ig = ((long)ia * ny + ib) * nz + ic;
                if (F && !dens) { F[3 * ig] += f.x; F[3 * ig + 1] += f.y; F[3 * ig + 2] += f.z; }
                if (E) E[ig] += e;
            }
}

/* evalRadialFF (ProbeParticle.cpp:556-575) for one model at separations
 * (r,0,0): Es, Fs (x force). */
void orc_radial(int model, double par, int n, const double *rs, const double *coefs, double *Es, double *Fs)
{
    for (int i = 0; i < n; i++) {
        v3 f = {0, 0, 0};
        v3 d = {rs[i], 0, 0};
        double e = model == 0 ? pair_lj(d, &f, coefs[0], coefs[1]) : pair_model(model, d, &f, coefs, par);
        Es[i] = e;
        Fs[i] = f.x;
    }
}

/* ------------------------------------------------------------- stiffness */

/* Mat3::eigenvals (Mat3.h:459-482) of a symmetric matrix m (row-major 3x3):
 * the characteristic cubic's coefficients scaled by the largest element,
 * then its three real roots in trigonometric form.  The reference writes
 * 1/3 and sqrt(3) as the decimals 0.33333333333 and 1.73205080757; they are
 * kept so that this rounds like the compiled reference.  Unsorted. */
static void ref_eigenvals(const double m[9], double ev[3])
{
    const double third = 0.33333333333, sqrt3 = 1.73205080757;
    double amax = m[0];
    for (int k = 1; k < 9; k++)
        if (m[k] > amax) amax = m[k];
    const double xx = m[0], xy = m[1], xz = m[2], yy = m[4], yz = m[5], zz = m[8];
    double det = xx * yy * zz + 2 * xy * xz * yz - xx * yz * yz - yy * xz * xz - zz * xy * xy;
    double minors = xx * yy - xy * xy + xx * zz - xz * xz + yy * zz - yz * yz;
    double tr = xx + yy + zz;
    const double amax2 = amax * amax;
    tr /= amax;
    minors /= amax2;
    det /= (amax2 * amax);
    const double t3 = tr * third;
    double p3 = (minors - tr * t3) * third;
    if (p3 > 0.0) p3 = 0.0;
    const double h = 0.5 * (det + t3 * (2.0 * t3 * t3 - minors));
    double disc = h * h + p3 * p3 * p3;
    if (disc > 0.0) disc = 0.0;
    const double mag = sqrt(-p3);
    const double ang = atan2(sqrt(-disc), h) * third;
    const double cs = cos(ang), sn = sin(ang);
    ev[0] = amax * (t3 + 2.0 * mag * cs);
    ev[1] = amax * (t3 - mag * (cs + sqrt3 * sn));
    ev[2] = amax * (t3 - mag * (cs - sqrt3 * sn));
}

/* Mat3::eigenvec (Mat3.h:484-500): the longest of the three cross products
 * of the rows of (m - lam I), normalised (NaN when all three vanish, as in
 * the reference). */
static void ref_eigenvec(const double m[9], double lam, double out[3])
{
    const v3 r0 = {m[0] - lam, m[1], m[2]}, r1 = {m[3], m[4] - lam, m[5]}, r2 = {m[6], m[7], m[8] - lam};
    const v3 c[3] = {{r0.y * r1.z - r0.z * r1.y, r0.z * r1.x - r0.x * r1.z, r0.x * r1.y - r0.y * r1.x},
                     {r0.y * r2.z - r0.z * r2.y, r0.z * r2.x - r0.x * r2.z, r0.x * r2.y - r0.y * r2.x},
                     {r1.y * r2.z - r1.z * r2.y, r1.z * r2.x - r1.x * r2.z, r1.x * r2.y - r1.y * r2.x}};
    const double d[3] = {v3_dot(c[0], c[0]), v3_dot(c[1], c[1]), v3_dot(c[2], c[2])};
    int k = 0;
    double dm = d[0];
    if (d[1] > dm) { dm = d[1]; k = 1; }
    if (d[2] > dm) k = 2;
    const double s = 1 / sqrt(d[k]);
    out[0] = c[k].x * s; out[1] = c[k].y * s; out[2] = c[k].z * s;
}

/* stiffnessMatrix (ProbeParticle.cpp:703-750): for each point, rows of the
 * dynamical matrix from central differences of getPPforce under +-ddisp
 * displacements along x, y, z (applied to and undone on r in place, as the
 * reference does), symmetrised; eigenvalues sorted ascending into evals
 * (n,3); for which >= 1/2/3 the eigenvectors of the smallest/middle/largest
 * into ev1/ev2/ev3 (n,3).  dynmat (n,9), optional, receives the symmetric
 * matrix.  Returns the number of points with a NaN eigenvalue (where the
 * reference exits). */
long orc_stiffness_spline(const int n[3], const double cell[9], const double *F, const double tip[8], int sn,
                          const double *sxs, const double *sydys, double ddisp, int which, long np,
                          const double *rtips, const double *rpps, double *evals, double *ev1, double *ev2,
                          double *ev3, double *dynmat)
{
    orc_field ff; orc_tip t;
    make_field(&ff, n, cell, F); make_tip(&t, tip);
    t.sn = sn; t.sxs = sxs; t.sydys = sydys;
    long nnan = 0;
    #pragma omp parallel for schedule(static) reduction(+ : nnan)
    for (long i = 0; i < np; i++) {
        const v3 rtip = {rtips[3 * i], rtips[3 * i + 1], rtips[3 * i + 2]};
        v3 r = {rpps[3 * i], rpps[3 * i + 1], rpps[3 * i + 2]};
        double m[9];
        v3 f1, f2;
        const double sc = -0.5 / ddisp;
        r.x -= ddisp; f1 = pp_force(&ff, &t, rtip, r); r.x += 2 * ddisp; f2 = pp_force(&ff, &t, rtip, r); r.x -= ddisp;
        m[0] = (f2.x - f1.x) * sc; m[1] = (f2.y - f1.y) * sc; m[2] = (f2.z - f1.z) * sc;
        r.y -= ddisp; f1 = pp_force(&ff, &t, rtip, r); r.y += 2 * ddisp; f2 = pp_force(&ff, &t, rtip, r); r.y -= ddisp;
        m[3] = (f2.x - f1.x) * sc; m[4] = (f2.y - f1.y) * sc; m[5] = (f2.z - f1.z) * sc;
        r.z -= ddisp; f1 = pp_force(&ff, &t, rtip, r); r.z += 2 * ddisp; f2 = pp_force(&ff, &t, rtip, r); r.z -= ddisp;
        m[6] = (f2.x - f1.x) * sc; m[7] = (f2.y - f1.y) * sc; m[8] = (f2.z - f1.z) * sc;
        double s;
        s = 0.5 * (m[1] + m[3]); m[1] = s; m[3] = s;
        s = 0.5 * (m[5] + m[7]); m[5] = s; m[7] = s;
        s = 0.5 * (m[6] + m[2]); m[6] = s; m[2] = s;
        double e[3];
        ref_eigenvals(m, e);
        if (isnan(e[0] * e[1] * e[2])) nnan++;
        if (e[0] > e[1]) { s = e[0]; e[0] = e[1]; e[1] = s; }
        if (e[1] > e[2]) { s = e[1]; e[1] = e[2]; e[2] = s; }
        if (e[0] > e[1]) { s = e[0]; e[0] = e[1]; e[1] = s; }
        evals[3 * i] = e[0]; evals[3 * i + 1] = e[1]; evals[3 * i + 2] = e[2];
        if (which > 0) ref_eigenvec(m, e[0], ev1 + 3 * i);
        if (which > 1) ref_eigenvec(m, e[1], ev2 + 3 * i);
        if (which > 2) ref_eigenvec(m, e[2], ev3 + 3 * i);
        if (dynmat) memcpy(dynmat + 9 * i, m, sizeof m);
    }
    return nnan;
}

long orc_stiffness(const int n[3], const double cell[9], const double *F, const double tip[8], double ddisp,
                   int which, long np, const double *rtips, const double *rpps, double *evals, double *ev1,
                   double *ev2, double *ev3, double *dynmat)
{
    return orc_stiffness_spline(n, cell, F, tip, 0, 0, 0, ddisp, which, np, rtips, rpps, evals, ev1, ev2, ev3,
                                dynmat);
}

/* ------------------------------------------------------- OpenCL-mode grid */

/* evalLJC_QZs_noPos (FF.cl:366-423) restated in FP64: LJ (getLJ, FF.cl:123-130)
 * plus four z-offset point charges of the tip (getCoulomb, FF.cl:115-121),
 * R2SAFE = 1e-6, Coulomb constant 14.399644, no clamp.  Output FE is
 * (nx*ny*nz, 4) with x fastest: index ia + nx*(ib + ny*ic).  atoms is
 * (N,4) xyzq in the grid frame *before* subtracting the origin p0. */
void orc_ljc_qzs_ff(const int n[3], const double p0[3], const double dA[3], const double dB[3],
                    const double dC[3], int natoms, const double *atoms, const double *cLJs,
                    const double Qs[4], const double QZs[4], double *FE)
{
    const int nx = n[0], ny = n[1], nz = n[2];
    const double r2safe = (double)1e-6f, kc = (double)14.399644f;
    double q[4];
    for (int k = 0; k < 4; k++) q[k] = Qs[k] * kc;
    #pragma omp parallel for collapse(3) schedule(static)
    for (int ic = 0; ic < nz; ic++)
        for (int ib = 0; ib < ny; ib++)
            for (int ia = 0; ia < nx; ia++) {
                double px = p0[0] + dA[0] * ia + dB[0] * ib + dC[0] * ic;
                double py = p0[1] + dA[1] * ia + dB[1] * ib + dC[1] * ic;
                double pz = p0[2] + dA[2] * ia + dB[2] * ib + dC[2] * ic;
                double fx = 0, fy = 0, fz = 0, e = 0;
                for (int i = 0; i < natoms; i++) {
                    const double *a = atoms + 4 * i;
                    double dx = px - a[0], dy = py - a[1], dz = pz - a[2];
                    double ir2 = 1.0 / (dx * dx + dy * dy + dz * dz + r2safe);
                    double ir6 = ir2 * ir2 * ir2;
                    double c6 = cLJs[2 * i], c12 = cLJs[2 * i + 1];
                    e += (c12 * ir6 - c6) * ir6;
                    double s = (12.0 * c12 * ir6 - 6.0 * c6) * ir6 * ir2;
                    fx += s * dx; fy += s * dy; fz += s * dz;
                    for (int k = 0; k < 4; k++) {
                        if (q[k] == 0) continue;
                        double dzk = dz + QZs[k];
                        double jr2 = 1.0 / (dx * dx + dy * dy + dzk * dzk + r2safe);
                        double ek = a[3] * sqrt(jr2);
                        double sk = ek * jr2 * q[k];
                        fx += sk * dx; fy += sk * dy; fz += sk * dzk;
                        e += ek * q[k];
                    }
                }
                long ig = (long)ia + (long)nx * ((long)ib + (long)ny * ic);
                FE[4 * ig] = fx; FE[4 * ig + 1] = fy; FE[4 * ig + 2] = fz; FE[4 * ig + 3] = e;
            }
}

/* -------------------------------------------------- OpenCL-mode relaxation */

/* relaxStrokesTilted (relax.cl:391-464) restated in FP64 with exact software
 * periodic trilinear interpolation in place of the hardware image filter
 * (interpFE, relax.cl:64-73: normalized coordinates + half texel with
 * ADDRESS_REPEAT and LINEAR filtering samples grid points floor(u) and
 * floor(u)+1 of u = pos.dinv, i.e. the same corners and weights as
 * Grid.h:66-111, but the hardware rounds the weights).
 *
 * FE grid: (nx,ny,nz,4) with x fastest (evalLJC_QZs_noPos layout).
 * poss: (npix,4) tip anchor positions (relative to the grid origin).
 * tipRot rows A,B,C (3x4; C[3] = -zstep), stiffness (4), dpos0 (4, tip frame),
 * relax_params (dt, damp), surfFF (x used), maxIters, convF2, F2SAFE.
 * Outputs FEs and paths (npix*nz, 4).  nevals (npix*nz, optional). */
/* Periodic cell of grid coordinate g along an axis of n points: lower
 * node lo, upper node hi, fractional position t.  The remainder is taken in
 * floating point (exact for every finite value), so a probe that has run far
 * away still indexes inside the grid; a NaN coordinate maps to node 0. */
static inline void xfast_axis(double g, int n, int *lo, int *hi, double *t)
{
    const double fl = floor(g);
    *t = g - fl;
    double m = fmod(fl, (double)n);
    if (m < 0) m += n;
    const int il = (m >= 0 && m < n) ? (int)m : 0;
    *lo = il;
    *hi = il + 1 == n ? 0 : il + 1;
}

static inline void fe_interp_xfast(const double *FE, const int n[3], double gx, double gy, double gz,
                                   double out[4])
{
    int i0[3], i1[3];
    double w0[3], w1[3];
    double g[3] = {gx, gy, gz};
    for (int a = 0; a < 3; a++) {
        double t;
        xfast_axis(g[a], n[a], &i0[a], &i1[a], &t);
        w0[a] = 1 - t; w1[a] = t;
    }
    for (int c = 0; c < 4; c++) out[c] = 0;
    for (int k = 0; k < 8; k++) {
        int ix = (k & 1) ? i1[0] : i0[0];
        int iy = (k & 2) ? i1[1] : i0[1];
        int iz = (k & 4) ? i1[2] : i0[2];
        double w = ((k & 1) ? w1[0] : w0[0]) * ((k & 2) ? w1[1] : w0[1]) * ((k & 4) ? w1[2] : w0[2]);
        const double *p = FE + 4 * ((long)ix + (long)n[0] * ((long)iy + (long)n[1] * iz));
        for (int c = 0; c < 4; c++) out[c] += w * p[c];
    }
}

void orc_relax_strokes_tilted(int npix, int nz, const int n[3], const double *FE, const double dinv[9],
                              const double tipRot[12], const double stiffness[4], const double dpos0[4],
                              const double relax_params[2], double surf_ff, int max_iters,
                              double conv_f2, double f2safe, const double *poss, double *FEs,
                              double *paths, int *nevals)
{
    #pragma omp parallel for schedule(dynamic, 16)
    for (int ip = 0; ip < npix; ip++) {
        const double *A = tipRot, *B = tipRot + 4, *C = tipRot + 8;
        double dtip[3] = {C[0] * C[3], C[1] * C[3], C[2] * C[3]};
        /* dpos0_ = rotMatT(dpos0) = A*x + B*y + C*z */
        double d0[3];
        for (int c = 0; c < 3; c++) d0[c] = A[c] * dpos0[0] + B[c] * dpos0[1] + C[c] * dpos0[2];
        double tp[3] = {poss[4 * ip], poss[4 * ip + 1], poss[4 * ip + 2]};
        double pos[3] = {tp[0] + d0[0], tp[1] + d0[1], tp[2] + d0[2]};
        double dt = relax_params[0], damp = relax_params[1];
        const double dtmax = dt, dtmin = dtmax * 0.1, damp0 = damp;
        for (int iz = 0; iz < nz; iz++) {
            double fe[4] = {0, 0, 0, 0};
            double v[3] = {0, 0, 0};
            int ne = 0;
            for (int i = 0; i < max_iters; i++) {
                double gx = pos[0] * dinv[0] + pos[1] * dinv[1] + pos[2] * dinv[2];
                double gy = pos[0] * dinv[3] + pos[1] * dinv[4] + pos[2] * dinv[5];
                double gz = pos[0] * dinv[6] + pos[1] * dinv[7] + pos[2] * dinv[8];
                fe_interp_xfast(FE, n, gx, gy, gz, fe);
                ne++;
                double f[3] = {fe[0], fe[1], fe[2]};
                double dp[3] = {pos[0] - tp[0], pos[1] - tp[1], pos[2] - tp[2]};
                double dq[3] = {dp[0] * A[0] + dp[1] * A[1] + dp[2] * A[2],
                                dp[0] * B[0] + dp[1] * B[1] + dp[2] * B[2],
                                dp[0] * C[0] + dp[1] * C[1] + dp[2] * C[2]};
                double rl = sqrt(dq[0] * dq[0] + dq[1] * dq[1] + dq[2] * dq[2]);
                double sr = stiffness[3] * (rl - dpos0[3]) / rl;
                double ft[3];
                for (int c = 0; c < 3; c++) ft[c] = (dq[c] - dpos0[c]) * stiffness[c] + dq[c] * sr;
                for (int c = 0; c < 3; c++) f[c] += A[c] * ft[0] + B[c] * ft[1] + C[c] * ft[2];
                for (int c = 0; c < 3; c++) f[c] += C[c] * surf_ff;
                /* update_FIRE (relax.cl:105-126) */
                double ff2 = f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
                double vv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
                double vf = v[0] * f[0] + v[1] * f[1] + v[2] * f[2];
                if (vf < 0) {
                    v[0] = v[1] = v[2] = 0;
                    dt = fmax(dtmin, dt * 0.5);
                    damp = damp0;
                } else {
                    double cf = damp * sqrt(vv / (ff2 + f2safe));
                    for (int c = 0; c < 3; c++) v[c] = v[c] * (1 - damp) + f[c] * cf;
                    dt = fmin(dtmax, dt * 1.1);
                    damp *= 0.99;
                }
                for (int c = 0; c < 3; c++) { v[c] += f[c] * dt; pos[c] += v[c] * dt; }
                if (ff2 < conv_f2) break;
            }
            long ind = (long)ip * nz + iz;
            FEs[4 * ind + 0] = fe[0] * A[0] + fe[1] * A[1] + fe[2] * A[2];
            FEs[4 * ind + 1] = fe[0] * B[0] + fe[1] * B[1] + fe[2] * B[2];
            FEs[4 * ind + 2] = fe[0] * C[0] + fe[1] * C[1] + fe[2] * C[2];
            FEs[4 * ind + 3] = fe[3];
            paths[4 * ind + 0] = pos[0]; paths[4 * ind + 1] = pos[1]; paths[4 * ind + 2] = pos[2];
            paths[4 * ind + 3] = 0;
            if (nevals) nevals[ind] = ne;
            for (int c = 0; c < 3; c++) { tp[c] += dtip[c]; pos[c] += dtip[c]; }
        }
    }
}

/* ------------------------------------- warm starts, OpenCL (tilted) semantics */

/* The warm method of orc_relax_strokes_warm on relaxStrokesTilted's
 * semantics (relax.cl:391-464): the tip springs act in the tip frame rows
 * A, B, C (tipForce, relax.cl:37-41), the constant surface force along C,
 * FIRE with dt 0.5, dtmin = dtmax / 10, F2SAFE in the square root, dt and
 * alpha kept along the stroke (update_FIRE, relax.cl:105-126), at most
 * max_iters steps, the stored force the last evaluation's grid force
 * rotated into the tip frame.  Each step opens with Newton iterations from
 * the continuation start (step 0: from starts0 when given); FIRE then ends
 * the step, with a Newton try every `retry` iterations.  Field FE is
 * x-fastest (nx, ny, nz, 4). */
typedef struct {
    const double *FE;
    int n[3];
    double dinv[9];
    double R[9];     /* rows A, B, C */
    double k[3], krad, L, d0[3], surf;
} tw_field;

/* Grid force fe (4 channels), total force f and its Jacobian at r.  fe and f
 * round as in orc_relax_strokes_tilted (the same interpolation and tip-force
 * expressions), so the warm method without Newton reproduces it exactly. */
static void tw_force_jac(const tw_field *t, v3 tp, v3 r, double fe[4], double f[3], double J[3][3])
{
    const double g[3] = {r.x * t->dinv[0] + r.y * t->dinv[1] + r.z * t->dinv[2],
                         r.x * t->dinv[3] + r.y * t->dinv[4] + r.z * t->dinv[5],
                         r.x * t->dinv[6] + r.y * t->dinv[7] + r.z * t->dinv[8]};
    fe_interp_xfast(t->FE, t->n, g[0], g[1], g[2], fe);
    /* the trilinear interpolant's derivative in grid coordinates, in the cell
     * fe_interp_xfast interpolates in */
    int lo[3], hi[3];
    double wl[3], wh[3];
    for (int a = 0; a < 3; a++) {
        xfast_axis(g[a], t->n[a], &lo[a], &hi[a], &wh[a]);
        wl[a] = 1 - wh[a];
    }
    double dFg[3][3] = {{0}};
    for (int bx = 0; bx < 2; bx++)
        for (int by = 0; by < 2; by++)
            for (int bz = 0; bz < 2; bz++) {
                const int ix = bx ? hi[0] : lo[0], iy = by ? hi[1] : lo[1], iz = bz ? hi[2] : lo[2];
                const double *c = t->FE + 4 * ((long)ix + (long)t->n[0] * ((long)iy + (long)t->n[1] * iz));
                const double w0 = bx ? wh[0] : wl[0], w1 = by ? wh[1] : wl[1], w2 = bz ? wh[2] : wl[2];
                const double s0 = bx ? 1 : -1, s1 = by ? 1 : -1, s2 = bz ? 1 : -1;
                for (int i = 0; i < 3; i++) {
                    dFg[i][0] += c[i] * s0 * w1 * w2;
                    dFg[i][1] += c[i] * w0 * s1 * w2;
                    dFg[i][2] += c[i] * w0 * w1 * s2;
                }
            }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            J[i][j] = dFg[i][0] * t->dinv[j] + dFg[i][1] * t->dinv[3 + j] + dFg[i][2] * t->dinv[6 + j];
    /* tip force in the tip frame (relax.cl:37-41), as orc_relax_strokes_tilted */
    const double *A = t->R, *B = t->R + 3, *C = t->R + 6;
    const double dp[3] = {r.x - tp.x, r.y - tp.y, r.z - tp.z};
    const double dq[3] = {dp[0] * A[0] + dp[1] * A[1] + dp[2] * A[2], dp[0] * B[0] + dp[1] * B[1] + dp[2] * B[2],
                          dp[0] * C[0] + dp[1] * C[1] + dp[2] * C[2]};
    const double rl = sqrt(dq[0] * dq[0] + dq[1] * dq[1] + dq[2] * dq[2]);
    const double sr = t->krad * (rl - t->L) / rl;
    double ft[3];
    for (int c = 0; c < 3; c++) ft[c] = (dq[c] - t->d0[c]) * t->k[c] + dq[c] * sr;
    for (int c = 0; c < 3; c++) f[c] = fe[c] + (A[c] * ft[0] + B[c] * ft[1] + C[c] * ft[2]);
    for (int c = 0; c < 3; c++) f[c] += C[c] * t->surf;
    /* spring Jacobian R^T Dq R, Dq = diag(k + a1) + a2 q q^T */
    const double a1 = t->krad * (1 - t->L / rl), a2 = t->krad * t->L / (rl * rl * rl);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0;
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) {
                    const double dqab = a2 * dq[a] * dq[b] + (a == b ? a1 + t->k[a] : 0);
                    s += t->R[3 * a + i] * dqab * t->R[3 * b + j];
                }
            J[i][j] += s;
        }
}

/* Total force and Jacobian of the tilted field at npts points (tests: the
 * finite-difference pin).  Arguments as orc_relax_strokes_tilted. */
void orc_tilted_force_jac(const int n[3], const double *FE, const double dinv[9], const double tipRot[12],
                          const double stiffness[4], const double dpos0[4], double surf_ff, int npts, const double *tips,
                          const double *rs, double *f_out, double *J_out)
{
    tw_field t;
    t.FE = FE;
    for (int a = 0; a < 3; a++) t.n[a] = n[a];
    for (int a = 0; a < 9; a++) t.dinv[a] = dinv[a];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) t.R[3 * i + j] = tipRot[4 * i + j];
    for (int a = 0; a < 3; a++) { t.k[a] = stiffness[a]; t.d0[a] = dpos0[a]; }
    t.krad = stiffness[3];
    t.L = dpos0[3];
    t.surf = surf_ff;
    for (int i = 0; i < npts; i++) {
        double fe[4], J[3][3];
        tw_force_jac(&t, (v3){tips[3 * i], tips[3 * i + 1], tips[3 * i + 2]}, (v3){rs[3 * i], rs[3 * i + 1], rs[3 * i + 2]},
                     fe, f_out + 3 * i, J);
        for (int a = 0; a < 9; a++) J_out[9 * i + a] = J[a / 3][a % 3];
    }
}

/* One damped Newton iteration (see newton_iter) on the tilted field; fe of
 * the old r. */
static int tw_newton_iter(const tw_field *t, v3 tp, v3 *r, double cap, int *stable, double *st2, double *f2, double fe[4])
{
    double f[3], J[3][3], dr[3];
    tw_force_jac(t, tp, *r, fe, f, J);
    *f2 = f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
    *stable = orc_stable3(&J[0][0]);
    const double b[3] = {-f[0], -f[1], -f[2]};
    if (!solve3(J, b, dr)) return 0;
    *st2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
    if (!(*st2 < 1e6)) return 0;
    const double sc = *st2 > cap * cap ? cap / sqrt(*st2) : 1.0;
    r->x += dr[0] * sc; r->y += dr[1] * sc; r->z += dr[2] * sc;
    return 1;
}

/* newton_run on the tilted field: 2 = a stable iterate meets the stop rule
 * (*q that iterate, fe its grid force), 1 = converged, 0 = failed. */
static int tw_newton_run(const tw_field *t, v3 tp, v3 *q, int kmax, double tol, double cap, double conv_f2, int *ne,
                         double fe[4])
{
    for (int k = 0; k < kmax; k++) {
        int stable = 0;
        double st2 = 0, f2 = 0;
        const v3 q0 = *q;
        const int moved = tw_newton_iter(t, tp, q, cap, &stable, &st2, &f2, fe);
        (*ne)++;
        if (stable && f2 < conv_f2) {
            *q = q0;
            return 2;
        }
        if (moved && stable && st2 < tol * tol) return 1;
        if (!(moved && stable)) return 0;
    }
    return 0;
}

/* FIRE of relaxStrokesTilted from *r with dt, alpha carried in *dt, *damp,
 * a Newton try every `retry` iterations (as fire_retry); `check`: *r is a
 * Newton zero.  fe: grid force of the last evaluation that ends the step.
 * Returns the FIRE iterations taken. */
static int tw_fire(const tw_field *t, v3 tp, v3 *r, double *dt, double *damp, double dtmax, double dtmin, double damp0,
                   int max_iters, double conv_f2, double f2safe, int check, int retry, int newton_max, double tol,
                   double cap, int *ne, double fe[4])
{
    v3 v = {0, 0, 0};
    int it = 0, next_try = retry;
    while (it < max_iters) {
        double f[3], J[3][3];
        tw_force_jac(t, tp, *r, fe, f, J);
        (*ne)++;
        const double ff2 = f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
        if (check) {
            check = 0;
            if (ff2 < conv_f2) return it;
        }
        if (retry > 0 && it >= next_try) {
            next_try = it + retry;
            v3 q = *r;
            int stable = 0;
            double st2 = 0, f2 = 0, feq[4];
            const int moved = tw_newton_iter(t, tp, &q, cap, &stable, &st2, &f2, feq);
            if (moved && stable) {
                int conv = st2 < tol * tol;
                if (!conv) conv = tw_newton_run(t, tp, &q, newton_max - 1, tol, cap, conv_f2, ne, feq);
                if (conv == 2) {
                    *r = q;
                    for (int c = 0; c < 4; c++) fe[c] = feq[c];
                    return it;
                }
                if (conv) {
                    *r = q;
                    v = (v3){0, 0, 0};
                    check = 1;
                }
                continue;
            }
        }
        const double vv = v.x * v.x + v.y * v.y + v.z * v.z, vf = v.x * f[0] + v.y * f[1] + v.z * f[2];
        if (vf < 0) {
            v = (v3){0, 0, 0};
            *dt = fmax(dtmin, *dt * 0.5);
            *damp = damp0;
        } else {
            const double cf = *damp * sqrt(vv / (ff2 + f2safe)), cv = 1 - *damp;
            v.x = v.x * cv + f[0] * cf; v.y = v.y * cv + f[1] * cf; v.z = v.z * cv + f[2] * cf;
            *dt = fmin(dtmax, *dt * 1.1);
            *damp *= 0.99;
        }
        v.x += f[0] * *dt; v.y += f[1] * *dt; v.z += f[2] * *dt;
        r->x += v.x * *dt; r->y += v.y * *dt; r->z += v.z * *dt;
        it++;
        if (ff2 < conv_f2) return it;
    }
    return max_iters;
}

long orc_relax_strokes_tilted_warm(int npix, int nz, const int n[3], const double *FE, const double dinv[9],
                                   const double tipRot[12], const double stiffness[4], const double dpos0[4],
                                   const double relax_params[2], double surf_ff, int max_iters, double conv_f2,
                                   double f2safe, const double *poss, const double *starts0, int newton_max,
                                   int newton_max0, int retry, double tol, double cap, double extrap,
                                   double res_max, double *FEs, double *paths, int *iters, int *nevals)
{
    tw_field t;
    t.FE = FE;
    for (int a = 0; a < 3; a++) t.n[a] = n[a];
    for (int a = 0; a < 9; a++) t.dinv[a] = dinv[a];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) t.R[3 * i + j] = tipRot[4 * i + j];
    for (int a = 0; a < 3; a++) { t.k[a] = stiffness[a]; t.d0[a] = dpos0[a]; }
    t.krad = stiffness[3];
    t.L = dpos0[3];
    t.surf = surf_ff;
    long total = 0;
    #pragma omp parallel for schedule(dynamic, 16) reduction(+ : total)
    for (int ip = 0; ip < npix; ip++) {
        const double *A = tipRot, *B = tipRot + 4, *C = tipRot + 8;
        const v3 dtip = {C[0] * C[3], C[1] * C[3], C[2] * C[3]};
        const v3 d0 = {A[0] * dpos0[0] + B[0] * dpos0[1] + C[0] * dpos0[2],
                       A[1] * dpos0[0] + B[1] * dpos0[1] + C[1] * dpos0[2],
                       A[2] * dpos0[0] + B[2] * dpos0[1] + C[2] * dpos0[2]};
        v3 tp = {poss[4 * ip], poss[4 * ip + 1], poss[4 * ip + 2]};
        v3 r = {tp.x + d0.x, tp.y + d0.y, tp.z + d0.z};
        double dt = relax_params[0], damp = relax_params[1];
        const double dtmax = dt, dtmin = dtmax * 0.1, damp0 = damp;
        v3 u1 = {0, 0, 0}, u2 = {0, 0, 0};  /* probe offsets from the tip one and two steps back */
        for (int iz = 0; iz < nz; iz++) {
            int ne = 0, check = 0, settled = 0, itr = 0;
            double fe[4] = {0, 0, 0, 0};
            const int kmax = iz == 0 ? newton_max0 : newton_max;
            v3 q = r;
            int est = 0;
            if (iz == 0 && starts0 && isfinite(starts0[3 * ip]) && isfinite(starts0[3 * ip + 1]) && isfinite(starts0[3 * ip + 2])) {
                q = (v3){starts0[3 * ip], starts0[3 * ip + 1], starts0[3 * ip + 2]};
                est = 1;
            }
            /* second-order predictor (as orc_relax_strokes_warm) */
            if (extrap > 0 && iz >= 2 && kmax > 0) {
                const v3 e = {u1.x - u2.x, u1.y - u2.y, u1.z - u2.z};
                if (e.x * e.x + e.y * e.y + e.z * e.z <= extrap * extrap) { q.x += e.x; q.y += e.y; q.z += e.z; }
            }
            /* residual gate at the step-0 estimate (as orc_relax_strokes_warm) */
            int gated = 0;
            if (est && res_max > 0 && kmax > 0) {
                double f0[3], J0[3][3], fe0[4];
                tw_force_jac(&t, tp, q, fe0, f0, J0);
                gated = !(f0[0] * f0[0] + f0[1] * f0[1] + f0[2] * f0[2] <= res_max * res_max);
                if (gated) ne++;
            }
            const int res = gated ? 0 : kmax == 0 ? -1 : tw_newton_run(&t, tp, &q, kmax, tol, cap, conv_f2, &ne, fe);
            if (res == -1) {
                r = q;
            } else if (res == 2) {
                r = q;
                settled = 1;
            } else if (res == 1) {
                r = q;
                check = 1;
            } else {
                total++;
            }
            if (!settled)
                itr = tw_fire(&t, tp, &r, &dt, &damp, dtmax, dtmin, damp0, max_iters, conv_f2, f2safe, check, retry,
                              newton_max, tol, cap, &ne, fe);
            const long ind = (long)ip * nz + iz;
            FEs[4 * ind + 0] = fe[0] * A[0] + fe[1] * A[1] + fe[2] * A[2];
            FEs[4 * ind + 1] = fe[0] * B[0] + fe[1] * B[1] + fe[2] * B[2];
            FEs[4 * ind + 2] = fe[0] * C[0] + fe[1] * C[1] + fe[2] * C[2];
            FEs[4 * ind + 3] = fe[3];
            paths[4 * ind + 0] = r.x; paths[4 * ind + 1] = r.y; paths[4 * ind + 2] = r.z; paths[4 * ind + 3] = 0;
            if (iters) iters[ind] = itr;
            if (nevals) nevals[ind] = ne;
            u2 = u1;
            u1 = (v3){r.x - tp.x, r.y - tp.y, r.z - tp.z};
            tp.x += dtip.x; tp.y += dtip.y; tp.z += dtip.z;
            r.x += dtip.x; r.y += dtip.y; r.z += dtip.z;
        }
    }
    return total;
}

/* ------------------------------------------------------------ z convolution */

/* Valid-mode correlation along the last axis: out[j] = sum_k in[j+k] * w[k],
 * nzout = nzin - nw + 1 (convolveZ, relax.cl:571-608, which is the same as
 * numpy.convolve(in, w[::-1], 'valid')).  ncol columns, stride nzin / nzout. */
void orc_convolve_z(long ncol, int nzin, int nw, const double *in, const double *w, double *out)
{
    int nzout = nzin - nw + 1;
    #pragma omp parallel for schedule(static)
    for (long c = 0; c < ncol; c++)
        for (int j = 0; j < nzout; j++) {
            double s = 0;
            for (int k = 0; k < nw; k++) s += in[c * nzin + j + k] * w[k];
            out[c * nzout + j] = s;
        }
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
