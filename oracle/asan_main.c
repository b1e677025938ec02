/* Sanitizer driver of the oracle (test infrastructure): every function the
 * tests call, on small seeded inputs in buffers of exactly the sizes the
 * functions are documented to touch, so AddressSanitizer reports any access
 * outside them and UndefinedBehaviorSanitizer any undefined operation.
 * Built and run by `make -C oracle asan` (serial: no OpenMP). */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

void orc_lj_ff(const int n[3], const double cell[9], int natoms, const double *Rs, const double *cLJs, double *F,
               double *E);
void orc_coulomb_ff(const int n[3], const double cell[9], int natoms, const double *Rs, const double *kQQs, int kind,
                    double *F, double *E);
void orc_limit_vec_field(long npts, double *F, double fmax);
long orc_relax_strokes(int nx, int ny, int nz, const int n[3], const double cell[9], const double *F, const double tip[8],
                       const double relax[7], int probe_start, int alg, const double *rtips, double *rs, double *fs,
                       int *iters, int *nevals);
long orc_relax_strokes_warm(int nx, int ny, int nz, const int n[3], const double cell[9], const double *F,
                            const double tip[8], const double relax[7], const double *rtips, const double *starts0,
                            int newton_max, int newton_max0, int retry, double tol, double cap, double extrap,
                            double res_max, double *rs, double *fs, int *iters, int *nevals);
long orc_stiffness(const int n[3], const double cell[9], const double *F, const double tip[8], double ddisp, int which,
                   long np, const double *rtips, const double *rpps, double *evals, double *ev1, double *ev2,
                   double *ev3, double *dynmat);
void orc_ljc_qzs_ff(const int n[3], const double p0[3], const double dA[3], const double dB[3], const double dC[3],
                    int natoms, const double *atoms, const double *cLJs, const double Qs[4], const double QZs[4],
                    double *FE);
void orc_relax_strokes_tilted(int npix, int nz, const int n[3], const double *FE, const double dinv[9],
                              const double tipRot[12], const double stiffness[4], const double dpos0[4],
                              const double relax_params[2], double surf_ff, int max_iters, double conv_f2, double f2safe,
                              const double *poss, double *FEs, double *paths, int *nevals);
long orc_relax_strokes_tilted_warm(int npix, int nz, const int n[3], const double *FE, const double dinv[9],
                                   const double tipRot[12], const double stiffness[4], const double dpos0[4],
                                   const double relax_params[2], double surf_ff, int max_iters, double conv_f2,
                                   double f2safe, const double *poss, const double *starts0, int newton_max,
                                   int newton_max0, int retry, double tol, double cap, double extrap,
                                   double res_max, double *FEs, double *paths, int *iters, int *nevals);
void orc_model_ff(int model, double par, const int n[3], const double cell[9], int natoms, const double *Rs,
                  const double *coefs, int ncoef, double *F, double *E);
void orc_convolve_z(long ncol, int nzin, int nw, const double *in, const double *w, double *out);

static double urand(unsigned *s)
{
    *s = *s * 1664525u + 1013904223u;
    return (*s >> 8) * (1.0 / 16777216.0);
}

static void *take(size_t n)
{
    void *p = calloc(n, 1);
    if (!p) {
        fprintf(stderr, "out of memory\n");
        exit(2);
    }
    return p;
}

int main(void)
{
    unsigned seed = 12345;
    const int n[3] = {20, 18, 24};
    const long npts = (long)n[0] * n[1] * n[2];
    const double cell[9] = {2.0, 0.0, 0.0, 0.1, 1.8, 0.0, 0.0, 0.05, 2.4};
    const int na = 5;
    double *Rs = take(sizeof(double) * 3 * na), *cLJs = take(sizeof(double) * 2 * na), *kq = take(sizeof(double) * na);
    for (int i = 0; i < na; i++) {
        Rs[3 * i] = 0.4 + 1.2 * urand(&seed);
        Rs[3 * i + 1] = 0.4 + 1.0 * urand(&seed);
        Rs[3 * i + 2] = 0.6 + 0.2 * urand(&seed);
        cLJs[2 * i] = 20 + 10 * urand(&seed);
        cLJs[2 * i + 1] = 1e4 + 2e4 * urand(&seed);
        kq[i] = 14.4 * (urand(&seed) - 0.5);
    }
    double *F = take(sizeof(double) * 3 * npts), *E = take(sizeof(double) * npts);
    orc_lj_ff(n, cell, na, Rs, cLJs, F, E);
    for (int kind = 0; kind < 3; kind++) orc_coulomb_ff(n, cell, na, Rs, kq, kind, F, E);
    /* the further pair models read up to 4 coefficients per atom (D3: c6, c8, R0^6, R0^8) */
    double *coef4 = take(sizeof(double) * 4 * na);
    for (int i = 0; i < na; i++) {
        coef4[4 * i] = 1.5 + urand(&seed);
        coef4[4 * i + 1] = 0.01 + 0.01 * urand(&seed);
        coef4[4 * i + 2] = 30 + 10 * urand(&seed);
        coef4[4 * i + 3] = 300 + 100 * urand(&seed);
    }
    for (int model = 4; model <= 18; model++) orc_model_ff(model, 0.5, n, cell, na, Rs, coef4, 4, F, E);
    orc_limit_vec_field(npts, F, 10.0);
    /* strokes: 3 x 2 columns of 5 steps above the atoms */
    const int sx = 3, sy = 2, sz = 5, ns = sx * sy * sz;
    double *tips = take(sizeof(double) * 3 * ns), *rs = take(sizeof(double) * 3 * ns), *fs = take(sizeof(double) * 3 * ns);
    int *it = take(sizeof(int) * ns), *ne = take(sizeof(int) * ns);
    for (int i = 0; i < sx; i++)
        for (int j = 0; j < sy; j++)
            for (int k = 0; k < sz; k++) {
                const int o = (i * sy + j) * sz + k;
                tips[3 * o] = 0.6 + 0.3 * i;
                tips[3 * o + 1] = 0.6 + 0.3 * j;
                tips[3 * o + 2] = 6.0 - 0.1 * k;
            }
    const double tip[8] = {3.0, 30 / -16.0217662, 0, 0, -3.0, 0.25 / -16.0217662, 0.25 / -16.0217662, 0};
    const double relax[7] = {1000, 1e-8, 0.1, 0.1, 1.1, 0.5, 0.99};
    for (int ps = -1; ps <= 1; ps++) orc_relax_strokes(sx, sy, sz, n, cell, F, tip, relax, ps, 1, tips, rs, fs, it, ne);
    orc_relax_strokes(sx, sy, sz, n, cell, F, tip, relax, 1, 0, tips, rs, fs, it, ne);
    orc_relax_strokes_warm(sx, sy, sz, n, cell, F, tip, relax, tips, 0, 8, 12, 8, 1e-3, 0.1, 0.05, 0.0, rs, fs, it, ne);
    double *ev = take(sizeof(double) * 3 * ns), *v1 = take(sizeof(double) * 3 * ns), *v2 = take(sizeof(double) * 3 * ns),
           *v3 = take(sizeof(double) * 3 * ns), *dm = take(sizeof(double) * 9 * ns);
    orc_stiffness(n, cell, F, tip, 0.05, 3, ns, tips, rs, ev, v1, v2, v3, dm);
    /* OpenCL-mode grid and tilted strokes */
    double *atoms = take(sizeof(double) * 4 * na), *FE = take(sizeof(double) * 4 * npts);
    for (int i = 0; i < na; i++) {
        for (int c = 0; c < 3; c++) atoms[4 * i + c] = Rs[3 * i + c];
        atoms[4 * i + 3] = kq[i] / 14.4;
    }
    const double p0[3] = {0, 0, 0}, dA[3] = {0.1, 0, 0}, dB[3] = {0.005, 0.1, 0}, dC[3] = {0, 0.002, 0.1};
    const double Qs[4] = {-10, 20, -10, 0}, QZs[4] = {0.1, 0, -0.1, 0};
    orc_ljc_qzs_ff(n, p0, dA, dB, dC, na, atoms, cLJs, Qs, QZs, FE);
    const double dinv[9] = {10, -0.5, 0, 0, 10, -0.2, 0, 0, 10};
    const double rot[12] = {0.98, 0, 0.17, 0, 0, 1, 0, 0, -0.17, 0, 0.98, -0.1};
    const double st[4] = {-0.016, -0.016, 0, -1.9}, d0[4] = {0, 0, -3, 3}, rp[2] = {0.5, 0.1};
    const int npix = sx * sy;
    double *poss = take(sizeof(double) * 4 * npix), *FEs = take(sizeof(double) * 4 * ns), *paths = take(sizeof(double) * 4 * ns);
    double *starts = take(sizeof(double) * 3 * npix);
    for (int p = 0; p < npix; p++) {
        poss[4 * p] = tips[3 * p * sz];
        poss[4 * p + 1] = tips[3 * p * sz + 1];
        poss[4 * p + 2] = tips[3 * p * sz + 2];
        for (int c = 0; c < 3; c++) starts[3 * p + c] = poss[4 * p + c] + (c == 2 ? -3.0 : 0.0);
    }
    orc_relax_strokes_tilted(npix, sz, n, FE, dinv, rot, st, d0, rp, 0.01, 128, 1e-8, 1e-8, poss, FEs, paths, ne);
    orc_relax_strokes_tilted_warm(npix, sz, n, FE, dinv, rot, st, d0, rp, 0.01, 128, 1e-8, 1e-8, poss, starts, 8, 12, 8,
                                  1e-3, 0.1, 0.05, 0.5, FEs, paths, it, ne);
    /* step-0 estimates with the residual gate (starts: one per column, C order of the columns) */
    orc_relax_strokes_warm(sx, sy, sz, n, cell, F, tip, relax, tips, starts, 8, 12, 8, 1e-3, 0.1, 0.05, 0.5, rs, fs, it,
                           ne);
    double *w = take(sizeof(double) * 4), *cv = take(sizeof(double) * npix * (sz - 3));
    for (int k = 0; k < 4; k++) w[k] = urand(&seed);
    orc_convolve_z(npix, sz, 4, fs, w, cv);
    double chk = 0;
    for (long i = 0; i < 3 * ns; i++) chk += rs[i] + fs[i];
    printf("oracle sanitizer driver: ok (checksum %.6g)\n", chk);
    free(Rs); free(cLJs); free(kq); free(F); free(E); free(tips); free(rs); free(fs); free(it); free(ne);
    free(ev); free(v1); free(v2); free(v3); free(dm); free(atoms); free(FE); free(poss); free(FEs); free(paths);
    free(starts); free(w); free(cv); free(coef4);
    return 0;
}
