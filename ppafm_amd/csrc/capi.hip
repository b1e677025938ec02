// C ABI of libppafm_hip.so: the reference-compatible host entry points
// (names and arguments of ppafm/core.py:33-324 over ppafm/cpp/
// ProbeParticle.cpp:319-701) and the device entry points (ppafm_*).

#include <cmath>
#include <mutex>
#include <vector>

#include "pp_common.hpp"

namespace pp {

namespace {
thread_local std::string g_err;
}

void set_error(const std::string &msg) { g_err = msg; }
const char *get_error() { return g_err.c_str(); }

hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

GridGeom make_geom(const int n[3], const double cell[9])
{
    // GridShape::setCell + Mat3::invert_T_to (Grid.h:29-36, Mat3.h:254-286)
    GridGeom g;
    for (int i = 0; i < 3; i++) g.n[i] = n[i];
    for (int r = 0; r < 3; r++) {
        const double s = 1.0 / n[r];
        for (int c = 0; c < 3; c++) g.dcell[3 * r + c] = cell[3 * r + c] * s;
    }
    const double *m = g.dcell;
    const double xx = m[0], xy = m[1], xz = m[2], yx = m[3], yy = m[4], yz = m[5], zx = m[6], zy = m[7], zz = m[8];
    const double c_xx = yy * zz - yz * zy, c_yx = yz * zx - yx * zz, c_zx = yx * zy - yy * zx;
    const double idet = 1 / (xx * c_xx + xy * c_yx + xz * c_zx);
    g.dinv[0] = (yy * zz - yz * zy) * idet;
    g.dinv[3] = (xz * zy - xy * zz) * idet;
    g.dinv[6] = (xy * yz - xz * yy) * idet;
    g.dinv[1] = (yz * zx - yx * zz) * idet;
    g.dinv[4] = (xx * zz - xz * zx) * idet;
    g.dinv[7] = (xz * yx - xx * yz) * idet;
    g.dinv[2] = (yx * zy - yy * zx) * idet;
    g.dinv[5] = (xy * zx - xx * zy) * idet;
    g.dinv[8] = (xx * yy - xy * yx) * idet;
    return g;
}

namespace {

// Grow-only device buffer.
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes)
    {
        if (bytes <= cap) return 0;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        PP_TRY_HIP(hipMalloc(&p, bytes));
        cap = bytes;
        return 0;
    }
};

// Process-wide state of the reference-compatible ABI (ProbeParticle.cpp:33-114).
struct RefState {
    std::mutex mu;
    int n[3] = {0, 0, 0};
    double cell[9] = {0};
    bool have_cell = false;
    double *F = nullptr;
    double *E = nullptr;
    // TIP
    double rpp0[3] = {0, 0, 0}, kspring[3] = {0, 0, 0};
    double lrad = 0, krad = 0;
    // RELAX / FIRE (ProbeParticle.cpp:76-103)
    int max_iters = 1000;
    double conv_f2 = 1.0e-8, dt = 0.1, damping = 0.1;
    double finc = 1.1, fdec = 0.5, falpha = 0.99;
    long long last_itrsum = 0;
    DevBuf dF, dE, dAtoms, dTips, dPos, dFs, dIters, dEv;
};

RefState &ref()
{
    static RefState s;
    return s;
}

int ref_grid_ready(RefState &S)
{
    PP_REQUIRE(S.n[0] > 0 && S.n[1] > 0 && S.n[2] > 0, "setGridN has not been called");
    PP_REQUIRE(S.have_cell, "setGridCell has not been called");
    return 0;
}

int ref_accum(int model, int natoms, const double *Rs, const double *coefs)
{
    RefState &S = ref();
    std::lock_guard<std::mutex> lk(S.mu);
    int e = ref_grid_ready(S);
    if (e) return e;
    PP_REQUIRE(natoms >= 0, "natoms must be >= 0");
    if (!S.F && !S.E) return 0;
    const GridGeom g = make_geom(S.n, S.cell);
    const size_t npts = (size_t)S.n[0] * S.n[1] * S.n[2];
    std::vector<AtomRec> h(natoms);
    for (int i = 0; i < natoms; i++) {
        AtomRec a{Rs[3 * i], Rs[3 * i + 1], Rs[3 * i + 2], 0, 0, 0};
        if (model == 0) { a.c6 = coefs[2 * i]; a.c12 = coefs[2 * i + 1]; }
        else a.q = coefs[i];
        h[i] = a;
    }
    if ((e = S.dAtoms.ensure(sizeof(AtomRec) * (natoms > 0 ? natoms : 1)))) return e;
    if (natoms) PP_TRY_HIP(hipMemcpy(S.dAtoms.p, h.data(), sizeof(AtomRec) * natoms, hipMemcpyHostToDevice));
    double *dF = nullptr, *dE = nullptr;
    if (S.F) {
        if ((e = S.dF.ensure(npts * 3 * sizeof(double)))) return e;
        PP_TRY_HIP(hipMemcpy(S.dF.p, S.F, npts * 3 * sizeof(double), hipMemcpyHostToDevice));
        dF = (double *)S.dF.p;
    }
    if (S.E) {
        if ((e = S.dE.ensure(npts * sizeof(double)))) return e;
        PP_TRY_HIP(hipMemcpy(S.dE.p, S.E, npts * sizeof(double), hipMemcpyHostToDevice));
        dE = (double *)S.dE.p;
    }
    if ((e = launch_ff_accum_f64(g, model, natoms, (const AtomRec *)S.dAtoms.p, dF, dE, nullptr))) return e;
    PP_TRY_HIP(hipGetLastError());
    if (S.F) PP_TRY_HIP(hipMemcpy(S.F, S.dF.p, npts * 3 * sizeof(double), hipMemcpyDeviceToHost));
    if (S.E) PP_TRY_HIP(hipMemcpy(S.E, S.dE.p, npts * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int ref_relax(int nx, int ny, int probeStart, int relaxAlg, int nstep, const double *rTips, double *rs,
              double *fs, const ppafm_spline_params *spline, long long *itrsum)
{
    RefState &S = ref();
    std::lock_guard<std::mutex> lk(S.mu);
    int e = ref_grid_ready(S);
    if (e) return e;
    PP_REQUIRE(spline == nullptr, "tip splines are not supported (pass None for the harmonic radial spring)");
    PP_REQUIRE(S.F != nullptr, "setFF_Fpointer has not been called");
    PP_REQUIRE(nx >= 0 && ny >= 0 && nstep >= 0, "negative scan size");
    const size_t npts = (size_t)S.n[0] * S.n[1] * S.n[2];
    const size_t npos = (size_t)nx * ny * nstep;
    *itrsum = 0;
    if (npos == 0) return 0;
    if ((e = S.dF.ensure(npts * 3 * sizeof(double)))) return e;
    if ((e = S.dTips.ensure(npos * 3 * sizeof(double)))) return e;
    if ((e = S.dPos.ensure(npos * 3 * sizeof(double)))) return e;
    if ((e = S.dFs.ensure(npos * 3 * sizeof(double)))) return e;
    if ((e = S.dIters.ensure(npos * sizeof(int)))) return e;
    PP_TRY_HIP(hipMemcpy(S.dF.p, S.F, npts * 3 * sizeof(double), hipMemcpyHostToDevice));
    PP_TRY_HIP(hipMemcpy(S.dTips.p, rTips, npos * 3 * sizeof(double), hipMemcpyHostToDevice));
    RelaxLaunch L;
    L.g = make_geom(S.n, S.cell);
    L.rp = ppafm_relax_params{0, relaxAlg, S.max_iters, probeStart, S.conv_f2, S.dt, S.damping,
                              S.finc, S.fdec, S.falpha, 0.0, 0.0};
    for (int k = 0; k < 9; k++) L.tip.rot[k] = (k % 4 == 0) ? 1.0 : 0.0;
    for (int k = 0; k < 3; k++) { L.tip.d0[k] = S.rpp0[k]; L.tip.kspring[k] = S.kspring[k]; }
    L.tip.krad = S.krad; L.tip.lrad = S.lrad; L.tip.surf_ff = 0;
    L.scan = ppafm_scan{nx, ny, nstep, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    L.tips = (const double *)S.dTips.p;
    L.out = ppafm_relax_out{nullptr, 0, nullptr, nullptr, nullptr};
    if ((e = launch_relax_f64grid(L, (const double *)S.dF.p, (double *)S.dPos.p, (double *)S.dFs.p,
                                  (int *)S.dIters.p, nullptr)))
        return e;
    PP_TRY_HIP(hipGetLastError());
    PP_TRY_HIP(hipMemcpy(rs, S.dPos.p, npos * 3 * sizeof(double), hipMemcpyDeviceToHost));
    PP_TRY_HIP(hipMemcpy(fs, S.dFs.p, npos * 3 * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<int> it(npos);
    PP_TRY_HIP(hipMemcpy(it.data(), S.dIters.p, npos * sizeof(int), hipMemcpyDeviceToHost));
    long long sum = 0;
    for (size_t k = 0; k < npos; k++) sum += it[k];
    *itrsum = sum;
    S.last_itrsum = sum;
    return 0;
}

// stiffnessMatrix (ProbeParticle.cpp:703-750) on the process-wide grid and
// tip: rTips, rPPs (n,3) host float64 in, eigenvalues (n,3) ascending and
// the first `which` eigenvectors (n,3) out.  The reference prints and exits
// the process on a NaN eigenvalue (ProbeParticle.cpp:735); this reports it
// as an error after writing every output.
int ref_stiffness(double ddisp, int which, int n, const double *rTips, const double *rPPs, double *evals,
                  double *ev1, double *ev2, double *ev3, const ppafm_spline_params *spline)
{
    RefState &S = ref();
    std::lock_guard<std::mutex> lk(S.mu);
    int e = ref_grid_ready(S);
    if (e) return e;
    PP_REQUIRE(spline == nullptr, "tip splines are not supported (pass None for the harmonic radial spring)");
    PP_REQUIRE(S.F != nullptr, "setFF_Fpointer has not been called");
    PP_REQUIRE(n >= 0, "n must be >= 0");
    if (n == 0) return 0;
    PP_REQUIRE(rTips && rPPs && evals, "null rTips, rPPs or eigenvals");
    const int nv = which < 0 ? 0 : (which > 3 ? 3 : which);
    double *hv[3] = {ev1, ev2, ev3};
    for (int k = 0; k < nv; k++) PP_REQUIRE(hv[k], "an eigenvector output requested by `which` is NULL");
    const size_t npts = (size_t)S.n[0] * S.n[1] * S.n[2];
    const size_t b3 = (size_t)n * 3 * sizeof(double);
    if ((e = S.dF.ensure(npts * 3 * sizeof(double)))) return e;
    if ((e = S.dTips.ensure(b3))) return e;
    if ((e = S.dPos.ensure(b3))) return e;
    if ((e = S.dFs.ensure(b3))) return e;
    if ((e = S.dEv.ensure(b3 * 3))) return e;
    PP_TRY_HIP(hipMemcpy(S.dF.p, S.F, npts * 3 * sizeof(double), hipMemcpyHostToDevice));
    PP_TRY_HIP(hipMemcpy(S.dTips.p, rTips, b3, hipMemcpyHostToDevice));
    PP_TRY_HIP(hipMemcpy(S.dPos.p, rPPs, b3, hipMemcpyHostToDevice));
    RelaxLaunch L;
    L.g = make_geom(S.n, S.cell);
    L.rp = ppafm_relax_params{0, 1, S.max_iters, 1, S.conv_f2, S.dt, S.damping, S.finc, S.fdec, S.falpha, 0.0, 0.0};
    for (int k = 0; k < 9; k++) L.tip.rot[k] = (k % 4 == 0) ? 1.0 : 0.0;
    for (int k = 0; k < 3; k++) { L.tip.d0[k] = S.rpp0[k]; L.tip.kspring[k] = S.kspring[k]; }
    L.tip.krad = S.krad; L.tip.lrad = S.lrad; L.tip.surf_ff = 0;
    L.scan = ppafm_scan{1, 1, n, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    L.tips = (const double *)S.dTips.p;
    L.out = ppafm_relax_out{nullptr, 0, nullptr, nullptr, nullptr};
    double *dv = (double *)S.dEv.p;
    if ((e = launch_stiffness(L, S.dF.p, 2, S.dPos.p, 0, ddisp, nv, (double *)S.dFs.p, dv, dv + 3 * (size_t)n,
                              dv + 6 * (size_t)n, nullptr)))
        return e;
    PP_TRY_HIP(hipGetLastError());
    PP_TRY_HIP(hipMemcpy(evals, S.dFs.p, b3, hipMemcpyDeviceToHost));
    for (int k = 0; k < nv; k++) PP_TRY_HIP(hipMemcpy(hv[k], dv + 3 * (size_t)k * n, b3, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++)
        if (std::isnan(evals[3 * i]) || std::isnan(evals[3 * i + 1]) || std::isnan(evals[3 * i + 2])) {
            char msg[160];
            snprintf(msg, sizeof msg, "stiffnessMatrix: NaN eigenvalues at point %d (first of them)", i);
            set_error(msg);
            return -1;
        }
    return 0;
}

// The reference ABI returns void: failures are reported on stderr (the
// reference prints and carries on, ProbeParticle.cpp:659-661) and through
// ppafm_last_error().
void report(int rc, const char *fn)
{
    if (rc) fprintf(stderr, "libppafm_hip: %s failed: %s\n", fn, get_error());
}

}  // namespace
}  // namespace pp

using namespace pp;

extern "C" {

// ----------------------------------------------------- reference-compatible

void setGridN(int *n)
{
    RefState &S = ref();
    std::lock_guard<std::mutex> lk(S.mu);
    for (int i = 0; i < 3; i++) S.n[i] = n[i];
}

void setGridCell(double *cell)
{
    RefState &S = ref();
    std::lock_guard<std::mutex> lk(S.mu);
    for (int i = 0; i < 9; i++) S.cell[i] = cell[i];
    S.have_cell = true;
}

void setFF_Fpointer(double *gridF)
{
    RefState &S = ref();
    std::lock_guard<std::mutex> lk(S.mu);
    S.F = gridF;
}

void setFF_Epointer(double *gridE)
{
    RefState &S = ref();
    std::lock_guard<std::mutex> lk(S.mu);
    S.E = gridE;
}

void deleteFF_Fpointer(void)
{
    RefState &S = ref();
    std::lock_guard<std::mutex> lk(S.mu);
    S.F = nullptr;
}

void deleteFF_Epointer(void)
{
    RefState &S = ref();
    std::lock_guard<std::mutex> lk(S.mu);
    S.E = nullptr;
}

void setTip(double lRad, double kRad, double *rPP0, double *kSpring)
{
    RefState &S = ref();
    std::lock_guard<std::mutex> lk(S.mu);
    S.lrad = lRad;
    S.krad = kRad;
    for (int i = 0; i < 3; i++) { S.rpp0[i] = rPP0[i]; S.kspring[i] = kSpring[i]; }
    // TIP::makeConsistent (ProbeParticle.cpp:64-69)
    if (std::fabs(kRad) > 1e-8) S.rpp0[2] = -std::sqrt(lRad * lRad - S.rpp0[0] * S.rpp0[0] - S.rpp0[1] * S.rpp0[1]);
}

void setRelax(int maxIters, double convF2, double dt, double damping)
{
    RefState &S = ref();
    std::lock_guard<std::mutex> lk(S.mu);
    S.max_iters = maxIters;
    S.conv_f2 = convF2;
    S.dt = dt;
    S.damping = damping;
}

void setFIRE(double finc, double fdec, double falpha)
{
    RefState &S = ref();
    std::lock_guard<std::mutex> lk(S.mu);
    S.finc = finc;
    S.fdec = fdec;
    S.falpha = falpha;
}

void getLennardJonesFF(int natoms, double *Rs, double *cLJs) { report(ref_accum(0, natoms, Rs, cLJs), "getLennardJonesFF"); }

void getCoulombFF(int natoms, double *Rs, double *kQQs, int kind)
{
    if (kind < 0 || kind > 2) {
        set_error("getCoulombFF: kind must be 0 (s), 1 (pz) or 2 (dz2)");
        report(-2, "getCoulombFF");
        return;
    }
    report(ref_accum(1 + kind, natoms, Rs, kQQs), "getCoulombFF");
}

int relaxTipStroke(int probeStart, int relaxAlg, int nstep, double *rTips, double *rs, double *fs,
                   ppafm_spline_params *splineParams)
{
    long long sum = 0;
    int rc = ref_relax(1, 1, probeStart, relaxAlg, nstep, rTips, rs, fs, splineParams, &sum);
    report(rc, "relaxTipStroke");
    return rc ? -1 : (int)sum;
}

int relaxTipStrokes_omp(int nx, int ny, int probeStart, int relaxAlg, int nstep, double *rTips, double *rs,
                        double *fs, ppafm_spline_params *splineParams)
{
    long long sum = 0;
    int rc = ref_relax(nx, ny, probeStart, relaxAlg, nstep, rTips, rs, fs, splineParams, &sum);
    report(rc, "relaxTipStrokes_omp");
    return rc ? -1 : 0;
}

void stiffnessMatrix(double ddisp, int which, int n, double *rTips, double *rPPs, double *eigenvals, double *evec1,
                     double *evec2, double *evec3, ppafm_spline_params *splineParams)
{
    report(ref_stiffness(ddisp, which, n, rTips, rPPs, eigenvals, evec1, evec2, evec3, splineParams),
           "stiffnessMatrix");
}

long long ppafm_last_itrsum(void) { return ref().last_itrsum; }

// ------------------------------------------------------------------- device

int ppafm_ff_build_dev(const ppafm_grid *grid, const ppafm_ff_model *model, int natoms, const ppafm_atom *atoms_dev,
                       void *out_dev, void *stream)
{
    PP_REQUIRE(grid && model && out_dev, "null argument");
    PP_REQUIRE(natoms == 0 || atoms_dev, "null atoms");
    const GridGeom g = make_geom(grid->n, grid->cell);
    int e = launch_ff_build(g, *model, natoms, (const AtomRec *)atoms_dev, out_dev, grid->f64 != 0, as_stream(stream));
    if (e) return e;
    PP_TRY_HIP(hipGetLastError());
    return 0;
}

int ppafm_pack_atoms(int natoms, const double *atoms_xyzq, const double *cLJs, double coulomb_const,
                     ppafm_atom *out_host)
{
    PP_REQUIRE(natoms >= 0 && (natoms == 0 || (atoms_xyzq && out_host)), "null argument");
    AtomRec *o = out_host;
    for (int i = 0; i < natoms; i++) {
        o[i].x = atoms_xyzq[4 * i];
        o[i].y = atoms_xyzq[4 * i + 1];
        o[i].z = atoms_xyzq[4 * i + 2];
        o[i].q = atoms_xyzq[4 * i + 3] * coulomb_const;
        o[i].c6 = cLJs ? cLJs[2 * i] : 0.0;
        o[i].c12 = cLJs ? cLJs[2 * i + 1] : 0.0;
    }
    return 0;
}

int ppafm_ff_build(const ppafm_grid *grid, const ppafm_ff_model *model, int natoms, const double *atoms_xyzq,
                   const double *cLJs, void *out_dev, void *stream)
{
    PP_REQUIRE(grid && model && out_dev, "null argument");
    static std::mutex mu;
    static DevBuf buf;
    std::lock_guard<std::mutex> lk(mu);
    std::vector<AtomRec> h(natoms > 0 ? natoms : 1);
    int e = ppafm_pack_atoms(natoms, atoms_xyzq, cLJs, model->coulomb_const, h.data());
    if (e) return e;
    // the staging buffer is reused: wait for earlier work on this stream
    PP_TRY_HIP(hipStreamSynchronize(as_stream(stream)));
    if ((e = buf.ensure(sizeof(AtomRec) * h.size()))) return e;
    PP_TRY_HIP(hipMemcpy(buf.p, h.data(), sizeof(AtomRec) * h.size(), hipMemcpyHostToDevice));
    return ppafm_ff_build_dev(grid, model, natoms, (const ppafm_atom *)buf.p, out_dev, stream);
}

int ppafm_relax(const ppafm_grid *grid, const void *grid_dev, const ppafm_relax_params *rp, const ppafm_tip *tip,
                const ppafm_scan *scan, const double *tips_dev, const ppafm_relax_out *out, void *stream)
{
    PP_REQUIRE(grid && grid_dev && rp && tip && scan && out, "null argument");
    RelaxLaunch L;
    L.g = make_geom(grid->n, grid->cell);
    L.rp = *rp;
    L.tip = *tip;
    L.scan = *scan;
    L.tips = tips_dev;
    L.out = *out;
    int e = launch_relax_dev(L, grid_dev, grid->f64 != 0, as_stream(stream));
    if (e) return e;
    PP_TRY_HIP(hipGetLastError());
    return 0;
}

int ppafm_stiffness(const ppafm_grid *grid, const void *grid_dev, const ppafm_tip *tip, const ppafm_scan *scan,
                    const double *tips_dev, const void *pos_dev, int pos_f32, double ddisp, int which,
                    double *evals_dev, double *evec1_dev, double *evec2_dev, double *evec3_dev, void *stream)
{
    PP_REQUIRE(grid && grid_dev && tip && scan && pos_dev && evals_dev, "null argument");
    RelaxLaunch L;
    L.g = make_geom(grid->n, grid->cell);
    L.rp = ppafm_relax_params{0, 1, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0};
    L.tip = *tip;
    L.scan = *scan;
    L.tips = tips_dev;
    L.out = ppafm_relax_out{nullptr, 0, nullptr, nullptr, nullptr};
    int e = launch_stiffness(L, grid_dev, grid->f64 ? 1 : 0, pos_dev, pos_f32, ddisp, which, evals_dev, evec1_dev,
                             evec2_dev, evec3_dev, as_stream(stream));
    if (e) return e;
    PP_TRY_HIP(hipGetLastError());
    return 0;
}

int ppafm_convolve_z(long ncol, int nz, int in_stride, int in_step, int in_offset, const void *in_dev, int nw,
                     const double *w, double scale, void *out_dev, int f64, void *stream)
{
    PP_REQUIRE(in_dev && out_dev && w, "null argument");
    int e = launch_convolve_z(ncol, nz, in_stride, in_step, in_offset, in_dev, nw, w, scale, out_dev, f64,
                              as_stream(stream));
    if (e) return e;
    PP_TRY_HIP(hipGetLastError());
    return 0;
}

int ppafm_limit_vec_field(double *F_dev, long npts, double fmax, void *stream)
{
    PP_REQUIRE(F_dev || npts == 0, "null argument");
    int e = launch_limit_vec_field(F_dev, npts, fmax, as_stream(stream));
    if (e) return e;
    PP_TRY_HIP(hipGetLastError());
    return 0;
}

int ppafm_copy_kernel(const void *src_dev, void *dst_dev, unsigned long long bytes, void *stream)
{
    PP_REQUIRE(bytes % 16 == 0, "copy size must be a multiple of 16 bytes");
    int e = launch_copy(src_dev, dst_dev, (long)(bytes / 16), as_stream(stream));
    if (e) return e;
    PP_TRY_HIP(hipGetLastError());
    return 0;
}

int ppafm_malloc(void **ptr, unsigned long long bytes)
{
    PP_REQUIRE(ptr, "null argument");
    PP_TRY_HIP(hipMalloc(ptr, bytes ? bytes : 1));
    return 0;
}

int ppafm_free(void *ptr)
{
    if (ptr) PP_TRY_HIP(hipFree(ptr));
    return 0;
}

int ppafm_memcpy_h2d(void *dst, const void *src, unsigned long long bytes, void *stream)
{
    if (!bytes) return 0;
    if (stream) PP_TRY_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)));
    else PP_TRY_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return 0;
}

int ppafm_memcpy_d2h(void *dst, const void *src, unsigned long long bytes, void *stream)
{
    if (!bytes) return 0;
    if (stream) PP_TRY_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
    else PP_TRY_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return 0;
}

int ppafm_memset(void *dst, int value, unsigned long long bytes, void *stream)
{
    if (!bytes) return 0;
    PP_TRY_HIP(hipMemsetAsync(dst, value, bytes, as_stream(stream)));
    return 0;
}

int ppafm_synchronize(void *stream)
{
    if (stream) PP_TRY_HIP(hipStreamSynchronize(as_stream(stream)));
    else PP_TRY_HIP(hipDeviceSynchronize());
    return 0;
}

int ppafm_event_create(void **ev)
{
    PP_REQUIRE(ev, "null argument");
    hipEvent_t e;
    PP_TRY_HIP(hipEventCreate(&e));
    *ev = (void *)e;
    return 0;
}

int ppafm_event_record(void *ev, void *stream)
{
    PP_TRY_HIP(hipEventRecord((hipEvent_t)ev, as_stream(stream)));
    return 0;
}

int ppafm_event_elapsed_ms(void *ev0, void *ev1, float *ms)
{
    PP_TRY_HIP(hipEventSynchronize((hipEvent_t)ev1));
    PP_TRY_HIP(hipEventElapsedTime(ms, (hipEvent_t)ev0, (hipEvent_t)ev1));
    return 0;
}

int ppafm_event_destroy(void *ev)
{
    if (ev) PP_TRY_HIP(hipEventDestroy((hipEvent_t)ev));
    return 0;
}

int ppafm_stream_create(void **stream)
{
    PP_REQUIRE(stream, "null argument");
    hipStream_t s;
    PP_TRY_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void *)s;
    return 0;
}

int ppafm_stream_destroy(void *stream)
{
    if (stream) PP_TRY_HIP(hipStreamDestroy(as_stream(stream)));
    return 0;
}

int ppafm_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int ppafm_device_info(char *name, int len, int *ncu)
{
    int dev = 0;
    PP_TRY_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    PP_TRY_HIP(hipGetDeviceProperties(&p, dev));
    if (name && len > 0) {
        std::snprintf(name, (size_t)len, "%s", p.gcnArchName);
    }
    if (ncu) *ncu = p.multiProcessorCount;
    return 0;
}

const char *ppafm_last_error(void) { return get_error(); }

void ppafm_clear_error(void) { set_error(""); }

}  // extern "C"
