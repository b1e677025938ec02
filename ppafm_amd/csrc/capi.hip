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
    // damping constants of the vdW models (ProbeParticle.cpp:48); a positive
    // ADamp passed to getVdWFF_RE / evalRadialFF replaces the model's value
    // for later calls too, as in the reference
    double adamp_const = 180.0, adamp_r2 = 0.5, adamp_r4 = 0.5, adamp_invr4 = 0.03, adamp_invr8 = 0.01;
    DevBuf dF, dE, dAtoms, dTips, dPos, dFs, dIters, dEv, dCoef, dSpline, dAux;
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

// evalCell<model> over the process-wide grid (ProbeParticle.cpp:262-278):
// += into the F and E grids that are set (E only for the densities, which
// the reference computes with the force grid switched off).
int ref_accum(int model, double par, int natoms, const double *Rs, const double *coefs, int ncoef,
              bool energy_only = false)
{
    RefState &S = ref();
    std::lock_guard<std::mutex> lk(S.mu);
    int e = ref_grid_ready(S);
    if (e) return e;
    PP_REQUIRE(natoms >= 0, "natoms must be >= 0");
    PP_REQUIRE(natoms == 0 || (Rs && coefs), "null atom positions or coefficients");
    double *hF = energy_only ? nullptr : S.F;
    if (!hF && !S.E) return 0;
    const GridGeom g = make_geom(S.n, S.cell);
    const size_t npts = (size_t)S.n[0] * S.n[1] * S.n[2];
    const size_t na = natoms > 0 ? natoms : 1;
    if ((e = S.dAtoms.ensure(na * 3 * sizeof(double)))) return e;
    if ((e = S.dCoef.ensure(na * (ncoef > 0 ? ncoef : 1) * sizeof(double)))) return e;
    if (natoms) {
        PP_TRY_HIP(hipMemcpy(S.dAtoms.p, Rs, (size_t)natoms * 3 * sizeof(double), hipMemcpyHostToDevice));
        PP_TRY_HIP(hipMemcpy(S.dCoef.p, coefs, (size_t)natoms * ncoef * sizeof(double), hipMemcpyHostToDevice));
    }
    double *dF = nullptr, *dE = nullptr;
    if (hF) {
        if ((e = S.dF.ensure(npts * 3 * sizeof(double)))) return e;
        PP_TRY_HIP(hipMemcpy(S.dF.p, hF, npts * 3 * sizeof(double), hipMemcpyHostToDevice));
        dF = (double *)S.dF.p;
    }
    if (S.E) {
        if ((e = S.dE.ensure(npts * sizeof(double)))) return e;
        PP_TRY_HIP(hipMemcpy(S.dE.p, S.E, npts * sizeof(double), hipMemcpyHostToDevice));
        dE = (double *)S.dE.p;
    }
    if ((e = launch_ff_models(g, model, par, natoms, (const double *)S.dAtoms.p, (const double *)S.dCoef.p, ncoef,
                              dF, dE, nullptr)))
        return e;
    PP_TRY_HIP(hipGetLastError());
    if (hF) PP_TRY_HIP(hipMemcpy(hF, S.dF.p, npts * 3 * sizeof(double), hipMemcpyDeviceToHost));
    if (S.E) PP_TRY_HIP(hipMemcpy(S.E, S.dE.p, npts * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

// Upload a tip spline (TIP::SplineParams: n knots xs, values and slopes
// ydys) into buf; fills L's spline fields.  NULL leaves the harmonic spring.
int upload_spline(DevBuf &buf, const ppafm_spline_params *sp, RelaxLaunch &L)
{
    if (!sp) return 0;
    PP_REQUIRE(sp->rff_n >= 2 && sp->rff_xs && sp->rff_ydys, "tip spline: need n >= 2 knots and non-NULL xs, ydys");
    int e = buf.ensure((size_t)sp->rff_n * 3 * sizeof(double));
    if (e) return e;
    double *d = (double *)buf.p;
    PP_TRY_HIP(hipMemcpy(d, sp->rff_xs, (size_t)sp->rff_n * sizeof(double), hipMemcpyHostToDevice));
    PP_TRY_HIP(hipMemcpy(d + sp->rff_n, sp->rff_ydys, (size_t)sp->rff_n * 2 * sizeof(double), hipMemcpyHostToDevice));
    L.spline_n = sp->rff_n;
    L.spline_xs = d;
    L.spline_ydys = d + sp->rff_n;
    return 0;
}

int ref_relax(int nx, int ny, int probeStart, int relaxAlg, int nstep, const double *rTips, double *rs,
              double *fs, const ppafm_spline_params *spline, long long *itrsum)
{
    RefState &S = ref();
    std::lock_guard<std::mutex> lk(S.mu);
    int e = ref_grid_ready(S);
    if (e) return e;
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
    if ((e = upload_spline(S.dSpline, spline, L))) return e;
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
    if ((e = upload_spline(S.dSpline, spline, L))) return e;
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

// Small host-array entry points: upload, one kernel, download.
struct Staged {
    std::vector<DevBuf> bufs;
    explicit Staged(int n) : bufs(n) {}
    ~Staged()
    {
        for (auto &b : bufs)
            if (b.p) (void)hipFree(b.p);
    }
    // device copy of a host array (bytes may be 0)
    int put(int k, const void *h, size_t bytes)
    {
        int e = bufs[k].ensure(bytes > 0 ? bytes : 8);
        if (e) return e;
        if (bytes) PP_TRY_HIP(hipMemcpy(bufs[k].p, h, bytes, hipMemcpyHostToDevice));
        return 0;
    }
    int alloc(int k, size_t bytes) { return bufs[k].ensure(bytes > 0 ? bytes : 8); }
    int get(int k, void *h, size_t bytes)
    {
        if (bytes) PP_TRY_HIP(hipMemcpy(h, bufs[k].p, bytes, hipMemcpyDeviceToHost));
        return 0;
    }
    template <class T> T *p(int k) { return (T *)bufs[k].p; }
};

#define PP_E(x)                                                                                  \
    do {                                                                                         \
        int _r = (x);                                                                            \
        if (_r) return _r;                                                                       \
    } while (0)

int host_points_lj(int npoints, const double *pts, double *fes, int natoms, const double *Rs, const double *cljs)
{
    PP_REQUIRE(npoints >= 0 && natoms >= 0, "negative count");
    if (npoints == 0) return 0;
    PP_REQUIRE(pts && fes && (natoms == 0 || (Rs && cljs)), "null argument");
    Staged S(4);
    PP_E(S.put(0, pts, (size_t)npoints * 3 * sizeof(double)));
    PP_E(S.put(1, Rs, (size_t)natoms * 3 * sizeof(double)));
    PP_E(S.put(2, cljs, (size_t)natoms * 2 * sizeof(double)));
    PP_E(S.alloc(3, (size_t)npoints * 4 * sizeof(double)));
    PP_E(launch_lj_points(npoints, S.p<double>(0), natoms, S.p<double>(1), S.p<double>(2), S.p<double>(3), nullptr));
    PP_TRY_HIP(hipGetLastError());
    return S.get(3, fes, (size_t)npoints * 4 * sizeof(double));
}

// evalRadialFF kinds: -3 LJ in (R, E), -2 LJ (c6, c12), -1 bare r^-6, 0 vdW
// with constant damping, 1..4 damping functions; others give zeros.
int host_radial(int n, const double *rs, const double *coefs, double *es, double *fs, int kind, double adamp)
{
    PP_REQUIRE(n >= 0, "n must be >= 0");
    if (n == 0) return 0;
    PP_REQUIRE(rs && es && fs, "null argument");
    RefState &R = ref();
    int model = -1;
    double par = 0;
    {
        std::lock_guard<std::mutex> lk(R.mu);
        double *slot[5] = {&R.adamp_const, &R.adamp_r2, &R.adamp_r4, &R.adamp_invr4, &R.adamp_invr8};
        if (kind >= 0 && kind <= 4) {
            if (adamp > 0) *slot[kind] = adamp;
            par = *slot[kind];
            model = kind == 0 ? (int)kModelVdW : (int)kModelVdWR2 + kind - 1;
        } else if (kind == -1) {
            model = kModelInvR6;
        } else if (kind == -2) {
            model = kModelLJ;
        } else if (kind == -3) {
            model = kModelLJRE;
        }
    }
    // the reference reads coefs[0] for the one-coefficient kinds and
    // coefs[0..1] for the others, nothing for unknown kinds
    const int nc = model < 0 ? 0 : ((kind == -1 || kind == 0) ? 1 : 2);
    PP_REQUIRE(nc == 0 || coefs, "null coefs");
    const double c0 = nc > 0 ? coefs[0] : 0.0, c1 = nc > 1 ? coefs[1] : 0.0;
    Staged S(3);
    PP_E(S.put(0, rs, (size_t)n * sizeof(double)));
    PP_E(S.alloc(1, (size_t)n * sizeof(double)));
    PP_E(S.alloc(2, (size_t)n * sizeof(double)));
    PP_E(launch_radial(n, S.p<double>(0), model, c0, c1, par, S.p<double>(1), S.p<double>(2), nullptr));
    PP_TRY_HIP(hipGetLastError());
    PP_E(S.get(1, es, (size_t)n * sizeof(double)));
    return S.get(2, fs, (size_t)n * sizeof(double));
}

int host_d3_coeffs(int natoms, const double *rs, const int *elems, const double *r_cov, const double *ref_cn,
                   const double *ref_c6, const double *r4r2, const double *k, const double *params, int elem_pp,
                   double *out)
{
    constexpr int kElem = 94, kRef = 5;  // ProbeParticle.cpp:28-29
    PP_REQUIRE(natoms >= 0, "natoms must be >= 0");
    if (natoms == 0) return 0;
    PP_REQUIRE(rs && elems && r_cov && ref_cn && ref_c6 && r4r2 && k && params && out, "null argument");
    PP_REQUIRE(elem_pp >= 1 && elem_pp <= kElem, "probe element out of range 1..94");
    for (int i = 0; i < natoms; i++)
        PP_REQUIRE(elems[i] >= 1 && elems[i] <= kElem, "atom element out of range 1..94");
    Staged S(7);
    PP_E(S.put(0, rs, (size_t)natoms * 3 * sizeof(double)));
    PP_E(S.put(1, elems, (size_t)natoms * sizeof(int)));
    PP_E(S.put(2, r_cov, kElem * sizeof(double)));
    PP_E(S.put(3, ref_cn, kElem * kRef * sizeof(double)));
    PP_E(S.put(4, ref_c6, (size_t)kElem * kElem * kRef * kRef * sizeof(double)));
    PP_E(S.put(5, r4r2, kElem * sizeof(double)));
    PP_E(S.alloc(6, (size_t)natoms * 4 * sizeof(double)));
    PP_E(launch_d3_coeffs(natoms, S.p<double>(0), S.p<int>(1), S.p<double>(2), S.p<double>(3), S.p<double>(4),
                          S.p<double>(5), k, params, elem_pp, S.p<double>(6), nullptr));
    PP_TRY_HIP(hipGetLastError());
    return S.get(6, out, (size_t)natoms * 4 * sizeof(double));
}

int host_spline(bool uniform, double x0, double dx, int n, const double *kx, const double *ydys, int m,
                const double *xs, double *ys)
{
    PP_REQUIRE(n >= 2, "a spline needs at least 2 knots");
    PP_REQUIRE(m >= 0, "m must be >= 0");
    if (m == 0) return 0;
    PP_REQUIRE(ydys && xs && ys && (uniform || kx), "null argument");
    Staged S(4);
    PP_E(S.put(0, ydys, (size_t)n * 2 * sizeof(double)));
    PP_E(S.put(1, xs, (size_t)m * sizeof(double)));
    PP_E(S.alloc(2, (size_t)m * sizeof(double)));
    if (uniform) {
        PP_E(launch_spline_uniform(x0, dx, n, S.p<double>(0), m, S.p<double>(1), S.p<double>(2), nullptr));
    } else {
        PP_E(S.put(3, kx, (size_t)n * sizeof(double)));
        PP_E(launch_spline_nonuniform(n, S.p<double>(3), S.p<double>(0), m, S.p<double>(1), S.p<double>(2), nullptr));
    }
    PP_TRY_HIP(hipGetLastError());
    return S.get(2, ys, (size_t)m * sizeof(double));
}

int host_test_force(int type, int n, const double *r0, const double *dr, const double *R, double *fs,
                    const ppafm_spline_params *sp)
{
    PP_REQUIRE(n >= 0, "n must be >= 0");
    if (n == 0) return 0;
    PP_REQUIRE(r0 && dr && R && fs, "null argument");
    RefState &RS = ref();
    double krad, lrad;
    {
        std::lock_guard<std::mutex> lk(RS.mu);
        krad = RS.krad;
        lrad = RS.lrad;
    }
    Staged S(2);
    int sn = 0;
    if (type == 1) {
        PP_REQUIRE(sp && sp->rff_n >= 2 && sp->rff_xs && sp->rff_ydys, "test_force type 1 needs a tip spline");
        sn = sp->rff_n;
        PP_E(S.alloc(0, (size_t)sn * 3 * sizeof(double)));
        PP_TRY_HIP(hipMemcpy(S.p<double>(0), sp->rff_xs, (size_t)sn * sizeof(double), hipMemcpyHostToDevice));
        PP_TRY_HIP(hipMemcpy(S.p<double>(0) + sn, sp->rff_ydys, (size_t)sn * 2 * sizeof(double), hipMemcpyHostToDevice));
    } else {
        PP_E(S.alloc(0, 8));
    }
    PP_E(S.alloc(1, (size_t)n * 3 * sizeof(double)));
    PP_E(launch_test_force(type, n, r0, dr, R, krad, lrad, sn, S.p<double>(0), S.p<double>(0) + sn, S.p<double>(1),
                           nullptr));
    PP_TRY_HIP(hipGetLastError());
    return S.get(1, fs, (size_t)n * 3 * sizeof(double));
}

int host_eigen3x3(const double *mat, double *evs)
{
    PP_REQUIRE(mat && evs, "null argument");
    Staged S(2);
    PP_E(S.put(0, mat, 9 * sizeof(double)));
    PP_E(S.alloc(1, 12 * sizeof(double)));
    PP_E(launch_eigen3x3(S.p<double>(0), S.p<double>(1), nullptr));
    PP_TRY_HIP(hipGetLastError());
    return S.get(1, evs, 12 * sizeof(double));
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

void getLennardJonesFF(int natoms, double *Rs, double *cLJs)
{
    report(ref_accum(kModelLJ, 0.0, natoms, Rs, cLJs, 2), "getLennardJonesFF");
}

void getCoulombFF(int natoms, double *Rs, double *kQQs, int kind)
{
    if (kind < 0 || kind > 2) {
        set_error("getCoulombFF: kind must be 0 (s), 1 (pz) or 2 (dz2)");
        report(-2, "getCoulombFF");
        return;
    }
    report(ref_accum(kModelCoulS + kind, 0.0, natoms, Rs, kQQs, 1), "getCoulombFF");
}

void getVdWFF(int natoms, double *Rs, double *cLJs)
{
    report(ref_accum(kModelVdW, ref().adamp_const, natoms, Rs, cLJs, 2), "getVdWFF");
}

// Kind 1..4 selects the damping function (R2, R4, invR4, invR8); other kinds
// add nothing, as in the reference.
void getVdWFF_RE(int natoms, double *Rs, double *REs, int kind, double ADamp)
{
    RefState &S = ref();
    double *slot[5] = {nullptr, &S.adamp_r2, &S.adamp_r4, &S.adamp_invr4, &S.adamp_invr8};
    if (kind < 1 || kind > 4) return;
    {
        std::lock_guard<std::mutex> lk(S.mu);
        if (ADamp > 0) *slot[kind] = ADamp;
    }
    report(ref_accum(kModelVdWR2 + kind - 1, *slot[kind], natoms, Rs, REs, 2), "getVdWFF_RE");
}

void getDFTD3FF(int natoms, double *Rs, double *d3_coeffs)
{
    report(ref_accum(kModelD3, 0.0, natoms, Rs, d3_coeffs, 4), "getDFTD3FF");
}

void getMorseFF(int natoms, double *Rs, double *REs, double alpha)
{
    report(ref_accum(kModelMorse, alpha, natoms, Rs, REs, 2), "getMorseFF");
}

void getGaussDensity(int natoms, double *Rs, double *cRAs)
{
    report(ref_accum(kModelGauss, 0.0, natoms, Rs, cRAs, 2, true), "getGaussDensity");
}

void getSlaterDensity(int natoms, double *Rs, double *cRAs)
{
    report(ref_accum(kModelSlater, 0.0, natoms, Rs, cRAs, 2, true), "getSlaterDensity");
}

void getDensityR4spline(int natoms, double *Rs, double *cRAs)
{
    report(ref_accum(kModelR4, 0.0, natoms, Rs, cRAs, 2, true), "getDensityR4spline");
}

void getInPoints_LJ(int npoints, double *points, double *FEs, int natoms, double *Rs, double *cLJs)
{
    report(host_points_lj(npoints, points, FEs, natoms, Rs, cLJs), "getInPoints_LJ");
}

void evalRadialFF(int n, double *rs, double *coefs, double *Es, double *Fs, int kind, double ADamp)
{
    report(host_radial(n, rs, coefs, Es, Fs, kind, ADamp), "evalRadialFF");
}

void computeD3Coeffs(int natoms, const double *rs, const int *elems, const double *r_cov, const double *r_cut,
                     const double *ref_cn, const double *ref_c6, const double *r4r2, const double *k,
                     const double *params, int elem_pp, double *d3_coeffs)
{
    (void)r_cut;  // passed by ppafm/core.py but not used by the reference either
    report(host_d3_coeffs(natoms, rs, elems, r_cov, ref_cn, ref_c6, r4r2, k, params, elem_pp, d3_coeffs),
           "computeD3Coeffs");
}

void subsample_uniform_spline(double x0, double dx, int n, double *ydys, int m, double *xs, double *ys)
{
    report(host_spline(true, x0, dx, n, nullptr, ydys, m, xs, ys), "subsample_uniform_spline");
}

void subsample_nonuniform_spline(int n, double *xs, double *ydys, int m, double *xs_, double *ys_)
{
    report(host_spline(false, 0.0, 0.0, n, xs, ydys, m, xs_, ys_), "subsample_nonuniform_spline");
}

void test_force(int type, int n, double *r0, double *dr, double *R, double *fs, ppafm_spline_params *splineParams)
{
    report(host_test_force(type, n, r0, dr, R, fs, splineParams), "test_force");
}

void test_eigen3x3(double *mat, double *evs) { report(host_eigen3x3(mat, evs), "test_eigen3x3"); }

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

int ppafm_xsf_write_data_dev(const char *fname, int append, const void *grid_dev, int f64, int channel, int nx, int ny,
                             int nz, int nthreads)
{
    PP_REQUIRE(grid_dev && fname, "null argument");
    PP_REQUIRE(channel >= 0 && channel < 4, "channel must be 0..3");
    PP_REQUIRE(nx > 0 && ny > 0 && nz > 0, "empty grid");
    const size_t npts = (size_t)nx * ny * nz;
    std::vector<double> h(npts);
    if (f64) {
        std::vector<double> g(npts * 4);
        PP_TRY_HIP(hipMemcpy(g.data(), grid_dev, npts * 4 * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < npts; i++) h[i] = g[4 * i + channel];
    } else {
        std::vector<float> g(npts * 4);
        PP_TRY_HIP(hipMemcpy(g.data(), grid_dev, npts * 4 * sizeof(float), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < npts; i++) h[i] = g[4 * i + channel];
    }
    return ppafm_xsf_write_data(fname, append, h.data(), nx, ny, nz, (long long)ny * nz, nz, 1, nthreads);
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

int ppafm_memcpy_d2d(void *dst, const void *src, unsigned long long bytes, void *stream)
{
    if (!bytes) return 0;
    if (stream) PP_TRY_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)));
    else PP_TRY_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
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
