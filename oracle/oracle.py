"""FP64 CPU oracle for the probe-particle hot path -- TEST INFRASTRUCTURE.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may
import this module, and only as the checker (or the timed CPU baseline), never
as the thing measured or shipped.  The product path lives in ``ppafm_amd`` and
runs on the GPU.

Two back ends:
  * ``liboracle.so`` built from ``pp_oracle.c`` (the clean FP64 restatement;
    citations are in that file's header);
  * ``_ref/ProbeParticle_lib.so``, the reference C++ compiled out of tree from
    the reference sources by ``make ref`` (only where the reference is present).

The Fz -> df pieces are numpy restatements:
  * :func:`get_df_weight`        -- ppafm/common.py:333-392 (Giessibl 2001 exact
    piecewise-linear weights, n = ceil(A/dz), n+1 taps).  Pinned by the golden
    vectors of tests/test_common.py:10-19 and by its closed form.
  * :func:`get_simple_df_weight` -- ppafm/common.py:395-418.  Pinned by the
    golden vector of tests/test_common.py:22-24.
  * :func:`fz2df`                -- ppafm/common.py:421-430 (valid-mode
    np.convolve along z, times units*f0/k0).  Textbook convolution; pinned by a
    brute-force loop in the tests.
"""

import ctypes
import os
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
REF_PATH = HERE / "_ref" / "ProbeParticle_lib.so"

eVA_Nm = 16.0217662  # common.py:20
CoulombConst = 14.399645  # common.py:21

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _ptr(a, t=_dp):
    if a is None:
        return None
    return a.ctypes.data_as(t)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


_lib = None


def lib():
    """Load liboracle.so (build it with ``make -C oracle``)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        _lib = load(LIB_PATH)
    return _lib


def load(path):
    """Load a build of pp_oracle.c from ``path`` with its argument types (the
    tests load copies built with planted mistakes this way)."""
    L = ctypes.CDLL(str(path))
    L.orc_grid_shape.argtypes = [_ip, _dp, _dp, _dp]
    L.orc_lj_ff.argtypes = [_ip, _dp, ctypes.c_int, _dp, _dp, _dp, _dp]
    L.orc_coulomb_ff.argtypes = [_ip, _dp, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, _dp]
    L.orc_limit_vec_field.argtypes = [ctypes.c_long, _dp, ctypes.c_double]
    L.orc_interp.argtypes = [_ip, _dp, _dp, ctypes.c_int, _dp, _dp]
    L.orc_relax_stroke.argtypes = [_ip, _dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _ip, _ip]
    L.orc_relax_stroke.restype = ctypes.c_int
    L.orc_relax_strokes.argtypes = [ctypes.c_int] * 3 + [_ip, _dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _ip, _ip]
    L.orc_relax_strokes.restype = ctypes.c_long
    L.orc_relax_strokes_subset.argtypes = [ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int, _ip, _dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _ip]
    L.orc_relax_strokes_subset.restype = ctypes.c_long
    L.orc_ljc_qzs_ff.argtypes = [_ip, _dp, _dp, _dp, _dp, ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
    L.orc_relax_strokes_tilted.argtypes = [
        ctypes.c_int, ctypes.c_int, _ip, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_double,
        ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp, _ip,
    ]
    L.orc_relax_strokes_tilted_warm.argtypes = [
        ctypes.c_int, ctypes.c_int, _ip, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_double,
        ctypes.c_double, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
        ctypes.c_double, ctypes.c_double, _dp, _dp, _ip, _ip,
    ]
    L.orc_relax_strokes_tilted_warm.restype = ctypes.c_long
    L.orc_tilted_force_jac.argtypes = [_ip, _dp, _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int, _dp, _dp, _dp, _dp]
    L.orc_convolve_z.argtypes = [ctypes.c_long, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
    L.orc_stable3.argtypes = [_dp]
    L.orc_stable3.restype = ctypes.c_int
    L.orc_force_jac.argtypes = [_ip, _dp, _dp, _dp, ctypes.c_int, _dp, _dp, _dp, _dp]
    L.orc_relax_strokes_warm.argtypes = [ctypes.c_int] * 3 + [_ip, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int,
                                                              ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                                              ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp,
                                                              _dp, _ip, _ip]
    L.orc_relax_strokes_warm.restype = ctypes.c_long
    L.orc_stiffness.argtypes = [_ip, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_long, _dp, _dp, _dp,
                                _dp, _dp, _dp, _dp]
    L.orc_stiffness.restype = ctypes.c_long
    L.orc_model_ff.argtypes = [ctypes.c_int, ctypes.c_double, _ip, _dp, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, _dp]
    L.orc_radial.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp, _dp, _dp, _dp]
    L.orc_relax_strokes_spline.argtypes = [ctypes.c_int] * 3 + [_ip, _dp, _dp, _dp, _dp, ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
    L.orc_relax_strokes_spline.restype = ctypes.c_long
    L.orc_stiffness_spline.argtypes = [_ip, _dp, _dp, _dp, ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_int,
                                       ctypes.c_long, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.orc_stiffness_spline.restype = ctypes.c_long
    L.orc_num_threads.restype = ctypes.c_int
    return L


def num_threads():
    return lib().orc_num_threads()


# ------------------------------------------------------------------ geometry


def grid_shape(n, cell):
    n = np.ascontiguousarray(n, dtype=np.int32)
    cell = _f64(cell, (3, 3))
    dcell = np.zeros((3, 3))
    dinv = np.zeros((3, 3))
    lib().orc_grid_shape(_ptr(n, _ip), _ptr(cell), _ptr(dcell), _ptr(dinv))
    return dcell, dinv


# ---------------------------------------------------------------- grid build


def lj_ff(n, cell, Rs, cLJs, F=None, E=None):
    """Accumulate the LJ grid (ProbeParticle.cpp:508).  Rs relative to the grid origin."""
    n = np.ascontiguousarray(n, dtype=np.int32)
    Rs = _f64(Rs, (-1, 3))
    cLJs = _f64(cLJs, (-1, 2))
    if F is None:
        F = np.zeros((n[0], n[1], n[2], 3))
    lib().orc_lj_ff(_ptr(n, _ip), _ptr(_f64(cell, (3, 3))), len(Rs), _ptr(Rs), _ptr(cLJs), _ptr(F), _ptr(E))
    return F


def coulomb_ff(n, cell, Rs, kQQs, kind=0, F=None, E=None):
    """Accumulate the unit-probe-charge Coulomb grid (ProbeParticle.cpp:540)."""
    n = np.ascontiguousarray(n, dtype=np.int32)
    Rs = _f64(Rs, (-1, 3))
    kQQs = _f64(kQQs, (-1,))
    if F is None:
        F = np.zeros((n[0], n[1], n[2], 3))
    lib().orc_coulomb_ff(_ptr(n, _ip), _ptr(_f64(cell, (3, 3))), len(Rs), _ptr(Rs), _ptr(kQQs), int(kind), _ptr(F), _ptr(E))
    return F


def limit_vec_field(F, Fmax=10.0):
    assert F.dtype == np.float64 and F.flags.c_contiguous
    lib().orc_limit_vec_field(F.size // 3, _ptr(F), float(Fmax))
    return F


def interp(n, cell, F, rs):
    n = np.ascontiguousarray(n, dtype=np.int32)
    rs = _f64(rs, (-1, 3))
    out = np.zeros_like(rs)
    lib().orc_interp(_ptr(n, _ip), _ptr(_f64(cell, (3, 3))), _ptr(F), len(rs), _ptr(rs), _ptr(out))
    return out


# ---------------------------------------------------------------- relaxation

DEFAULT_RELAX = (1000, 1.0e-8, 0.1, 0.1, 1.1, 0.5, 0.99)  # ProbeParticle.cpp:76-103


def tip_params(lRadial, kRadial, rPP0, kSpring):
    """Pack setTip arguments (ProbeParticle.cpp:371)."""
    return np.array([lRadial, kRadial, rPP0[0], rPP0[1], rPP0[2], kSpring[0], kSpring[1], kSpring[2]], dtype=np.float64)


def relax_strokes(n, cell, F, tip, rTips, relax=DEFAULT_RELAX, probeStart=1, relaxAlg=1, want_iters=True):
    """relaxTipStrokes_omp semantics on (nx,ny,nz,3) tip positions.

    Returns rs, fs, iters (per position, or None), total iteration sum."""
    rTips = _f64(rTips)
    nx, ny, nz, _ = rTips.shape
    rs = np.zeros_like(rTips)
    fs = np.zeros_like(rTips)
    iters = np.zeros((nx, ny, nz), dtype=np.int32) if want_iters else None
    n = np.ascontiguousarray(n, dtype=np.int32)
    F = _f64(F)
    tot = lib().orc_relax_strokes(
        nx, ny, nz, _ptr(n, _ip), _ptr(_f64(cell, (3, 3))), _ptr(F), _ptr(_f64(tip)), _ptr(_f64(relax)),
        int(probeStart), int(relaxAlg), _ptr(rTips), _ptr(rs), _ptr(fs), _ptr(iters, _ip), None,
    )
    return rs, fs, iters, int(tot)


def stable3(J):
    """1 when every eigenvalue of the 3x3 matrix J has a negative real part."""
    return bool(lib().orc_stable3(_ptr(_f64(J, (3, 3)))))


def force_jac(n, cell, F, tip, rtips, rs):
    """Total probe force and its Jacobian dF/dr (trilinear field + springs)
    at positions rs (m, 3) for tips rtips (m, 3): (m, 3), (m, 3, 3)."""
    rtips = _f64(rtips, (-1, 3))
    rs = _f64(rs, (-1, 3))
    f = np.zeros_like(rs)
    J = np.zeros((len(rs), 3, 3))
    n = np.ascontiguousarray(n, dtype=np.int32)
    lib().orc_force_jac(_ptr(n, _ip), _ptr(_f64(cell, (3, 3))), _ptr(_f64(F)), _ptr(_f64(tip)), len(rs), _ptr(rtips),
                        _ptr(rs), _ptr(f), _ptr(J))
    return f, J


# Oracle model ids (pp_oracle.c pair_model) of the reference's grid builders
# and evalRadialFF kinds.
MODEL = dict(vdw=4, vdw_r2=5, vdw_r4=6, vdw_invr4=7, vdw_invr8=8, d3=9, morse=10, gauss=11, slater=12, r4=13,
             lj_re=14, inv_r6=15, lj=0, vdw_nodamp_ocl=16, vdw_invr4_ocl=17, vdw_invr8_ocl=18)
RADIAL_KIND = {-3: "lj_re", -2: "lj", -1: "inv_r6", 0: "vdw", 1: "vdw_r2", 2: "vdw_r4", 3: "vdw_invr4", 4: "vdw_invr8"}
ADAMP_DEFAULT = {"vdw": 180.0, "vdw_r2": 0.5, "vdw_r4": 0.5, "vdw_invr4": 0.03, "vdw_invr8": 0.01}  # ProbeParticle.cpp:48


def model_ff(name, n, cell, Rs, coefs, par=0.0, F=None, E=None):
    """Accumulate one of the further grid models (getVdWFF, getVdWFF_RE,
    getDFTD3FF, getMorseFF, the densities; ProbeParticle.cpp:515-631) into F
    (nx,ny,nz,3) and E (nx,ny,nz); Rs relative to the grid origin."""
    n = np.ascontiguousarray(n, dtype=np.int32)
    Rs = _f64(Rs, (-1, 3))
    coefs = _f64(coefs)
    coefs = coefs.reshape(len(Rs), -1)
    if F is None:
        F = np.zeros((n[0], n[1], n[2], 3))
    lib().orc_model_ff(MODEL[name], float(par), _ptr(n, _ip), _ptr(_f64(cell, (3, 3))), len(Rs), _ptr(Rs), _ptr(coefs),
                       coefs.shape[1], _ptr(F), _ptr(E))
    return F


def radial(kind, rs, coefs, adamp=None):
    """evalRadialFF (ProbeParticle.cpp:556-575): Es, Fs at (r,0,0); adamp None
    = the kind's default constant."""
    name = RADIAL_KIND[kind]
    par = ADAMP_DEFAULT.get(name, 0.0) if adamp is None else adamp
    rs = _f64(rs)
    c = np.zeros(4)
    cf = _f64(coefs).ravel()
    c[: len(cf)] = cf
    Es, Fs = np.zeros(len(rs)), np.zeros(len(rs))
    lib().orc_radial(MODEL[name], float(par), len(rs), _ptr(rs), _ptr(c), _ptr(Es), _ptr(Fs))
    return Es, Fs


def relax_strokes_spline(n, cell, F, tip, rTips, spline, relax=DEFAULT_RELAX):
    """relaxTipStrokes_omp (probeStart 1, FIRE) with the radial tip force of
    spline = (xs, ydys) (forceRSpline, ProbeParticle.cpp:211-224)."""
    rTips = _f64(rTips)
    nx, ny, nz, _ = rTips.shape
    rs, fs = np.zeros_like(rTips), np.zeros_like(rTips)
    xs, ydys = _f64(spline[0]), _f64(spline[1])
    n = np.ascontiguousarray(n, dtype=np.int32)
    lib().orc_relax_strokes_spline(nx, ny, nz, _ptr(n, _ip), _ptr(_f64(cell, (3, 3))), _ptr(_f64(F)), _ptr(_f64(tip)),
                                   _ptr(_f64(relax)), len(xs), _ptr(xs), _ptr(ydys), _ptr(rTips), _ptr(rs), _ptr(fs))
    return rs, fs


def hermite(x, y0, y1, dy0, dy1):
    """Spline_Hermite::val (spline_hermite.h:45-51)."""
    y01 = y0 - y1
    return y0 + x * (dy0 + x * (-3 * y01 - 2 * dy0 - dy1 + x * (2 * y01 + dy0 + dy1)))


def find_index(i, di, x, xs):
    """Spline_Hermite::find_index (spline_hermite.h:159-168)."""
    while di > 1:
        di >>= 1
        if xs[i + di] < x:
            i += di
    return i


def spline_uniform(x0, dx, ydys, xs):
    """subsample_uniform_spline (ProbeParticle.cpp:752-763); samples beyond
    the last interval extrapolate it."""
    n = len(ydys)
    ys = np.zeros(len(xs))
    for j, x in enumerate(xs):
        u = (x - x0) * (1 / dx)
        i = min(max(int(u), 0), n - 2)
        u -= i
        ys[j] = hermite(u, ydys[i, 0], ydys[i + 1, 0], ydys[i, 1] * dx, ydys[i + 1, 1] * dx)
    return ys


def spline_nonuniform(kx, ydys, xs):
    """subsample_nonuniform_spline (ProbeParticle.cpp:765-784): the interval
    found for one sample is kept while later samples stay below its end."""
    n = len(kx)
    ys = np.zeros(len(xs))
    i, x0, x1, dx, denom = 0, 0.0, -1e300, 1.0, 1.0
    for j, x in enumerate(xs):
        if x > x1:
            i = min(find_index(i, n - i, x, kx), n - 2)
            x0, x1 = kx[i], kx[i + 1]
            dx = x1 - x0
            denom = 1 / dx
        u = (x - x0) * denom
        ys[j] = hermite(u, ydys[i, 0], ydys[i + 1, 0], ydys[i, 1] * dx, ydys[i + 1, 1] * dx)
    return ys


def d3_coeffs(Rs, iZs, tables, params, iZPP):
    """computeD3Coeffs (ProbeParticle.cpp:397-506) in numpy: coordination
    numbers, Gaussian reference weights (kept in FP32 as there), C6, C8 and
    the damping radii; the probe's unused reference slots get weight 0."""
    Rs = _f64(Rs, (-1, 3))
    k1, k2, k3 = tables["K"][0], tables["K"][1], -tables["K"][2]
    k12 = -k1 * k2
    s6, s8, a1, a2 = params
    rc, rcn, c6t, r4r2 = tables["R_COV"], np.asarray(tables["REF_CN"]).reshape(94, 5), np.asarray(tables["REF_C6"]).reshape(94, 94, 5, 5), tables["R4R2"]
    pp = iZPP - 1
    lpp = np.array([np.exp(k3 * c * c) if c >= 0.0 else 0.0 for c in rcn[pp]])
    out = np.zeros((len(Rs), 4))
    for ia in range(len(Rs)):
        el = iZs[ia] - 1
        cn = 0.0
        for j in range(len(Rs)):
            if j == ia:
                continue
            d = np.sqrt(np.sum((Rs[j] - Rs[ia]) ** 2))
            cn += 1.0 / (1.0 + np.exp(k12 * (rc[iZs[j] - 1] + rc[el]) / d + k1))
        L = np.zeros((5, 5), dtype=np.float32)
        norm = 0.0
        na = 0
        while na < 5 and rcn[el, na] >= 0.0:
            la = np.exp(k3 * (rcn[el, na] - cn) ** 2)
            for b in range(5):
                norm += la * lpp[b]
                L[na, b] = la * lpp[b]
            na += 1
        if norm == 0 and na > 0:
            for b in range(5):
                norm += lpp[b]
                L[na - 1, b] = lpp[b]
        c6 = 0.0
        for a in range(na):
            for b in range(5):
                c6 += float(L[a, b]) * c6t[el, pp, a, b]
        c6 /= norm
        qq = 3 * r4r2[el] * r4r2[pp]
        r0 = a1 * np.sqrt(qq) + a2
        out[ia] = (c6 * s6, qq * c6 * s8, r0 ** 6, r0 ** 8)
    return out


def stiffness(n, cell, F, tip, rTips, rPPs, which=0, ddisp=0.05, want_dynmat=False, spline=None):
    """stiffnessMatrix (ProbeParticle.cpp:703-750) at positions rPPs (m,3)
    for tips rTips (m,3): eigenvalues (m,3) ascending, the list of three
    eigenvector arrays (the first ``which`` filled, others None), the number
    of NaN points and, with want_dynmat, the symmetric matrices (m,3,3)."""
    rTips = _f64(rTips, (-1, 3))
    rPPs = _f64(rPPs, (-1, 3))
    m = len(rTips)
    ev = np.zeros((m, 3))
    vecs = [np.zeros((m, 3)) if k < which else None for k in range(3)]
    D = np.zeros((m, 3, 3)) if want_dynmat else None
    n = np.ascontiguousarray(n, dtype=np.int32)
    sxs, sydys = (None, None) if spline is None else (_f64(spline[0]), _f64(spline[1]))
    nnan = lib().orc_stiffness_spline(_ptr(n, _ip), _ptr(_f64(cell, (3, 3))), _ptr(_f64(F)), _ptr(_f64(tip)),
                                      0 if spline is None else len(sxs), _ptr(sxs), _ptr(sydys), float(ddisp),
                                      int(which), m, _ptr(rTips), _ptr(rPPs), _ptr(ev), *[_ptr(v) for v in vecs], _ptr(D))
    out = (ev, vecs, int(nnan))
    return out + (D,) if want_dynmat else out


def relax_strokes_warm(n, cell, F, tip, rTips, starts0=None, newton_max=16, newton_max0=12, retry=0, tol=1e-3,
                       cap=0.3, relax=DEFAULT_RELAX, extrap=1.0, res_max=0.0):
    """Warm-start strokes (Newton from the continuation start, then FIRE
    with a Newton try every ``retry`` iterations; probeStart 1) on
    (nx,ny,nz,3) tip positions; starts0 (nx,ny,3) optional step-0 estimates;
    res_max > 0: a column whose total force at its estimate exceeds res_max
    runs step 0 as plain FIRE from the plain start.  Returns rs, fs, iters,
    nevals (per position) and the number of fallback steps."""
    rTips = _f64(rTips)
    nx, ny, nz, _ = rTips.shape
    rs = np.zeros_like(rTips)
    fs = np.zeros_like(rTips)
    iters = np.zeros((nx, ny, nz), dtype=np.int32)
    nev = np.zeros((nx, ny, nz), dtype=np.int32)
    n = np.ascontiguousarray(n, dtype=np.int32)
    st = None if starts0 is None else _f64(starts0, (nx, ny, 3))
    nfb = lib().orc_relax_strokes_warm(nx, ny, nz, _ptr(n, _ip), _ptr(_f64(cell, (3, 3))), _ptr(_f64(F)), _ptr(_f64(tip)),
                                       _ptr(_f64(relax)), _ptr(rTips), None if st is None else _ptr(st), int(newton_max),
                                       int(newton_max0), int(retry), float(tol), float(cap), float(extrap),
                                       float(res_max), _ptr(rs),
                                       _ptr(fs), _ptr(iters, _ip),
                                       _ptr(nev, _ip))
    return rs, fs, iters, nev, int(nfb)


def relax_strokes_subset(cols, n, cell, F, tip, rTips, relax=DEFAULT_RELAX, probeStart=1, relaxAlg=1):
    """Relax only the listed (ix,iy) columns (for a bounded CPU timing sample)."""
    rTips = _f64(rTips)
    nx, ny, nz, _ = rTips.shape
    cols = np.ascontiguousarray(cols, dtype=np.int32).reshape(-1, 2)
    rs = np.zeros_like(rTips)
    fs = np.zeros_like(rTips)
    iters = np.zeros((nx, ny, nz), dtype=np.int32)
    n = np.ascontiguousarray(n, dtype=np.int32)
    tot = lib().orc_relax_strokes_subset(
        len(cols), _ptr(cols, _ip), nx, nz, _ptr(n, _ip), _ptr(_f64(cell, (3, 3))), _ptr(_f64(F)), _ptr(_f64(tip)),
        _ptr(_f64(relax)), int(probeStart), int(relaxAlg), _ptr(rTips), _ptr(rs), _ptr(fs), _ptr(iters, _ip),
    )
    return rs, fs, iters, int(tot)


# ----------------------------------------------------------- OpenCL-mode path


def ljc_qzs_ff(nDim, lvec, atoms, cLJs, Qs, QZs):
    """evalLJC_QZs_noPos in FP64; returns FE (nx,ny,nz,4) in C order of (iz,iy,ix)
    reshaped to (nx,ny,nz,4) x-fastest semantics: FE[ix,iy,iz]."""
    nx, ny, nz = (int(v) for v in nDim[:3])
    lvec = _f64(lvec, (4, 3))
    d = np.array([lvec[1] / nx, lvec[2] / ny, lvec[3] / nz])
    out = np.zeros((nz, ny, nx, 4))
    n = np.array([nx, ny, nz], dtype=np.int32)
    atoms = _f64(atoms, (-1, 4))
    lib().orc_ljc_qzs_ff(
        _ptr(n, _ip), _ptr(_f64(lvec[0])), _ptr(_f64(d[0])), _ptr(_f64(d[1])), _ptr(_f64(d[2])), len(atoms), _ptr(atoms),
        _ptr(_f64(cLJs, (-1, 2))), _ptr(_f64(Qs)), _ptr(_f64(QZs)), _ptr(out),
    )
    return out.transpose(2, 1, 0, 3).copy()


def relax_tilted(FE, dinv, poss, nz, tipRot, stiffness, dpos0, relax_params, surfFF=0.0, max_iters=128, conv_f2=1e-8, f2safe=1e-8,
                 _lib=None):
    """relaxStrokesTilted semantics in FP64 with exact trilinear sampling.

    FE (nx,ny,nz_g,4) indexed FE[ix,iy,iz]; dinv (3,3) maps a position relative
    to the grid origin to grid-index coordinates; poss (npix,4); tipRot (3,4).
    Returns FEs, paths (npix, nz, 4) and nevals (npix, nz)."""
    nx, ny, nzg = FE.shape[:3]
    fe_x = np.ascontiguousarray(np.asarray(FE, dtype=np.float64).transpose(2, 1, 0, 3))
    poss = _f64(poss, (-1, 4))
    npix = len(poss)
    FEs = np.zeros((npix, nz, 4))
    paths = np.zeros((npix, nz, 4))
    nevals = np.zeros((npix, nz), dtype=np.int32)
    n = np.array([nx, ny, nzg], dtype=np.int32)
    (_lib or lib()).orc_relax_strokes_tilted(
        npix, int(nz), _ptr(n, _ip), _ptr(fe_x), _ptr(_f64(dinv, (3, 3))), _ptr(_f64(tipRot, (3, 4))), _ptr(_f64(stiffness)),
        _ptr(_f64(dpos0)), _ptr(_f64(relax_params[:2])), float(surfFF), int(max_iters), float(conv_f2), float(f2safe),
        _ptr(poss), _ptr(FEs), _ptr(paths), _ptr(nevals, _ip),
    )
    return FEs, paths, nevals


def relax_tilted_warm(FE, dinv, poss, nz, tipRot, stiffness, dpos0, relax_params, surfFF=0.0, max_iters=128,
                      conv_f2=1e-8, f2safe=1e-8, starts0=None, newton_max=16, newton_max0=12, retry=0, tol=1e-3, cap=0.3,
                      extrap=1.0, res_max=0.0):
    """The warm method (see relax_strokes_warm) on relaxStrokesTilted
    semantics, arguments as relax_tilted; starts0 (npix, 3) optional step-0
    estimates (res_max: the residual gate of relax_strokes_warm).  Returns
    FEs, paths (npix, nz, 4), FIRE iterations and force evaluations
    (npix, nz) and the number of fallback steps."""
    nx, ny, nzg = FE.shape[:3]
    fe_x = np.ascontiguousarray(np.asarray(FE, dtype=np.float64).transpose(2, 1, 0, 3))
    poss = _f64(poss, (-1, 4))
    npix = len(poss)
    FEs = np.zeros((npix, nz, 4))
    paths = np.zeros((npix, nz, 4))
    iters = np.zeros((npix, nz), dtype=np.int32)
    nev = np.zeros((npix, nz), dtype=np.int32)
    n = np.array([nx, ny, nzg], dtype=np.int32)
    st = None if starts0 is None else _f64(starts0, (npix, 3))
    nfb = lib().orc_relax_strokes_tilted_warm(
        npix, int(nz), _ptr(n, _ip), _ptr(fe_x), _ptr(_f64(dinv, (3, 3))), _ptr(_f64(tipRot, (3, 4))), _ptr(_f64(stiffness)),
        _ptr(_f64(dpos0)), _ptr(_f64(relax_params[:2])), float(surfFF), int(max_iters), float(conv_f2), float(f2safe),
        _ptr(poss), None if st is None else _ptr(st), int(newton_max), int(newton_max0), int(retry), float(tol), float(cap),
        float(extrap), float(res_max), _ptr(FEs), _ptr(paths), _ptr(iters, _ip), _ptr(nev, _ip),
    )
    return FEs, paths, iters, nev, int(nfb)


def tilted_force_jac(FE, dinv, tipRot, stiffness, dpos0, tips, rs, surfFF=0.0):
    """Total probe force and its Jacobian in relaxStrokesTilted semantics at
    positions rs (m, 3) for tips (m, 3): (m, 3), (m, 3, 3)."""
    nx, ny, nzg = FE.shape[:3]
    fe_x = np.ascontiguousarray(np.asarray(FE, dtype=np.float64).transpose(2, 1, 0, 3))
    tips = _f64(tips, (-1, 3))
    rs = _f64(rs, (-1, 3))
    f = np.zeros_like(rs)
    J = np.zeros((len(rs), 3, 3))
    n = np.array([nx, ny, nzg], dtype=np.int32)
    lib().orc_tilted_force_jac(_ptr(n, _ip), _ptr(fe_x), _ptr(_f64(dinv, (3, 3))), _ptr(_f64(tipRot, (3, 4))),
                               _ptr(_f64(stiffness)), _ptr(_f64(dpos0)), float(surfFF), len(rs), _ptr(tips), _ptr(rs),
                               _ptr(f), _ptr(J))
    return f, J


def convolve_z(F, w):
    """out[..., j] = sum_k F[..., j+k] w[k] (convolveZ, relax.cl:571-608)."""
    F = _f64(F)
    w = _f64(w)
    shp = F.shape
    nzin = shp[-1]
    ncol = F.size // nzin
    out = np.zeros(shp[:-1] + (nzin - len(w) + 1,))
    lib().orc_convolve_z(ncol, nzin, len(w), _ptr(F), _ptr(w), _ptr(out))
    return out


# ---------------------------------------------------------------- Fz -> df


def get_df_weight(Amp, dz=0.1):
    """Giessibl exact weights (common.py:333-392), restated.

    t_k runs from -1 + 2 n dz / A down to -1 (n+1 nodes, n = ceil(A/dz)).
    With f(t) = sqrt(1-t^2) and f2(t) = (arccos(-t) - t f(t))/2 on interior
    nodes, f = 0 and f2 = pi/2 at the first node (t >= 1 treated as t = 1) and
    f = f2 = 0 at the last node, each interval [t_{k+1}, t_k] adds
    t_k (f_k - f_{k+1}) + (f2_k - f2_{k+1}) to w[k+1] and subtracts
    t_{k+1} (f_k - f_{k+1}) + (f2_k - f2_{k+1}) from w[k]; all over pi dz."""
    n = int(np.ceil(Amp / dz))
    t = np.linspace(-1 + 2 * n * dz / Amp, -1, n + 1)
    f = np.zeros(n + 1)
    f2 = np.zeros(n + 1)
    if n > 1:
        ti = t[1:-1]
        f[1:-1] = np.sqrt(1 - ti * ti)
        f2[1:-1] = (np.arccos(-ti) - ti * f[1:-1]) / 2.0
    f2[0] = np.pi / 2
    dfk = f[:-1] - f[1:]
    df2k = f2[:-1] - f2[1:]
    w = np.zeros(n + 1)
    w[1:] += t[:-1] * dfk + df2k
    w[:-1] -= t[1:] * dfk + df2k
    return w / (np.pi * dz)


def get_simple_df_weight(n=10, dz=0.1):
    """Simplified weights (common.py:395-418), restated."""
    n = max(int(n), 2)
    t = np.linspace(-1 + 0.5 / n, 1 - 0.5 / n, n - 1)
    s = np.sqrt(1 - t * t)
    w = np.zeros(n)
    w[:-1] += s
    w[1:] -= s
    tn = np.linspace(-1, 1, n)
    return w / (np.sum(tn * w) * n * dz)


def fz2df(F, dz, k0, f0, amplitude=1.0, units=16.0217656):
    """common.py:421-430: valid-mode np.convolve along the last axis."""
    W = get_df_weight(amplitude, dz=dz)
    return convolve_z(F, W[::-1]) * units * f0 / k0


# ------------------------------------------------------- compiled reference


class Reference:
    """Drives the reference C++ (oracle/_ref) through its own C ABI.

    The reference keeps global state (ProbeParticle.cpp:33-89), so one instance
    drives one configuration at a time.  Argument types follow core.py:33-324."""

    def __init__(self, path=REF_PATH):
        if not pathlib.Path(path).exists():
            raise RuntimeError(f"reference library missing: {path} (run `make -C oracle ref`)")
        L = ctypes.CDLL(str(path))
        L.setGridN.argtypes = [_ip]
        L.setGridCell.argtypes = [_dp]
        L.setFF_Fpointer.argtypes = [_dp]
        L.setFF_Epointer.argtypes = [_dp]
        L.deleteFF_Fpointer.argtypes = []
        L.deleteFF_Epointer.argtypes = []
        L.setTip.argtypes = [ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.setRelax.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.setFIRE.argtypes = [ctypes.c_double] * 3
        L.getLennardJonesFF.argtypes = [ctypes.c_int, _dp, _dp]
        L.getCoulombFF.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int]
        L.relaxTipStroke.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, ctypes.c_void_p]
        L.relaxTipStroke.restype = ctypes.c_int
        L.relaxTipStrokes_omp.argtypes = [ctypes.c_int] * 5 + [_dp, _dp, _dp, ctypes.c_void_p]
        L.relaxTipStrokes_omp.restype = ctypes.c_int
        L.stiffnessMatrix.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int] + [_dp] * 6 + [ctypes.c_void_p]
        L.stiffnessMatrix.restype = None
        for name in ("getVdWFF", "getDFTD3FF", "getGaussDensity", "getSlaterDensity", "getDensityR4spline"):
            getattr(L, name).argtypes = [ctypes.c_int, _dp, _dp]
        L.getVdWFF_RE.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_double]
        L.getMorseFF.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_double]
        L.getInPoints_LJ.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, _dp]
        L.evalRadialFF.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_double]
        L.computeD3Coeffs.argtypes = [ctypes.c_int, _dp, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int, _dp]
        L.subsample_uniform_spline.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp]
        L.subsample_nonuniform_spline.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, _dp]
        L.test_force.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_void_p]
        L.test_eigen3x3.argtypes = [_dp, _dp]
        self.L = L
        self._keep = []

    def set_grid(self, n, cell, F, E=None):
        n = np.ascontiguousarray(n, dtype=np.int32)
        cell = _f64(cell, (3, 3))
        self.L.setGridN(_ptr(n, _ip))
        self.L.setGridCell(_ptr(cell))
        self.L.setFF_Fpointer(_ptr(F))
        if E is not None:
            self.L.setFF_Epointer(_ptr(E))
        else:
            self.L.deleteFF_Epointer()
        self._keep = [n, cell, F, E]

    def set_tip(self, lRadial, kRadial, rPP0, kSpring):
        rPP0 = _f64(rPP0)
        kSpring = _f64(kSpring)
        self.L.setTip(float(lRadial), float(kRadial), _ptr(rPP0), _ptr(kSpring))

    def set_relax(self, maxIters=1000, convF2=1e-8, dt=0.1, damping=0.1):
        self.L.setRelax(int(maxIters), float(convF2), float(dt), float(damping))

    def lj(self, Rs, cLJs):
        Rs = _f64(Rs, (-1, 3))
        cLJs = _f64(cLJs, (-1, 2))
        self.L.getLennardJonesFF(len(Rs), _ptr(Rs), _ptr(cLJs))

    def coulomb(self, Rs, kQQs, kind=0):
        Rs = _f64(Rs, (-1, 3))
        kQQs = _f64(kQQs, (-1,))
        self.L.getCoulombFF(len(Rs), _ptr(Rs), _ptr(kQQs), int(kind))

    def relax_strokes(self, rTips, probeStart=1, relaxAlg=1):
        rTips = _f64(rTips)
        nx, ny, nz, _ = rTips.shape
        rs = np.zeros_like(rTips)
        fs = np.zeros_like(rTips)
        self.L.relaxTipStrokes_omp(nx, ny, int(probeStart), int(relaxAlg), nz, _ptr(rTips), _ptr(rs), _ptr(fs), None)
        return rs, fs

    def relax_stroke(self, rTips, probeStart=1, relaxAlg=1):
        rTips = _f64(rTips, (-1, 3))
        rs = np.zeros_like(rTips)
        fs = np.zeros_like(rTips)
        itr = self.L.relaxTipStroke(int(probeStart), int(relaxAlg), len(rTips), _ptr(rTips), _ptr(rs), _ptr(fs), None)
        return rs, fs, itr


    def stiffness(self, rTips, rPPs, which=0, ddisp=0.05, spline=None):
        """stiffnessMatrix on the current grid and tip (rTips, rPPs (m,3))."""
        rTips = _f64(rTips, (-1, 3))
        rPPs = _f64(rPPs, (-1, 3))
        m = len(rTips)
        ev = np.zeros((m, 3))
        vecs = [np.zeros((m, 3)) if k < which else None for k in range(3)]
        self.L.stiffnessMatrix(float(ddisp), int(which), m, _ptr(rTips), _ptr(rPPs), _ptr(ev), *[_ptr(v) for v in vecs],
                               self._spline(spline))
        return ev, vecs

    def _spline(self, spline):
        """TIP::SplineParams (ProbeParticle.cpp:58-62) for (xs, ydys) or None."""
        if spline is None:
            return None
        xs, ydys = _f64(spline[0]), _f64(spline[1])

        class SP(ctypes.Structure):
            _fields_ = [("n", ctypes.c_int), ("xs", _dp), ("ydys", _dp)]

        sp = SP(len(xs), _ptr(xs), _ptr(ydys))
        self._keep_spline = (xs, ydys, sp)
        return ctypes.cast(ctypes.pointer(sp), ctypes.c_void_p)

    def relax_strokes_spline(self, rTips, spline, probeStart=1, relaxAlg=1):
        rTips = _f64(rTips)
        nx, ny, nz, _ = rTips.shape
        rs = np.zeros_like(rTips)
        fs = np.zeros_like(rTips)
        self.L.relaxTipStrokes_omp(nx, ny, int(probeStart), int(relaxAlg), nz, _ptr(rTips), _ptr(rs), _ptr(fs), self._spline(spline))
        return rs, fs

    def grid_model(self, name, Rs, coefs, *extra):
        """One of the grid builders that take (natoms, Rs, coefs, ...)."""
        Rs = _f64(Rs, (-1, 3))
        coefs = _f64(coefs)
        getattr(self.L, name)(len(Rs), _ptr(Rs), _ptr(coefs), *extra)

    def points_lj(self, ps, Rs, cLJs):
        ps, Rs, cLJs = _f64(ps, (-1, 3)), _f64(Rs, (-1, 3)), _f64(cLJs, (-1, 2))
        FEs = np.zeros((len(ps), 4))
        self.L.getInPoints_LJ(len(ps), _ptr(ps), _ptr(FEs), len(Rs), _ptr(Rs), _ptr(cLJs))
        return FEs

    def radial(self, rs, coefs, kind, ADamp=-1.0):
        rs, coefs = _f64(rs), _f64(coefs)
        Es, Fs = np.zeros(len(rs)), np.zeros(len(rs))
        self.L.evalRadialFF(len(rs), _ptr(rs), _ptr(coefs), _ptr(Es), _ptr(Fs), int(kind), float(ADamp))
        return Es, Fs

    def d3_coeffs(self, Rs, iZs, tables, params, iZPP):
        Rs = _f64(Rs, (-1, 3))
        iZs = np.ascontiguousarray(iZs, dtype=np.int32)
        t = {k: _f64(v).ravel() for k, v in tables.items()}
        out = np.zeros(4 * len(Rs))
        self.L.computeD3Coeffs(len(Rs), _ptr(Rs), _ptr(iZs, _ip), _ptr(t["R_COV"]), _ptr(t["R0"]), _ptr(t["REF_CN"]),
                               _ptr(t["REF_C6"]), _ptr(t["R4R2"]), _ptr(t["K"]), _ptr(_f64(params)), int(iZPP), _ptr(out))
        return out.reshape(-1, 4)

    def spline_uniform(self, x0, dx, ydys, xs):
        ydys, xs = _f64(ydys), _f64(xs)
        ys = np.zeros(len(xs))
        self.L.subsample_uniform_spline(float(x0), float(dx), len(ydys), _ptr(ydys), len(xs), _ptr(xs), _ptr(ys))
        return ys

    def spline_nonuniform(self, kx, ydys, xs):
        kx, ydys, xs = _f64(kx), _f64(ydys), _f64(xs)
        ys = np.zeros(len(xs))
        self.L.subsample_nonuniform_spline(len(kx), _ptr(kx), _ptr(ydys), len(xs), _ptr(xs), _ptr(ys))
        return ys

    def test_force(self, typ, n, r0, dr, R, spline=None):
        fs = np.zeros((n, 3))
        self.L.test_force(int(typ), n, _ptr(_f64(r0)), _ptr(_f64(dr)), _ptr(_f64(R)), _ptr(fs), self._spline(spline))
        return fs

    def eigen3x3(self, mat):
        evs = np.zeros((4, 3))
        self.L.test_eigen3x3(_ptr(_f64(mat)), _ptr(evs))
        return evs


def have_reference():
    return REF_PATH.exists()


REF_GU_PATH = HERE / "_ref" / "GridUtils_lib.so"


def have_reference_gridutils():
    return REF_GU_PATH.exists()


def reference_read_numbers(fname, dims, noline):
    """The reference's ReadNumsUpTo_C (GridUtils.cpp:32-71) from oracle/_ref:
    prod(dims) numbers after noline lines.  It prints progress to stdout."""
    L = ctypes.CDLL(str(REF_GU_PATH))
    L.ReadNumsUpTo_C.argtypes = [ctypes.c_char_p, _dp, _ip, ctypes.c_int]
    L.ReadNumsUpTo_C.restype = ctypes.c_int
    dims = np.ascontiguousarray(dims, dtype=np.int32)
    out = np.zeros(int(np.prod(dims.astype(np.int64))))
    L.ReadNumsUpTo_C(str(fname).encode(), _ptr(out), _ptr(dims, _ip), int(noline))
    return out


if __name__ == "__main__":  # pragma: no cover
    print("oracle threads:", num_threads(), "ref:", have_reference(), os.getpid())


# ------------------------------------------------------------- XSF grid text


def xsf_data_text(data_xyz):
    """The data block saveXSFData writes (io.py:575-599), restated with
    Python's own %-formatting: the (nx,ny,nz) grid extended by its first
    plane along each axis, x fastest, one "%10.5e" value per line, and the
    two END lines."""
    a = np.asarray(data_xyz, dtype=np.float64)
    ext = np.concatenate([a, a[:1]], axis=0)
    ext = np.concatenate([ext, ext[:, :1]], axis=1)
    ext = np.concatenate([ext, ext[:, :, :1]], axis=2)
    flat = ext.transpose(2, 1, 0).ravel()
    return "".join("%10.5e\n" % v for v in flat) + "   END_DATAGRID_3D\nEND_BLOCK_DATAGRID_3D\n"


def read_numbers(text, noline, count):
    """ReadNumsUpTo_C (GridUtils.cpp:32-71) on a string: skip noline lines
    (fgets with a 5000-byte buffer: at most 4999 characters each), then the
    next count whitespace-separated fields as floats."""
    pos = 0
    for _ in range(noline):
        nl = text.find("\n", pos, pos + 4999)
        pos = nl + 1 if nl >= 0 else min(len(text), pos + 4999)
    fields = text[pos:].split()
    return np.array([float(v) for v in fields[:count]])
