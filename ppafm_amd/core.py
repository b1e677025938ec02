"""Drop-in for ``ppafm.core``: the same functions, arguments and array
layouts (ppafm/core.py:33-324), backed by libppafm_hip.so on the GPU.

Arrays are C-contiguous float64 numpy arrays exactly as the reference
passes them to ProbeParticle.cpp; the library copies them to the device,
runs the HIP kernels in FP64 and writes the results back in place.  Like
the reference the grid, tip and relaxation settings are process-wide state.
"""

import weakref

import numpy as np

from . import _lib
from . import common as PPU
from ._lib import lib

_keep = {}


def _c64(a, ndim=None, name="array"):
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous):
        raise TypeError(f"{name} must be a C-contiguous float64 numpy array")
    if ndim is not None and a.ndim != ndim:
        raise TypeError(f"{name} must have {ndim} dimensions, got {a.ndim}")
    return a


def setGridN(n):
    n = np.array(n, dtype=np.int32).copy()
    lib().setGridN(_lib.iptr(n))


def setGridCell(cell):
    cell = np.array(cell[-3:, 0:3], copy=True, dtype=np.float64)
    lib().setGridCell(_lib.dptr(np.ascontiguousarray(cell)))


def _release(key, token):
    # Clear the library pointer only if the collected array is still the one
    # registered; the reference's finalizer (core.py:61,71) clears it even
    # when a newer array has been set since, which leaves the grid unset.
    cur = _keep.get(key)
    if cur is not None and cur[1] == token:
        (lib().deleteFF_Fpointer if key == "F" else lib().deleteFF_Epointer)()
        _keep.pop(key, None)


_token = [0]


def _register(key, arr):
    _token[0] += 1
    tok = _token[0]
    _keep[key] = (weakref.ref(arr), tok)
    weakref.finalize(arr, _release, key, tok)


def setFF_Fpointer(gridF):
    _c64(gridF, 4, "gridF")
    lib().setFF_Fpointer(_lib.dptr(gridF))
    _register("F", gridF)


def setFF_Epointer(gridE):
    _c64(gridE, 3, "gridE")
    lib().setFF_Epointer(_lib.dptr(gridE))
    _register("E", gridE)


def deleteFF_Fpointer():
    lib().deleteFF_Fpointer()
    _keep.pop("F", None)


def deleteFF_Epointer():
    lib().deleteFF_Epointer()
    _keep.pop("E", None)


def setRelax(maxIters=1000, convF2=1.0e-4, dt=0.1, damping=0.1):
    """Relaxation settings.  The convergence argument is the force threshold
    and is squared before it reaches the library, which is what the
    reference intends (its body, ppafm/core.py:97-98, squares ``convF``)."""
    lib().setRelax(int(maxIters), float(convF2) * float(convF2), float(dt), float(damping))


def setFIRE(finc=1.1, fdec=0.5, falpha=0.99):
    lib().setFIRE(float(finc), float(fdec), float(falpha))


def setTip(lRadial=None, kRadial=None, rPP0=None, kSpring=None, parameters=None):
    if lRadial is None:
        lRadial = parameters.r0Probe[2]
    if kRadial is None:
        kRadial = parameters.krad / -PPU.eVA_Nm
    if rPP0 is None:
        rPP0 = np.array((parameters.r0Probe[0], parameters.r0Probe[1], 0.0))
    if kSpring is None:
        kSpring = np.array((parameters.klat, parameters.klat, 0.0)) / -PPU.eVA_Nm
    rPP0 = np.ascontiguousarray(rPP0, dtype=np.float64)
    kSpring = np.ascontiguousarray(kSpring, dtype=np.float64)
    lib().setTip(float(lRadial), float(kRadial), _lib.dptr(rPP0), _lib.dptr(kSpring))


def getLennardJonesFF(Rs, cLJs):
    Rs = _c64(Rs, 2, "Rs")
    cLJs = _c64(cLJs, 2, "cLJs")
    lib().ppafm_clear_error()
    lib().getLennardJonesFF(len(Rs), _lib.dptr(Rs), _lib.dptr(cLJs))
    _check_last("getLennardJonesFF")


def getCoulombFF(Rs, kQQs, kind=0):
    Rs = _c64(Rs, 2, "Rs")
    kQQs = _c64(kQQs, 1, "kQQs")
    lib().ppafm_clear_error()
    lib().getCoulombFF(len(Rs), _lib.dptr(Rs), _lib.dptr(kQQs), int(kind))
    _check_last("getCoulombFF")


def getVdWFF(Rs, cLJs):
    """vdW with constant damping (ppafm/core.py:203-205, ProbeParticle.cpp:515)."""
    _grid_call("getVdWFF", Rs, cLJs, 2)


def getVdWFF_RE(Rs, REs, kind=0, ADamp=-1.0):
    """vdW with damping function ``kind`` 1..4 (ppafm/core.py:213-215, ProbeParticle.cpp:577)."""
    _grid_call("getVdWFF_RE", Rs, REs, 2, int(kind), float(ADamp))


def getDFTD3FF(Rs, d3_coeffs):
    """DFT-D3 grid from computeD3Coeffs' coefficients (ppafm/core.py:223-225)."""
    _grid_call("getDFTD3FF", Rs, d3_coeffs, 4)


def computeD3Coeffs(Rs, iZs, iZPP, df_params, tables):
    """DFT-D3 coefficients (natoms,4) (ppafm/core.py:236-249).  ppafm loads the
    reference tables from its own ppafm.defaults.d3 module; here they are
    passed in ``tables``: dict with R_COV (94), R0 (ignored), REF_CN (94,5),
    REF_C6 (94,94,5,5), R4R2 (94), K (K1, K2, K3).  df_params: s6, s8, a1 and
    a2 (a2 in Angstrom)."""
    Rs = np.ascontiguousarray(Rs, dtype=np.float64).reshape(-1, 3)
    iZs = np.ascontiguousarray(iZs, dtype=np.int32)
    if len(iZs) != len(Rs):
        raise ValueError("iZs and Rs differ in length")
    t = {k: np.ascontiguousarray(v, dtype=np.float64).ravel() for k, v in tables.items() if k != "K"}
    k = np.ascontiguousarray(tables["K"], dtype=np.float64)
    p = np.array([df_params["s6"], df_params["s8"], df_params["a1"], df_params["a2"]], dtype=np.float64)
    r_cut = t.get("R0", np.zeros(1))
    out = np.empty(4 * len(Rs))
    lib().ppafm_clear_error()
    lib().computeD3Coeffs(len(Rs), _lib.dptr(Rs), _lib.iptr(iZs), _lib.dptr(t["R_COV"]), _lib.dptr(r_cut), _lib.dptr(t["REF_CN"]),
                          _lib.dptr(t["REF_C6"]), _lib.dptr(t["R4R2"]), _lib.dptr(k), _lib.dptr(p), int(iZPP), _lib.dptr(out))
    _check_last("computeD3Coeffs")
    return out.reshape(-1, 4)


def getMorseFF(Rs, REs, alpha=None, parameters=None):
    """Morse grid (ppafm/core.py:257-262); alpha defaults to parameters.aMorse."""
    if alpha is None:
        alpha = parameters.aMorse
    _grid_call("getMorseFF", Rs, REs, 2, float(alpha))


def getGaussDensity(Rs, cRAs):
    _grid_call("getGaussDensity", Rs, cRAs, 2)


def getSlaterDensity(Rs, cRAs):
    _grid_call("getSlaterDensity", Rs, cRAs, 2)


def getDensityR4spline(Rs, cRAs, bNormalize=True):
    """As ppafm/core.py:300-304: with bNormalize the amplitudes in cRAs are
    divided in place by the kernel's integral (32 pi / 105) R^3."""
    if bNormalize:
        cRAs[:, 0] /= ((np.pi * 32) / 105) * cRAs[:, 1] ** 3
    _grid_call("getDensityR4spline", Rs, cRAs, 2)


def getInPoints_LJ(ps, Rs, cLJs, FEs=None):
    """LJ (F, E) at points ps (n,3) (ppafm/core.py:163-169) -> FEs (n,4)."""
    _c64(ps, 2, "ps"), _c64(Rs, 2, "Rs"), _c64(cLJs, 2, "cLJs")
    if FEs is None:
        FEs = np.zeros((len(ps), 4))
    _c64(FEs, 2, "FEs")
    lib().ppafm_clear_error()
    lib().getInPoints_LJ(len(ps), _lib.dptr(ps), _lib.dptr(FEs), len(Rs), _lib.dptr(Rs), _lib.dptr(cLJs))
    _check_last("getInPoints_LJ")
    return FEs


def evalRadialFF(rs, coefs, Es=None, Fs=None, kind=1, ADamp=-1):
    """One pair term along x at distances rs (ppafm/core.py:177-185) -> Es, Fs."""
    _c64(rs, 1, "rs")
    n = len(rs)
    Es = np.zeros(n) if Es is None else _c64(Es, 1, "Es")
    Fs = np.zeros(n) if Fs is None else _c64(Fs, 1, "Fs")
    coefs = np.array(coefs, dtype=np.float64)
    lib().ppafm_clear_error()
    lib().evalRadialFF(n, _lib.dptr(rs), _lib.dptr(coefs), _lib.dptr(Es), _lib.dptr(Fs), int(kind), float(ADamp))
    _check_last("evalRadialFF")
    return Es, Fs


def _grid_call(name, Rs, coefs, ncoef, *extra):
    _c64(Rs, 2, "Rs")
    _c64(coefs, 2, "coefficients")
    if coefs.shape != (Rs.shape[0], ncoef):
        raise ValueError(f"{name}: need ({len(Rs)}, {ncoef}) coefficients, got {coefs.shape}")
    lib().ppafm_clear_error()
    getattr(lib(), name)(len(Rs), _lib.dptr(Rs), _lib.dptr(coefs), *extra)
    _check_last(name)


class SplineParameters(_lib.SplineParams):
    """Radial tip spline (ppafm/core.py:137-152): knots xs and (n,2) values
    and slopes ydys; pass as ``tip_spline`` to the relaxation and stiffness
    calls."""

    def __init__(self, xs, ydys):
        super().__init__()
        self.np_rff_xs = np.ascontiguousarray(xs, dtype=np.float64)
        self.np_rff_ydys = np.ascontiguousarray(ydys, dtype=np.float64)
        self.rff_n = len(self.np_rff_xs)
        self.rff_xs = _lib.dptr(self.np_rff_xs)
        self.rff_ydys = _lib.dptr(self.np_rff_ydys)


def _spline_arg(tip_spline):
    import ctypes

    if tip_spline is None or isinstance(tip_spline, ctypes._Pointer):
        return tip_spline
    return ctypes.pointer(tip_spline)


def subsample_uniform_spline(x0, dx, ydys, xs_, ys_=None):
    """Hermite spline on uniform knots sampled at xs_ (ppafm/core.py:353-359)."""
    _c64(ydys, 2, "ydys"), _c64(xs_, 1, "xs_")
    ys_ = np.zeros(len(xs_)) if ys_ is None else _c64(ys_, 1, "ys_")
    lib().ppafm_clear_error()
    lib().subsample_uniform_spline(float(x0), float(dx), len(ydys), _lib.dptr(ydys), len(xs_), _lib.dptr(xs_), _lib.dptr(ys_))
    _check_last("subsample_uniform_spline")
    return ys_


def subsample_nonuniform_spline(xs, ydys, xs_, ys_=None):
    """Hermite spline on knots xs sampled at xs_ (ppafm/core.py:367-373)."""
    _c64(xs, 1, "xs"), _c64(ydys, 2, "ydys"), _c64(xs_, 1, "xs_")
    ys_ = np.zeros(len(xs_)) if ys_ is None else _c64(ys_, 1, "ys_")
    lib().ppafm_clear_error()
    lib().subsample_nonuniform_spline(len(xs), _lib.dptr(xs), _lib.dptr(ydys), len(xs_), _lib.dptr(xs_), _lib.dptr(ys_))
    _check_last("subsample_nonuniform_spline")
    return ys_


def test_force(typ, r0, dr, R, fs, spline_parameters=None):
    """Radial tip force along r0 + k dr (ppafm/core.py:381-384)."""
    r0, dr, R = (np.ascontiguousarray(v, dtype=np.float64) for v in (r0, dr, R))
    _c64(fs, 2, "fs")
    lib().ppafm_clear_error()
    lib().test_force(int(typ), len(fs), _lib.dptr(r0), _lib.dptr(dr), _lib.dptr(R), _lib.dptr(fs), _spline_arg(spline_parameters))
    _check_last("test_force")
    return fs


def test_eigen3x3(mat):
    """Eigenvalues and eigenvectors of a 3x3 matrix (ppafm/core.py:392-395): (4,3)."""
    mat = np.ascontiguousarray(mat, dtype=np.float64)
    evs = np.zeros((4, 3))
    lib().ppafm_clear_error()
    lib().test_eigen3x3(_lib.dptr(mat), _lib.dptr(evs))
    _check_last("test_eigen3x3")
    return evs


def relaxTipStroke(rTips, rs, fs, probeStart=1, relaxAlg=1, tip_spline=None):
    _c64(rTips, 2, "rTips"), _c64(rs, 2, "rs"), _c64(fs, 2, "fs")
    n = len(rTips)
    r = lib().relaxTipStroke(int(probeStart), int(relaxAlg), n, _lib.dptr(rTips), _lib.dptr(rs), _lib.dptr(fs), _spline_arg(tip_spline))
    if r < 0:
        raise _lib.HipError("relaxTipStroke failed: " + lib().ppafm_last_error().decode())
    return r


def relaxTipStrokes_omp(rTips, rs, fs, probeStart=1, relaxAlg=1, tip_spline=None):
    _c64(rTips, 4, "rTips"), _c64(rs, 4, "rs"), _c64(fs, 4, "fs")
    nx, ny, nz, _ = rTips.shape
    r = lib().relaxTipStrokes_omp(nx, ny, int(probeStart), int(relaxAlg), nz, _lib.dptr(rTips), _lib.dptr(rs), _lib.dptr(fs), _spline_arg(tip_spline))
    if r < 0:
        raise _lib.HipError("relaxTipStrokes_omp failed: " + lib().ppafm_last_error().decode())
    return r


def stiffnessMatrix(rTips, rPPs, which=0, ddisp=0.05, tip_spline=None):
    """Eigenvalues (n,3), ascending, of the probe's dynamical matrix at the
    relaxed positions rPPs (n,3) for tips rTips (n,3), and a list of three
    (n,3) eigenvector arrays of which the first ``which`` are filled (the
    rest None), as ppafm/core.py:332-345 (ProbeParticle.cpp:703-750)."""
    _c64(rTips, 2, "rTips"), _c64(rPPs, 2, "rPPs")
    n = len(rTips)
    if rPPs.shape != rTips.shape or rTips.shape[1:] != (3,):
        raise ValueError(f"rTips and rPPs must both be (n,3), got {rTips.shape} and {rPPs.shape}")
    eigenvals = np.zeros((n, 3))
    evecs = [None, None, None]
    for i in range(which):
        evecs[i] = np.zeros((n, 3))
    lib().ppafm_clear_error()
    lib().stiffnessMatrix(float(ddisp), int(which), n, _lib.dptr(rTips), _lib.dptr(rPPs), _lib.dptr(eigenvals),
                          _lib.dptr(evecs[0]), _lib.dptr(evecs[1]), _lib.dptr(evecs[2]), _spline_arg(tip_spline))
    _check_last("stiffnessMatrix")
    return eigenvals, evecs


def last_itrsum():
    """Sum of FIRE iteration indices of the last relaxTipStrokes_omp call."""
    return int(lib().ppafm_last_itrsum())


def _check_last(what):
    # the void entry points report failures through ppafm_last_error (and stderr)
    msg = lib().ppafm_last_error()
    if msg:
        raise _lib.HipError(f"{what} failed: {msg.decode(errors='replace')}")
