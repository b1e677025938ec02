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


def relaxTipStroke(rTips, rs, fs, probeStart=1, relaxAlg=1, tip_spline=None):
    _c64(rTips, 2, "rTips"), _c64(rs, 2, "rs"), _c64(fs, 2, "fs")
    n = len(rTips)
    r = lib().relaxTipStroke(int(probeStart), int(relaxAlg), n, _lib.dptr(rTips), _lib.dptr(rs), _lib.dptr(fs), tip_spline)
    if r < 0:
        raise _lib.HipError("relaxTipStroke failed: " + lib().ppafm_last_error().decode())
    return r


def relaxTipStrokes_omp(rTips, rs, fs, probeStart=1, relaxAlg=1, tip_spline=None):
    _c64(rTips, 4, "rTips"), _c64(rs, 4, "rs"), _c64(fs, 4, "fs")
    nx, ny, nz, _ = rTips.shape
    r = lib().relaxTipStrokes_omp(nx, ny, int(probeStart), int(relaxAlg), nz, _lib.dptr(rTips), _lib.dptr(rs), _lib.dptr(fs), tip_spline)
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
                          _lib.dptr(evecs[0]), _lib.dptr(evecs[1]), _lib.dptr(evecs[2]), tip_spline)
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
