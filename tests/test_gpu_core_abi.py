"""GPU parity of the reference-compatible host ABI (ppafm_amd.core, FP64 on
the device) against the FP64 oracle: the library rounds every operation like
ProbeParticle.cpp built for x86-64, so grids and relaxed strokes agree to a
few ulp (bit-exact where the operation order matches)."""

import numpy as np
import pytest

import oracle as O
from cases import case, tip_params
from ppafm_amd import common as C
from ppafm_amd import core
from ppafm_amd import molgen as M

pytestmark = pytest.mark.gpu


def _setup(k, F, E=None):
    core.setGridN(k["cfg"]["n"])
    core.setGridCell(k["lvec"])
    core.setFF_Fpointer(F)
    if E is not None:
        core.setFF_Epointer(E)
    else:
        core.deleteFF_Epointer()


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("model", ["lj", "coul0", "coul1", "coul2"])
def test_grid_fp64(model):
    print("case: small molecule, 64x64x80 grid, model", model, flush=True)
    k = case(0, small=True)
    n = k["cfg"]["n"]
    F = np.zeros(tuple(n) + (3,))
    E = np.zeros(tuple(n))
    _setup(k, F, E)
    Fr = np.zeros_like(F)
    Er = np.zeros_like(E)
    Rs0 = np.ascontiguousarray(k["Rs0"])
    if model == "lj":
        core.getLennardJonesFF(Rs0, np.ascontiguousarray(k["cLJs"]))
        O.lj_ff(n, k["lvec"][1:], Rs0, k["cLJs"], F=Fr, E=Er)
    else:
        kind = int(model[-1])
        kq = np.ascontiguousarray(k["qs"] * C.CoulombConst)
        core.getCoulombFF(Rs0, kq, kind=kind)
        O.coulomb_ff(n, k["lvec"][1:], Rs0, kq, kind=kind, F=Fr, E=Er)
    rf, re = rel(F, Fr), rel(E, Er)
    print(f"{model}: max rel dF {rf:.3g}, dE {re:.3g}; bit-identical F: {np.array_equal(F, Fr)}")
    assert rf < 1e-13 and re < 1e-13


def test_grid_accumulates():
    k = case(0, small=True)
    n = k["cfg"]["n"]
    F = np.ones(tuple(n) + (3,))
    _setup(k, F)
    core.getLennardJonesFF(np.ascontiguousarray(k["Rs0"]), np.ascontiguousarray(k["cLJs"]))
    Fr = O.lj_ff(n, k["lvec"][1:], k["Rs0"], k["cLJs"], F=np.ones_like(F))
    assert rel(F, Fr) < 1e-13


def test_relax_strokes_fp64():
    print("case: small molecule, 24x24x20 scan through the core ABI", flush=True)
    k = case(0, small=True)
    cfg = k["cfg"]
    F = np.ascontiguousarray(k["FF"])
    _setup(k, F)
    tp = tip_params()
    core.setTip(lRadial=tp[0], kRadial=tp[1], rPP0=np.array([0.0, 0.0, 0.0]), kSpring=tp[5:8].copy())
    core.setRelax(1000, 1e-4, 0.1, 0.1)
    core.setFIRE(1.1, 0.5, 0.99)
    tips = np.ascontiguousarray(M.scan_tips(cfg) - k["lvec"][0])
    rs = np.zeros_like(tips)
    fs = np.zeros_like(tips)
    core.relaxTipStrokes_omp(tips, rs, fs)
    rr, fr, it, tot = O.relax_strokes(cfg["n"], k["lvec"][1:], F, tp, tips)
    d = np.abs(rs - rr).max()
    print(f"max |drs| {d:.3g}, max |dfs| {np.abs(fs - fr).max():.3g}, itrsum gpu {core.last_itrsum()} oracle {tot}; identical: {np.array_equal(rs, rr)}")
    assert d < 1e-9
    assert core.last_itrsum() == tot


def test_relax_single_stroke_returns_itrsum():
    k = case(0, small=True)
    F = np.ascontiguousarray(k["FF"])
    _setup(k, F)
    tp = tip_params()
    core.setTip(lRadial=tp[0], kRadial=tp[1], rPP0=np.array([0.0, 0.0, 0.0]), kSpring=tp[5:8].copy())
    core.setRelax(1000, 1e-4, 0.1, 0.1)
    tips = np.ascontiguousarray((M.scan_tips(k["cfg"]) - k["lvec"][0])[3, 5])
    rs = np.zeros_like(tips)
    fs = np.zeros_like(tips)
    itr = core.relaxTipStroke(tips, rs, fs)
    n = np.ascontiguousarray(k["cfg"]["n"], dtype=np.int32)
    rr = np.zeros_like(tips)
    fr = np.zeros_like(tips)
    it = np.zeros(len(tips), dtype=np.int32)
    tot = O.lib().orc_relax_stroke(n.ctypes.data_as(O._ip), O._ptr(np.ascontiguousarray(k["lvec"][1:])), O._ptr(F), O._ptr(tp), O._ptr(np.array(O.DEFAULT_RELAX, dtype=np.float64)), 1, 1, len(tips), O._ptr(tips), O._ptr(rr), O._ptr(fr), O._ptr(it, O._ip), None)
    assert itr == tot
    assert np.abs(rs - rr).max() < 1e-9


def test_malformed_tip_spline_rejected():
    """A spline with fewer than 2 knots (here an empty SplineParams) is an error."""
    import ctypes

    from ppafm_amd import _lib

    k = case(0, small=True)
    F = np.ascontiguousarray(k["FF"])
    _setup(k, F)
    tips = np.zeros((1, 1, 2, 3))
    sp = _lib.SplineParams()
    with pytest.raises(_lib.HipError):
        core.relaxTipStrokes_omp(tips, np.zeros_like(tips), np.zeros_like(tips), tip_spline=ctypes.pointer(sp))
