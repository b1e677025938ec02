"""GPU parity of the further reference-ABI entry points against the compiled
reference's golden fixtures (tests/golden/models.npz) and the FP64 oracle:
the vdW, damped-vdW, Morse, DFT-D3 and density grids, getInPoints_LJ,
evalRadialFF (with the damping constants it leaves behind for later calls),
computeD3Coeffs, the spline subsamplers, test_force, test_eigen3x3, and the
relaxation and stiffness matrix with a radial tip spline.

Grids, points and radial values are FP64 in the reference's operation
order: bar 1e-13 relative to the largest value (exp in Morse, the densities
and D3 may differ by an ulp).  Eigenvalues of test_eigen3x3 come from a
different closed form (bar 1e-9, the reference's 11-digit constants cost it
~1e-11).  The tests in this file run in order: evalRadialFF with a custom
damping constant changes the process-wide value that later getVdWFF(_RE)
calls use, exactly as in the reference."""

import os
import sys

import numpy as np
import pytest

import oracle as O
from ppafm_amd import core

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REL = 1e-13


def _load():
    return np.load(os.path.join(GOLDEN, "models.npz")), np.load(os.path.join(GOLDEN, "grid.npz"))


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _grid(g, with_e=True):
    F = np.zeros(tuple(g["n"]) + (3,))
    E = np.zeros(tuple(g["n"]))
    core.setGridN(g["n"])
    core.setGridCell(np.asarray(g["cell"]))
    core.setFF_Fpointer(F)
    if with_e:
        core.setFF_Epointer(E)
    else:
        core.deleteFF_Epointer()
    return F, E


def _check(tag, F, E, m):
    rf, re = _rel(F, m["F_" + tag]), _rel(E, m["E_" + tag])
    print(f"{tag}: max rel dF {rf:.3g}, dE {re:.3g}")
    assert rf < REL and re < REL, tag


def test_model_grids_match_reference():
    print("case: golden 11x9x13 sheared grid, 5 atoms, vdW / damped vdW / Morse / D3 / densities", flush=True)
    m, g = _load()
    Rs = np.ascontiguousarray(g["Rs"])
    runs = [("vdw", core.getVdWFF, (np.ascontiguousarray(g["cLJs"]),))]
    for k in (1, 2, 3, 4):
        runs.append((f"vdwre{k}", core.getVdWFF_RE, (np.ascontiguousarray(m["REs"]), k, -1.0)))
    runs += [("morse", core.getMorseFF, (np.ascontiguousarray(m["REs"]), -1.6)),
             ("d3", core.getDFTD3FF, (np.ascontiguousarray(m["d3_coeffs"]),)),
             ("gauss", core.getGaussDensity, (np.ascontiguousarray(m["cRAs"]),)),
             ("slater", core.getSlaterDensity, (np.ascontiguousarray(m["cRAs"]),)),
             ("r4", lambda R, c: core.getDensityR4spline(R, c, bNormalize=False), (np.ascontiguousarray(m["cRAs"]),))]
    for tag, fn, args in runs:
        F, E = _grid(g)
        fn(Rs, *args)
        _check(tag, F, E, m)


def test_points_and_radial_then_persistent_damping():
    print("case: getInPoints_LJ at 40 points; evalRadialFF kinds -3..4, default then custom damping", flush=True)
    m, g = _load()
    fe = core.getInPoints_LJ(np.ascontiguousarray(m["points"]), np.ascontiguousarray(g["Rs"]), np.ascontiguousarray(g["cLJs"]))
    print(f"points: max rel {_rel(fe, m['points_FE']):.3g}")
    assert _rel(fe, m["points_FE"]) < REL
    rs = np.ascontiguousarray(m["radial_rs"])
    cf = {-3: (3.4, 0.004), -2: (23.6, 2.4e4), -1: (23.6,), 0: (23.6,), 1: (3.4, 0.004), 2: (3.4, 0.004), 3: (3.4, 0.004), 4: (3.4, 0.004)}
    for kind in range(-3, 5):
        E, F = core.evalRadialFF(rs, cf[kind], kind=kind)
        assert _rel(E, m[f"radial_E{kind}"]) < REL or np.abs(m[f"radial_E{kind}"]).max() == 0
        assert _rel(F, m[f"radial_F{kind}"]) < REL, kind
    for kind, ad in ((0, 90.0), (1, 0.7), (2, 0.9), (3, 0.05), (4, 0.02)):
        E, F = core.evalRadialFF(rs, cf[kind], kind=kind, ADamp=ad)
        assert _rel(F, m[f"radialc_F{kind}"]) < REL, kind
    # the grid builders now use the constants evalRadialFF left behind
    Rs = np.ascontiguousarray(g["Rs"])
    for k in (1, 2, 3, 4):
        F, E = _grid(g)
        core.getVdWFF_RE(Rs, np.ascontiguousarray(m["REs"]), kind=k, ADamp=-1.0)
        _check(f"vdwre{k}_after", F, E, m)
    F, E = _grid(g)
    core.getVdWFF(Rs, np.ascontiguousarray(g["cLJs"]))
    _check("vdw_after", F, E, m)
    # back to the reference's defaults for whatever runs next in this process
    for kind, ad in ((0, 180.0), (1, 0.5), (2, 0.5), (3, 0.03), (4, 0.01)):
        core.evalRadialFF(rs[:1].copy(), cf[kind], kind=kind, ADamp=ad)


def test_many_atoms_span_several_tiles():
    """300 atoms (two LDS tiles of the grid kernel) against the oracle."""
    print("case: 12x10x14 grid, 300 atoms, Morse and damped vdW", flush=True)
    rng = np.random.default_rng(3)
    n = np.array([12, 10, 14], dtype=np.int32)
    cell = np.array([[3.0, 0.0, 0.0], [0.3, 2.6, 0.0], [0.0, 0.2, 3.4]])
    Rs = np.ascontiguousarray(rng.uniform(-1.0, 4.0, (300, 3)))
    REs = np.ascontiguousarray(np.column_stack([rng.uniform(2.5, 3.5, 300), rng.uniform(0.001, 0.01, 300)]))
    g = dict(n=n, cell=cell)
    for name, call, par in (("morse", lambda: core.getMorseFF(Rs, REs, -1.6), -1.6),
                            ("vdw_invr8", lambda: core.getVdWFF_RE(Rs, REs, kind=4, ADamp=0.015), 0.015)):
        F, E = _grid(g)
        call()
        Fo, Eo = np.zeros_like(F), np.zeros_like(E)
        O.model_ff(name, n, cell, Rs, REs, par, F=Fo, E=Eo)
        print(f"{name}: max rel dF {_rel(F, Fo):.3g}, dE {_rel(E, Eo):.3g}")
        assert _rel(F, Fo) < REL and _rel(E, Eo) < REL


def test_d3_coeffs_match_reference():
    sys.path.insert(0, GOLDEN)
    import make_golden as MG

    m, _ = _load()
    p = m["d3_params"]
    c = core.computeD3Coeffs(m["d3_Rs"], m["d3_iZs"], 8, dict(s6=p[0], s8=p[1], a1=p[2], a2=p[3]), MG.d3_tables())
    print(f"d3 coefficients: max rel {_rel(c, m['d3_coeffs']):.3g}")
    assert _rel(c, m["d3_coeffs"]) < REL


def test_splines_test_force_and_eigen3x3():
    m, _ = _load()
    yu = core.subsample_uniform_spline(1.0, 0.2, np.ascontiguousarray(m["spline_ydys_u"]), np.ascontiguousarray(m["spline_xs_u"]))
    yn = core.subsample_nonuniform_spline(np.ascontiguousarray(m["spline_kx"]), np.ascontiguousarray(m["spline_ydys"]),
                                          np.ascontiguousarray(m["spline_xs_n"]))
    print(f"splines: uniform {_rel(yu, m['spline_u']):.3g}, non-uniform {_rel(yn, m['spline_n']):.3g}")
    assert _rel(yu, m["spline_u"]) < REL and _rel(yn, m["spline_n"]) < REL
    core.setTip(lRadial=1.0, kRadial=25 / -16.0217662, rPP0=np.zeros(3), kSpring=np.array([0.3, 0.3, 0.0]) / -16.0217662)
    sp = core.SplineParameters(m["tipspline_xs"], m["tipspline_ydys"])
    for typ, spl in ((1, sp), (2, None)):
        fs = core.test_force(typ, m["tf_r0"], m["tf_dr"], m["tf_R"], np.zeros((25, 3)), spl)
        print(f"test_force type {typ}: max rel {_rel(fs, m[f'tf_type{typ}']):.3g}")
        assert _rel(fs, m[f"tf_type{typ}"]) < REL
    for k in range(3):
        evs = core.test_eigen3x3(m["eig_mats"][k])
        ref = m["eig_out"][k]
        scale = np.abs(ref[0]).max()
        print(f"eigen3x3 {k}: values {np.abs(evs[0] - ref[0]).max() / scale:.3g}, vectors {np.abs(evs[1:] - ref[1:]).max():.3g}")
        assert np.abs(evs[0] - ref[0]).max() < 1e-9 * scale
        assert np.abs(evs[1:] - ref[1:]).max() < 1e-7


def test_spline_tip_relax_and_stiffness():
    print("case: golden 16x16x50 grid, 4x3x6 strokes with a radial tip spline", flush=True)
    m, _ = _load()
    r = np.load(os.path.join(GOLDEN, "relax.npz"))
    F = np.ascontiguousarray(r["FF"])
    core.setGridN(r["n"])
    core.setGridCell(np.asarray(r["cell"]))
    core.setFF_Fpointer(F)
    core.deleteFF_Epointer()
    core.setTip(lRadial=float(r["tip_lRadial"]), kRadial=float(r["tip_kRadial"]), rPP0=np.array(r["tip_rPP0"]),
                kSpring=np.array(r["tip_kSpring"]))
    core.setRelax(1000, 1e-4, 0.1, 0.1)
    sp = core.SplineParameters(m["tipspline_xs"], m["tipspline_ydys"])
    tips = np.ascontiguousarray(r["tips"])
    rs, fs = np.zeros_like(tips), np.zeros_like(tips)
    core.relaxTipStrokes_omp(tips, rs, fs, tip_spline=sp)
    d = np.abs(rs - m["spline_rs"]).max()
    print(f"relax with spline: max |drs| {d:.3g} A, identical {np.array_equal(rs, m['spline_rs'])}")
    assert d < 1e-9
    ev, vecs = core.stiffnessMatrix(np.ascontiguousarray(tips.reshape(-1, 3)), np.ascontiguousarray(rs.reshape(-1, 3)),
                                    which=3, tip_spline=sp)
    evr = m["spline_evals"]
    scale = np.abs(evr).max(axis=1, keepdims=True)
    dev = (np.abs(ev - evr) / scale).max()
    print(f"stiffness with spline: max rel d eigenvalue {dev:.3g}")
    assert dev < 1e-8
