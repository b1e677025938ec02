"""GPU parity of HighLevel.computeLJ (HighLevel.py:260-307) for its three
force-field models -- LJ, Morse (getMorseFF with aMorse) and vdW
(getVdWFF for vdWDampKind 0, getVdWFF_RE otherwise) -- against the FP64
oracle on the same atoms, after the reference's |F| <= Fmax clamp.  The
geometry is a seeded molecule written to an xyz file with a periodic cell
in its comment line (no PBC images)."""

import numpy as np
import pytest

import oracle as O
from ppafm_amd import HighLevel as PPH
from ppafm_amd import common as C
from ppafm_amd import molgen as M

pytestmark = pytest.mark.gpu


def _write_xyz(path, xyz, Zs, lvec):
    sym = {1: "H", 6: "C", 7: "N", 8: "O"}
    L = lvec[1:].ravel()
    with open(path, "w") as f:
        f.write(f"{len(Zs)}\n")
        f.write('Lattice="' + " ".join(f"{v:.6f}" for v in L) + '"\n')
        for z, p in zip(Zs, xyz):
            f.write(f"{sym[int(z)]} {p[0]:.6f} {p[1]:.6f} {p[2]:.6f}\n")


def reset_damping():
    """Put the process-wide vdW damping constants back to the reference's
    defaults (a positive ADamp given to evalRadialFF replaces them)."""
    from ppafm_amd import core

    for kind, name in ((0, "vdw"), (1, "vdw_r2"), (2, "vdw_r4"), (3, "vdw_invr4"), (4, "vdw_invr8")):
        core.evalRadialFF(np.array([3.0]), [3.0, 0.01], kind=kind, ADamp=O.ADAMP_DEFAULT[name])


@pytest.mark.parametrize("model,damp", [("LJ", 0), ("Morse", 0), ("vdW", 0), ("vdW", 2), ("vdW", 3)])
def test_computeLJ_models_match_oracle(tmp_path, model, damp):
    print(f"case: 16-atom molecule, 40x36x50 grid, ffModel {model}, vdWDampKind {damp}", flush=True)
    xyz, Zs, _ = M.molecule(16, 11)
    xyz = xyz + np.array([2.0, 1.8, 2.5])
    lvec = np.array([[0.0, 0.0, 0.0], [4.0, 0.0, 0.0], [0.0, 3.6, 0.0], [0.0, 0.0, 5.0]])
    path = str(tmp_path / "mol.xyz")
    _write_xyz(path, xyz, Zs, lvec)
    reset_damping()
    p = C.PpafmParameters(gridN=[40, 36, 50], PBC=False, probeType="O", vdWDampKind=damp, aMorse=-1.6)
    FF, V, nDim, lv = PPH.computeLJ(path, ffModel=model, computeVpot=True, parameters=p)
    sp = C.loadSpecies()
    n = np.array([40, 36, 50])
    Rs0 = np.ascontiguousarray(np.round(xyz, 6) - lv[0])
    Fo = np.zeros(tuple(n) + (3,))
    Eo = np.zeros(tuple(n))
    if model == "LJ":
        O.lj_ff(n, lv[1:], Rs0, C.getAtomsLJ(8, Zs, sp), F=Fo, E=Eo)
    elif model == "Morse":
        O.model_ff("morse", n, lv[1:], Rs0, C.getAtomsRE(8, Zs, sp), -1.6, F=Fo, E=Eo)
    elif damp == 0:
        O.model_ff("vdw", n, lv[1:], Rs0, C.getAtomsLJ(8, Zs, sp), O.ADAMP_DEFAULT["vdw"], F=Fo, E=Eo)
    else:
        name = {2: "vdw_r4", 3: "vdw_invr4"}[damp]
        O.model_ff(name, n, lv[1:], Rs0, C.getAtomsRE(8, Zs, sp), O.ADAMP_DEFAULT[name], F=Fo, E=Eo)
    O.limit_vec_field(Fo, PPH.Fmax_DEFAULT)
    Eo[Eo > PPH.Vmax_DEFAULT] = PPH.Vmax_DEFAULT
    rf = np.abs(FF - Fo).max() / np.abs(Fo).max()
    re = np.abs(V - Eo).max() / max(np.abs(Eo).max(), 1e-300)
    print(f"{model}/{damp}: max rel dF {rf:.3g}, dE {re:.3g}")
    assert rf < 1e-13
    assert re < 1e-13 or np.abs(Eo).max() == 0


def test_relaxedScan3D_tilted_trajectory_matches_oracle():
    """relaxedScan3D along trjByDir (HighLevel.py:36-41, 63-101): column
    (ix, iy) is the stroke trj + (x, y, 0); vs the oracle relaxing the same
    explicit strokes.  Untilted, it equals relaxedScan3D_omp."""
    from cases import case, tip_params

    from ppafm_amd import core

    print("case: small molecule grid 64x64x80, 9x7 columns of 12 steps tilted by (0.02, -0.01, -0.1)", flush=True)
    k = case(0, small=True)
    F = np.ascontiguousarray(k["FF"])
    core.setGridN(k["cfg"]["n"])
    core.setGridCell(k["lvec"])
    core.setFF_Fpointer(F)
    tp = tip_params()
    core.setTip(lRadial=tp[0], kRadial=tp[1], rPP0=np.zeros(3), kSpring=tp[5:8].copy())
    core.setRelax(1000, 1e-4, 0.1, 0.1)
    core.setFIRE(1.1, 0.5, 0.99)
    o = k["lvec"][0]
    xs = np.linspace(-0.8, 0.8, 9) - o[0]
    ys = np.linspace(-0.6, 0.6, 7) - o[1]
    zs = np.linspace(6.0, 7.1, 12) - o[2]
    trj = PPH.trjByDir(len(zs), d=[0.02, -0.01, -0.1], p0=[0.0, 0.0, zs[0]])
    fzs, pos = PPH.relaxedScan3D(xs, ys, zs, trj=trj, bF3d=True)
    tips = np.zeros((9, 7, 12, 3))
    tips[..., 0] = trj[None, None, :, 0] + xs[:, None, None]
    tips[..., 1] = trj[None, None, :, 1] + ys[None, :, None]
    tips[..., 2] = trj[None, None, :, 2]
    rr, fr, it, tot = O.relax_strokes(k["cfg"]["n"], k["lvec"][1:], F, tp, tips)
    d = np.abs(pos - rr[:, :, ::-1, :]).max()
    df = np.abs(fzs - fr[:, :, ::-1, :]).max()
    print(f"tilted: max |dr| {d:.3g} A, max |dF| {df:.3g} eV/A")
    assert d < 1e-9 and df < 1e-9
    f1, p1 = PPH.relaxedScan3D(xs, ys, zs)
    f2, p2 = PPH.relaxedScan3D_omp(xs, ys, zs)
    assert np.array_equal(p1, p2) and np.array_equal(f1, f2.transpose(2, 1, 0))
