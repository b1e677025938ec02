"""GPU parity of stiffnessMatrix (ProbeParticle.cpp:703-750): the probe's
dynamical matrix at relaxed positions and its eigen-decomposition.

The six force evaluations per point round like the reference (FP64, the
reference's operation order), so the dynamical matrices D are the oracle's
(which equal the compiled reference's bit for bit).  The eigenvalues come
from a different closed form of the same cubic, so they are pinned by
numpy's symmetric eigensolver on the oracle's D: within 1e-12 of the largest
|eigenvalue|, or 1e-7 where two eigenvalues lie closer than 1e-4 of it
(trigonometric root formulas lose half the digits at a double root).  The
reference writes 1/3 and sqrt(3) as 11-digit decimals, which costs it up to
~2e-9 near close pairs; its eigenvalues must differ from ours by no more
than its own distance from eigh.  Eigenvectors must satisfy
|D v - lambda v| < 1e-9 max|lambda| and |v| = 1 everywhere, and equal the
oracle's (same selection and sign rule) wherever their eigenvalue is
separated from the others by more than 1e-3 max|lambda|, to 1e-7 plus
4 |d lambda| / gap (an eigenvalue error turns the vector by about that).

  * host ABI (core.stiffnessMatrix) against the golden fixture written by the
    compiled reference (tests/golden/stiff.npz), which = 3, two ddisp;
  * `which` controls which eigenvector arrays are written;
  * device path (PlainScan.stiffness on its double4 grid and float32 relaxed
    positions) against the oracle on the same grid and positions."""

import os

import numpy as np
import pytest

import oracle as O
from cases import small_config
from ppafm_amd import core
from ppafm_amd import molgen as M
from ppafm_amd.scan import PlainScan

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EV_TOL = 1e-9  # eigen-residual, of max |eigenvalue|
VEC_TOL = 1e-7
GAP = 1e-3  # of max |eigenvalue|


def check_eigenvalues(ev, evr, D):
    """GPU eigenvalues ev against eigh of the oracle's matrices D and against
    the oracle's/reference's evr; returns (worst vs eigh, worst vs evr)."""
    ev = ev.reshape(-1, 3)
    w = np.linalg.eigvalsh(D)
    scale = np.abs(w).max(axis=1)
    gaps = np.minimum(w[:, 1] - w[:, 0], w[:, 2] - w[:, 1])
    tol = np.where(gaps > 1e-4 * scale, 1e-12, 1e-7) * scale
    d_exact = np.abs(ev - w).max(axis=1)
    assert (d_exact < tol).all(), (d_exact / scale).max()
    d_ref = np.abs(ev - evr).max(axis=1)
    ref_err = np.abs(evr - w).max(axis=1)
    assert (d_ref <= ref_err + tol).all(), (d_ref / scale).max()
    return (d_exact / scale).max(), (d_ref / scale).max()


def check_vectors(ev, vecs, evr, vr, D):
    """Eigen-residual against the oracle's matrices D (m,3,3), unit norm, and
    equality with the oracle's vectors where the eigenvalue is separated."""
    scale = np.abs(evr).max(axis=1)
    worst_res = worst_dv = 0.0
    ncmp = 0
    for k in range(3):
        v = vecs[k].reshape(-1, 3)
        lam = ev.reshape(-1, 3)[:, k]
        res = np.linalg.norm(np.einsum("mij,mj->mi", D, v) - lam[:, None] * v, axis=1) / scale
        assert res.max() < EV_TOL, (k, res.max())
        assert np.abs(np.linalg.norm(v, axis=1) - 1).max() < 1e-12
        others = np.delete(evr, k, axis=1)
        gap = np.abs(others - evr[:, k:k + 1]).min(axis=1)
        sep = gap > GAP * scale
        bound = VEC_TOL + 4 * np.abs(lam - evr[:, k]) / np.maximum(gap, 1e-300)
        dv = np.abs(v - vr[k]).max(axis=1)
        assert (dv[sep] < bound[sep]).all(), (k, dv[sep].max())
        dv = dv[sep]
        worst_res = max(worst_res, res.max())
        worst_dv = max(worst_dv, dv.max() if dv.size else 0.0)
        ncmp += int(sep.sum())
    return worst_res, worst_dv, ncmp


def _load_golden():
    r = np.load(os.path.join(GOLDEN, "relax.npz"))
    s = np.load(os.path.join(GOLDEN, "stiff.npz"))
    F = np.ascontiguousarray(r["FF"])
    core.setGridN(r["n"])
    core.setGridCell(np.asarray(r["cell"]))
    core.setFF_Fpointer(F)
    core.deleteFF_Epointer()
    core.setTip(lRadial=float(r["tip_lRadial"]), kRadial=float(r["tip_kRadial"]), rPP0=np.array(r["tip_rPP0"]),
                kSpring=np.array(r["tip_kSpring"]))
    return r, s, F


@pytest.mark.parametrize("tag", ["05", "02"])
def test_host_abi_matches_reference_golden(tag):
    print(f"case: golden 16x16x50 grid, 144 points, ddisp 0.{tag}", flush=True)
    r, s, F = _load_golden()
    tips = np.ascontiguousarray(s["tips"])
    rpps = np.ascontiguousarray(s["rpps"])
    dd = int(tag) / 100
    ev, vecs = core.stiffnessMatrix(tips, rpps, which=3, ddisp=dd)
    evr = s[f"evals_{tag}"]
    vr = [s[f"evec{k + 1}_{tag}"] for k in range(3)]
    tip = O.tip_params(float(r["tip_lRadial"]), float(r["tip_kRadial"]), r["tip_rPP0"], r["tip_kSpring"])
    _, _, _, D = O.stiffness(r["n"], r["cell"], F, tip, tips, rpps, which=0, ddisp=dd, want_dynmat=True)
    d_exact, d_ref = check_eigenvalues(ev, evr, D)
    res, dv, ncmp = check_vectors(ev, vecs, evr, vr, D)
    print(f"eigenvalues: max rel diff vs eigh {d_exact:.3g}, vs reference {d_ref:.3g}; eigen-residual {res:.3g}, "
          f"max |d eigenvector| {dv:.3g} over {ncmp} separated of {3 * len(ev)}")


def test_which_selects_eigenvectors():
    r, s, F = _load_golden()
    tips = np.ascontiguousarray(s["tips"][:20])
    rpps = np.ascontiguousarray(s["rpps"][:20])
    ev0, v0 = core.stiffnessMatrix(tips, rpps, which=0)
    assert v0 == [None, None, None]
    ev1, v1 = core.stiffnessMatrix(tips, rpps, which=1)
    assert v1[1] is None and v1[2] is None
    assert np.array_equal(ev0, ev1)
    ev3, v3 = core.stiffnessMatrix(tips, rpps, which=3)
    assert np.array_equal(ev0, ev3) and np.array_equal(v1[0], v3[0])


def test_device_path_matches_oracle():
    cfg, seed = small_config(scan_n=(16, 16), nz=12)
    print("case: small molecule (seed 7), 64x64x80 grid, 16x16x12 scan, which 3", flush=True)
    xyz, Zs, qs = M.molecule(cfg["natoms"], seed)
    lvec = M.config_lvec(cfg)
    sc = PlainScan(lvec, cfg["n"], cfg, tipR0=M.TIP_R0, k_lat=M.K_LAT, k_rad=M.K_RAD, probe_q=M.PROBE_Q)
    sc.set_molecule(xyz, Zs, qs)
    sc.run()
    ev, vecs = sc.stiffness(which=3, ddisp=0.05)
    G = sc.grid_host()
    F = np.ascontiguousarray(G[..., :3])
    pos = sc.pos.download().astype(np.float64).reshape(-1, 3)
    s = sc.scan_d
    i, j, k = np.meshgrid(np.arange(s.nx), np.arange(s.ny), np.arange(s.nz), indexing="ij")
    tips = np.stack([s.p0[c] + i * s.da[c] + j * s.db[c] + k * s.dstep[c] for c in range(3)], axis=-1).reshape(-1, 3)
    tip = O.tip_params(sc.tip.lrad, sc.tip.krad, tuple(sc.tip.d0), tuple(sc.tip.kspring))
    evr, vr, nnan, D = O.stiffness(cfg["n"], lvec[1:], F, tip, tips, pos, which=3, ddisp=0.05, want_dynmat=True)
    assert nnan == 0
    ev = ev.reshape(-1, 3)
    d_exact, d_ref = check_eigenvalues(ev, evr, D)
    res, dv, ncmp = check_vectors(ev, vecs, evr, vr, D)
    print(f"{len(ev)} points: eigenvalues max rel diff vs eigh {d_exact:.3g}, vs oracle {d_ref:.3g}; eigen-residual {res:.3g}, "
          f"max |d eigenvector| {dv:.3g} over {ncmp} separated of {3 * len(ev)}; "
          f"smallest eigenvalue range {evr[:, 0].min():.3g} .. {evr[:, 0].max():.3g} eV/A^2")
