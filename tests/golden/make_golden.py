"""Write the golden fixtures of tests/golden/ from the reference C++.

Runs the reference ProbeParticle.cpp compiled out of tree by
``make -C oracle ref`` (oracle/_ref/ProbeParticle_lib.so) through its own C
ABI on small seeded inputs and stores inputs and outputs as .npz data:

  grid.npz   LJ (F and E) and Coulomb kinds 0/1/2 on an 11x9x13 grid of a
             non-orthogonal cell, 5 atoms (getLennardJonesFF / getCoulombFF,
             ProbeParticle.cpp:508-554)
  relax.npz  relaxTipStrokes_omp on a 16x16x50 LJ+Coulomb grid (clamped to
             10 eV/A like io.limit_vec_field), 4x3 columns
             of 6 tip positions, for probeStart 1/0/-1 (FIRE) and the damped
             leap-frog (relaxAlg 0), plus relaxTipStroke's iteration sum
             (ProbeParticle.cpp:633-701)
  stiff.npz  stiffnessMatrix on that grid at the relaxed positions of the
             probeStart-1 strokes (and at positions displaced from them by
             up to 0.3 A), which = 3, ddisp 0.05 and 0.02
             (ProbeParticle.cpp:703-750)

Usage: python tests/golden/make_golden.py   (needs oracle/_ref built)
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402


def grid_case(ref):
    rng = np.random.default_rng(20240601)
    n = np.array([11, 9, 13], dtype=np.int32)
    cell = np.array([[1.1, 0.0, 0.0], [0.15, 0.9, 0.0], [0.05, 0.02, 1.3]])
    Rs = rng.uniform(-0.4, 1.6, (5, 3))
    cLJs = np.column_stack([rng.uniform(5, 30, 5), rng.uniform(1e3, 3e4, 5)])
    kQQs = rng.uniform(-0.3, 0.3, 5) * 14.399645
    out = dict(n=n, cell=cell, Rs=Rs, cLJs=cLJs, kQQs=kQQs)
    F = np.zeros((11, 9, 13, 3))
    E = np.zeros((11, 9, 13))
    ref.set_grid(n, cell, F, E)
    ref.lj(Rs, cLJs)
    out["F_lj"], out["E_lj"] = F.copy(), E.copy()
    for kind in (0, 1, 2):
        F[:] = 0
        E[:] = 0
        ref.coulomb(Rs, kQQs, kind)
        out[f"F_coul{kind}"], out[f"E_coul{kind}"] = F.copy(), E.copy()
    return out


def relax_case(ref):
    rng = np.random.default_rng(20240602)
    n = np.array([16, 16, 50], dtype=np.int32)
    cell = np.diag([1.6, 1.6, 5.0])
    Rs = np.column_stack([rng.uniform(0.3, 1.3, 3), rng.uniform(0.3, 1.3, 3), np.full(3, 0.5)])
    cLJs = np.tile([23.6, 2.4e4], (3, 1)) * rng.uniform(0.8, 1.2, (3, 1))
    kQQs = np.array([0.2, -0.1, -0.1]) * 14.399645
    F = np.zeros(tuple(n) + (3,))
    ref.set_grid(n, cell, F)
    ref.lj(Rs, cLJs)
    Fel = np.zeros_like(F)
    ref.set_grid(n, cell, Fel)
    ref.coulomb(Rs, kQQs, 0)
    O.limit_vec_field(F, 10.0)
    O.limit_vec_field(Fel, 10.0)
    FF = np.ascontiguousarray(F - 0.1 * Fel)
    tip = dict(lRadial=1.0, kRadial=30 / -16.0217662, rPP0=np.array([0.05, -0.03, 0.0]), kSpring=np.array([0.25, 0.25, 0.0]) / -16.0217662)
    tips = np.zeros((4, 3, 6, 3))
    tips[..., 0] = np.linspace(0.4, 1.2, 4)[:, None, None]
    tips[..., 1] = np.linspace(0.5, 1.1, 3)[None, :, None]
    tips[..., 2] = (4.0 - 0.1 * np.arange(6))[None, None, :]
    out = dict(n=n, cell=cell, FF=FF, tips=tips, tip_lRadial=tip["lRadial"], tip_kRadial=tip["kRadial"], tip_rPP0=tip["rPP0"], tip_kSpring=tip["kSpring"])
    ref.set_grid(n, cell, FF)
    ref.set_tip(tip["lRadial"], tip["kRadial"], tip["rPP0"], tip["kSpring"])
    ref.set_relax(1000, 1e-8, 0.1, 0.1)
    for ps in (1, 0, -1):
        rs, fs = ref.relax_strokes(tips, probeStart=ps, relaxAlg=1)
        out[f"rs_ps{ps}"], out[f"fs_ps{ps}"] = rs, fs
    ref.set_relax(300, 1e-8, 0.1, 0.1)
    rs, fs = ref.relax_strokes(tips, probeStart=1, relaxAlg=0)
    out["rs_alg0"], out["fs_alg0"] = rs, fs
    ref.set_relax(1000, 1e-8, 0.1, 0.1)
    rs1, fs1, itr = ref.relax_stroke(tips[2, 1], probeStart=1, relaxAlg=1)
    out["stroke_rs"], out["stroke_fs"], out["stroke_itrsum"] = rs1, fs1, np.array(itr)
    return out


def stiff_case(ref, rc):
    rng = np.random.default_rng(20240603)
    ref.set_grid(rc["n"], rc["cell"], rc["FF"])
    ref.set_tip(rc["tip_lRadial"], rc["tip_kRadial"], rc["tip_rPP0"], rc["tip_kSpring"])
    tips = rc["tips"].reshape(-1, 3)
    rpps = np.concatenate([rc["rs_ps1"].reshape(-1, 3), rc["rs_ps1"].reshape(-1, 3) + rng.uniform(-0.3, 0.3, (len(tips), 3))])
    tips = np.concatenate([tips, tips])
    out = dict(tips=tips, rpps=rpps)
    for dd in (0.05, 0.02):
        ev, vecs = ref.stiffness(tips, rpps, which=3, ddisp=dd)
        tag = f"{int(round(dd * 100)):02d}"
        out[f"evals_{tag}"] = ev
        for k in range(3):
            out[f"evec{k + 1}_{tag}"] = vecs[k]
    return out


def main():
    ref = O.Reference()
    np.savez_compressed(os.path.join(HERE, "grid.npz"), **grid_case(ref))
    rc = relax_case(ref)
    np.savez_compressed(os.path.join(HERE, "relax.npz"), **rc)
    np.savez_compressed(os.path.join(HERE, "stiff.npz"), **stiff_case(ref, rc))
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main()
