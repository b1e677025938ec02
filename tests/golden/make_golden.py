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
  models.npz the further grid models on the grid.npz geometry (getVdWFF,
             getVdWFF_RE kinds 1-4, getMorseFF, getDFTD3FF, the three
             densities), getInPoints_LJ, evalRadialFF (kinds -3..4, default
             then custom damping constants, which persist), computeD3Coeffs
             on seeded synthetic tables (d3_tables() below), the spline
             subsamplers, test_force, test_eigen3x3, and relaxTipStrokes_omp
             and stiffnessMatrix with a radial tip spline on the relax.npz
             grid (ProbeParticle.cpp:379-631, 703-815)

core_symbols.txt  the C functions ppafm/core.py binds at import.

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


def d3_tables(seed=20240605):
    """Synthetic DFT-D3 tables of the reference's shapes (94 elements, 5
    reference coordination numbers, slots < 0 unused).  The probe element 8
    has all five slots valid; elements 1, 6 and 7 have 2, 3 and 4."""
    rng = np.random.default_rng(seed)
    ne, nr = 94, 5
    r_cov = rng.uniform(0.5, 1.6, ne)
    ref_cn = -np.ones((ne, nr))
    nvalid = rng.integers(1, nr + 1, ne)
    nvalid[[0, 5, 6, 7]] = [2, 3, 4, 5]
    for e in range(ne):
        ref_cn[e, : nvalid[e]] = np.sort(rng.uniform(0.0, 4.0, nvalid[e]))
    ref_c6 = rng.uniform(1.0, 60.0, (ne, ne, nr, nr))
    r4r2 = rng.uniform(1.5, 6.0, ne)
    r0 = rng.uniform(2.0, 5.0, (ne, ne))
    return dict(R_COV=r_cov, R0=r0, REF_CN=ref_cn, REF_C6=ref_c6, R4R2=r4r2, K=np.array([16.0, 4.0 / 3.0, 4.0]))


D3_PARAMS = np.array([1.0, 0.7875, 0.4289, 4.4407 * 0.52917721])  # s6, s8, a1, a2 [A]


def tip_spline(L=1.0, kR=30 / -16.0217662, c=-0.4, n=14):
    """Radial tip force k (x - L) + c (x - L)^3 on n knots over [0.3, 3.2] A,
    values and slopes."""
    xs = np.linspace(0.3, 3.2, n)
    ydys = np.column_stack([kR * (xs - L) + c * (xs - L) ** 3, kR + 3 * c * (xs - L) ** 2])
    return xs, np.ascontiguousarray(ydys)


def models_case(ref, gc, rc):
    rng = np.random.default_rng(20240604)
    n, cell, Rs = gc["n"], gc["cell"], gc["Rs"]
    na = len(Rs)
    shp = tuple(n)
    out = dict()
    REs = np.column_stack([rng.uniform(2.8, 3.6, na), np.sqrt(rng.uniform(1e-5, 5e-5, na))])
    cRAs = np.column_stack([rng.uniform(0.5, 2.0, na), rng.uniform(0.6, 1.4, na)])
    out.update(REs=REs, cRAs=cRAs)
    F = np.zeros(shp + (3,))
    E = np.zeros(shp)

    def run(tag, name, coefs, *extra):
        F[:] = 0
        E[:] = 0
        ref.set_grid(n, cell, F, E)
        ref.grid_model(name, Rs, coefs, *extra)
        out["F_" + tag], out["E_" + tag] = F.copy(), E.copy()

    run("vdw", "getVdWFF", gc["cLJs"])
    for kind in (1, 2, 3, 4):
        run(f"vdwre{kind}", "getVdWFF_RE", REs, kind, -1.0)
    run("morse", "getMorseFF", REs, -1.6)
    # D3: coefficients from synthetic tables for atoms of elements 1/6/7/8
    t = d3_tables()
    iZs = np.array([6, 1, 7, 8, 6], dtype=np.int32)
    Rd3 = np.array([[0.0, 0.0, 0.0], [1.09, 0.0, 0.0], [-0.7, 1.2, 0.1], [-0.6, -1.25, -0.1], [2.2, 1.3, 0.4]])
    out.update(d3_iZs=iZs, d3_Rs=Rd3, d3_params=D3_PARAMS)
    out["d3_coeffs"] = ref.d3_coeffs(Rd3, iZs, t, D3_PARAMS, 8)
    run("d3", "getDFTD3FF", out["d3_coeffs"])
    run("gauss", "getGaussDensity", cRAs)
    run("slater", "getSlaterDensity", cRAs)
    run("r4", "getDensityR4spline", cRAs)
    # points
    ps = rng.uniform(-0.5, 2.0, (40, 3))
    out["points"] = ps
    out["points_FE"] = ref.points_lj(ps, Rs, gc["cLJs"])
    # radial: default constants, then custom ones (these persist), then the
    # grid builders that read them again
    rs = np.linspace(0.8, 6.0, 60)
    out["radial_rs"] = rs
    cf = {-3: (3.4, 0.004), -2: (23.6, 2.4e4), -1: (23.6,), 0: (23.6,), 1: (3.4, 0.004), 2: (3.4, 0.004), 3: (3.4, 0.004), 4: (3.4, 0.004)}
    for kind in range(-3, 5):
        out[f"radial_E{kind}"], out[f"radial_F{kind}"] = ref.radial(rs, np.array(cf[kind]), kind)
    for kind, ad in ((0, 90.0), (1, 0.7), (2, 0.9), (3, 0.05), (4, 0.02)):
        out[f"radialc_E{kind}"], out[f"radialc_F{kind}"] = ref.radial(rs, np.array(cf[kind]), kind, ad)
    for kind in (1, 2, 3, 4):
        run(f"vdwre{kind}_after", "getVdWFF_RE", REs, kind, -1.0)
    run("vdw_after", "getVdWFF", gc["cLJs"])
    # splines
    kx = np.sort(rng.uniform(0.5, 4.0, 10))
    ydys = np.column_stack([np.sin(kx), np.cos(kx)])
    xs_u = np.sort(rng.uniform(1.0, 1.0 + 0.2 * 10.99, 50))
    ydys_u = np.column_stack([np.sin(1.0 + 0.2 * np.arange(12)), np.cos(1.0 + 0.2 * np.arange(12))])
    xs_n = np.concatenate([np.sort(rng.uniform(kx[0], kx[-1], 40)), rng.uniform(kx[0], kx[-1], 10)])
    out.update(spline_kx=kx, spline_ydys=ydys, spline_xs_u=xs_u, spline_ydys_u=ydys_u, spline_xs_n=xs_n)
    out["spline_u"] = ref.spline_uniform(1.0, 0.2, ydys_u, xs_u)
    out["spline_n"] = ref.spline_nonuniform(kx, ydys, xs_n)
    # test_force (type 2 uses setTip's spring) and test_eigen3x3
    sx, sy = tip_spline()
    out.update(tipspline_xs=sx, tipspline_ydys=sy)
    ref.set_tip(1.0, 25 / -16.0217662, np.array([0.0, 0.0, 0.0]), np.array([0.3, 0.3, 0.0]) / -16.0217662)
    r0, dr, R = np.array([0.1, -0.2, 0.6]), np.array([0.03, 0.02, 0.09]), np.array([0.0, 0.0, 0.0])
    out.update(tf_r0=r0, tf_dr=dr, tf_R=R)
    out["tf_type1"] = ref.test_force(1, 25, r0, dr, R, (sx, sy))
    out["tf_type2"] = ref.test_force(2, 25, r0, dr, R)
    mats = [rng.normal(size=(3, 3)), None, None]
    mats[0] = mats[0] + mats[0].T + 4 * np.eye(3)
    mats[1] = rng.normal(size=(3, 3))
    mats[2] = -(mats[0] + 10 * np.eye(3))
    out["eig_mats"] = np.array(mats)
    out["eig_out"] = np.array([ref.eigen3x3(m) for m in mats])
    # relaxation and stiffness with the tip spline on the relax.npz grid
    ref.set_grid(rc["n"], rc["cell"], rc["FF"])
    ref.set_tip(rc["tip_lRadial"], rc["tip_kRadial"], rc["tip_rPP0"], rc["tip_kSpring"])
    ref.set_relax(1000, 1e-8, 0.1, 0.1)
    rs_s, fs_s = ref.relax_strokes_spline(rc["tips"], (sx, sy))
    out["spline_rs"], out["spline_fs"] = rs_s, fs_s
    ev, vecs = ref.stiffness(rc["tips"].reshape(-1, 3), rs_s.reshape(-1, 3), which=3, ddisp=0.05, spline=(sx, sy))
    out["spline_evals"] = ev
    for k in range(3):
        out[f"spline_evec{k + 1}"] = vecs[k]
    return out


def core_symbols(path="/root/reference/ppafm/core.py"):
    """Names of the C functions ppafm/core.py binds when it is imported
    (lib.<name>.argtypes); a drop-in library must export every one."""
    import re

    names = sorted(set(re.findall(r"lib\.([A-Za-z_0-9]+)\.argtypes", open(path).read())))
    with open(os.path.join(HERE, "core_symbols.txt"), "w") as f:
        f.write("# C functions bound by ppafm/core.py at import (lib.<name>.argtypes)\n")
        f.write("\n".join(names) + "\n")


def main():
    core_symbols()
    ref = O.Reference()
    gc = grid_case(ref)
    np.savez_compressed(os.path.join(HERE, "grid.npz"), **gc)
    rc = relax_case(ref)
    np.savez_compressed(os.path.join(HERE, "relax.npz"), **rc)
    np.savez_compressed(os.path.join(HERE, "stiff.npz"), **stiff_case(ref, rc))
    np.savez_compressed(os.path.join(HERE, "models.npz"), **models_case(ref, gc, rc))
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main()
