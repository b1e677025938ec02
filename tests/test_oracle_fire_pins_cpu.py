"""Pins of the oracle's OpenCL-mode FIRE (orc_relax_strokes_tilted, the
restatement of relaxStrokesTilted, ppafm/ocl/cl/relax.cl:391-464) on
hand-derived trajectories (tests/fire_cases.py), and a check that each pin
fails when the matching line of oracle/pp_oracle.c is changed: the file is
copied with one planted mistake, compiled with gcc into a temporary
directory and run on the same cases."""

import pathlib
import shutil
import subprocess

import numpy as np
import pytest

import fire_cases as FC
import oracle as O

ORACLE_C = pathlib.Path(O.__file__).resolve().parent / "pp_oracle.c"
F2 = 1e-8  # F2CONV and F2SAFE of the restatement (FP64 values)


def run(c, lib=None):
    tipRot = np.zeros((3, 4))
    tipRot[:, :3] = c["R"]
    tipRot[2, 3] = c["step"]
    poss = np.append(c["tip0"], 0.0)[None]
    FEs, paths, nev = O.relax_tilted(c["FE"], np.eye(3) / FC.H, poss, c["nz"], tipRot, np.zeros(4), c["dpos0"],
                                     (FC.DT, FC.DAMP), max_iters=c["max_iters"], conv_f2=F2, f2safe=F2, _lib=lib)
    return FEs[0], paths[0, :, :3], nev[0]


def check(c, lib=None, tol=1e-9):
    """Paths (A), stored forces and iteration counts of the case; returns the
    list of mismatches (empty when the run follows the hand derivation)."""
    FEs, paths, nev = run(c, lib)
    bad = []
    if not np.array_equal(nev, c["iters"]):
        bad.append(f"iterations {nev} != {c['iters']}")
    if c["paths"] is not None and not (np.abs(paths - c["paths"]).max() <= tol):
        bad.append(f"paths off by {np.abs(paths - c['paths']).max():.3g} A")
    if c["paths"] is None and not np.isfinite(paths).all():
        bad.append("paths not finite")
    if not (np.abs(FEs - c["forces"]).max() <= 1e-9):
        bad.append(f"stored forces off by {np.abs(FEs - c['forces']).max():.3g}")
    return bad


def cases():
    return {
        "constant": FC.constant_field(F2),
        "stiff": FC.stiff_field(FC.DT * 0.1),
        "zero": FC.zero_field(),
        "weak": FC.weak_field(),
        "strong": FC.strong_field(),
    }


@pytest.mark.parametrize("name", ["constant", "stiff", "zero", "weak", "strong"])
def test_oracle_follows_hand_derivation(name):
    c = cases()[name]
    assert check(c, tol=1e-12 if name != "stiff" else 1e-9) == []


# ------------------------------------------------------------ planted mistakes
#
# (name, text of pp_oracle.c, replacement, cases that must fail).  Each
# replacement is a plausible misreading of relax.cl.

_LOOP_HEAD = ("        for (int iz = 0; iz < nz; iz++) {\n            double fe[4] = {0, 0, 0, 0};\n"
              "            double v[3] = {0, 0, 0};\n")
MUTANTS = [
    ("dt_reset_per_step", _LOOP_HEAD, _LOOP_HEAD + "            dt = dtmax;\n", ["stiff"]),
    ("alpha_reset_per_step", _LOOP_HEAD, _LOOP_HEAD + "            damp = damp0;\n", ["constant"]),
    ("velocity_kept_along_stroke", _LOOP_HEAD,
     "        double v[3] = {0, 0, 0};\n        for (int iz = 0; iz < nz; iz++) {\n            double fe[4] = {0, 0, 0, 0};\n",
     ["constant"]),
    ("no_dtmin_floor", "dtmin = dtmax * 0.1,", "dtmin = dtmax * 0.01,", ["stiff"]),
    ("dtmin_from_dt0_tenth_wrong", "dtmin = dtmax * 0.1,", "dtmin = dtmax * 0.2,", ["stiff"]),
    ("no_f2safe", "sqrt(vv / (ff2 + f2safe))", "sqrt(vv / ff2)", ["zero", "constant"]),
    ("f2safe_outside_sqrt", "sqrt(vv / (ff2 + f2safe))", "(sqrt(vv / ff2) + f2safe)", ["zero", "constant"]),
    ("cap_off_by_one", "for (int i = 0; i < max_iters; i++) {", "for (int i = 0; i <= max_iters; i++) {",
     ["stiff", "strong"]),
    ("force_after_final_move", "            long ind = (long)ip * nz + iz;\n",
     "            {\n                double hx = pos[0] * dinv[0] + pos[1] * dinv[1] + pos[2] * dinv[2];\n"
     "                double hy = pos[0] * dinv[3] + pos[1] * dinv[4] + pos[2] * dinv[5];\n"
     "                double hz = pos[0] * dinv[6] + pos[1] * dinv[7] + pos[2] * dinv[8];\n"
     "                fe_interp_xfast(FE, n, hx, hy, hz, fe);\n            }\n"
     "            long ind = (long)ip * nz + iz;\n", ["stiff"]),
    ("test_before_move", "                if (ff2 < conv_f2) break;\n", "", ["weak"]),
]


def _build(tmp, name, old, new):
    full = ORACLE_C.read_text()
    # the change goes into orc_relax_strokes_tilted only
    i = full.index("void orc_relax_strokes_tilted(int npix")
    j = full.index("/* ----", i)
    src = full[i:j]
    assert src.count(old) == 1, f"{name}: the line to change is not unique in orc_relax_strokes_tilted"
    if name == "test_before_move":
        anchor = "                double vf = v[0] * f[0] + v[1] * f[1] + v[2] * f[2];\n"
        assert src.count(anchor) == 1
        src = src.replace(old, new).replace(anchor, anchor + "                if (ff2 < conv_f2) break;\n")
    else:
        src = src.replace(old, new)
    src = full[:i] + src + full[j:]
    c_file = tmp / f"{name}.c"
    so = tmp / f"{name}.so"
    c_file.write_text(src)
    subprocess.run(["gcc", "-O2", "-fPIC", "-fopenmp", "-shared", "-std=c11", str(c_file), "-o", str(so), "-lm"], check=True,
                   capture_output=True)
    return O.load(so)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
@pytest.mark.parametrize("name,old,new,must_fail", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_planted_mistake_is_caught(tmp_path, name, old, new, must_fail):
    lib = _build(tmp_path, name, old, new)
    all_cases = cases()
    for cname in must_fail:
        assert check(all_cases[cname], lib) != [], f"planted '{name}' passes the '{cname}' case"
