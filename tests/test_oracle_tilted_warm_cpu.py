"""Pins of the oracle's warm method in relaxStrokesTilted semantics
(orc_relax_strokes_tilted_warm; the device counterpart is the mode-1 warm
column of relax_dev.hip):

* without Newton iterations it is the oracle's tilted FIRE bit for bit (that
  FIRE is itself pinned on hand-derived trajectories,
  tests/test_oracle_fire_pins_cpu.py);
* the Jacobian it solves with equals central differences of its force;
* its results meet the reference stop rule at stable points, lie within the
  stop rule's scatter of plain FIRE's, and take far fewer evaluations."""

import functools

import numpy as np

import oracle as O
from ppafm_amd import common as C
from ppafm_amd import molgen as M


def _rot_y(deg):
    t = np.deg2rad(deg)
    return np.array([[np.cos(t), 0, np.sin(t)], [0, 1, 0], [-np.sin(t), 0, np.cos(t)]])


@functools.lru_cache(maxsize=1)
def _setup():
    n = (40, 40, 80)
    lvec = np.array([[-2.0, -2.0, -2.0], [4.0, 0, 0], [0, 4.0, 0], [0, 0, 8.0]])
    xyz, Zs, qs = M.molecule(10, 21)
    cLJs = C.REA2LJ(C.getAtomsREA(8, Zs, C.loadSpecies(), alphaFac=-1.0))
    FE = O.ljc_qzs_ff(n, lvec, np.concatenate([xyz, qs[:, None]], 1), cLJs, np.array([-0.1, 0, 0, 0.0]), np.zeros(4))
    dinv = np.linalg.inv(lvec[1:] / np.array(n)[:, None]).T
    R = _rot_y(10.0)
    tipRot = np.zeros((3, 4))
    tipRot[:, :3] = R
    tipRot[2, 3] = -0.1
    st = np.array([0.25, 0.25, 0.0, 30.0]) / -C.eVA_Nm
    dpos0 = np.array([0.0, 0.0, -3.0, 3.0])
    ii, jj = np.meshgrid(np.linspace(-0.8, 0.8, 12), np.linspace(-0.8, 0.8, 12), indexing="ij")
    poss = np.zeros((144, 4))
    for c in range(3):
        poss[:, c] = (np.array([0.0, 0.0, 6.6]) - lvec[0])[c] + ii.ravel() * R[0, c] + jj.ravel() * R[1, c]
    return FE, dinv, tipRot, st, dpos0, poss


KW = dict(max_iters=128, conv_f2=float(np.float32(1e-8)), f2safe=float(np.float32(1e-8)))


def test_no_newton_is_tilted_fire_bit_for_bit():
    FE, dinv, tipRot, st, dpos0, poss = _setup()
    FEs, paths, nev = O.relax_tilted(FE, dinv, poss, 16, tipRot, st, dpos0, (0.5, 0.1), surfFF=0.01, **KW)
    FEw, pw, itw, nevw, nfb = O.relax_tilted_warm(FE, dinv, poss, 16, tipRot, st, dpos0, (0.5, 0.1), surfFF=0.01,
                                                  newton_max=0, newton_max0=0, retry=0, **KW)
    assert np.array_equal(paths, pw) and np.array_equal(FEs, FEw)
    assert np.array_equal(nev, nevw) and np.array_equal(nev, itw) and nfb == 0


def test_jacobian_equals_central_differences():
    FE, dinv, tipRot, st, dpos0, poss = _setup()
    rng = np.random.default_rng(2)
    tips = poss[:20, :3]
    rs = tips + tipRot[:, :3].T @ dpos0[:3] + rng.normal(0, 0.05, (20, 3))
    # keep each point and its +-h neighbours inside one grid cell, where the
    # trilinear force is linear along every axis
    g = rs @ dinv.T
    frac = g - np.floor(g)
    rs = rs + ((np.clip(frac, 0.01, 0.99) - frac) @ np.linalg.inv(dinv).T)
    f, J = O.tilted_force_jac(FE, dinv, tipRot, st, dpos0, tips, rs, surfFF=0.01)
    h = 1e-6
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        fp, _ = O.tilted_force_jac(FE, dinv, tipRot, st, dpos0, tips, rs + e, surfFF=0.01)
        fm, _ = O.tilted_force_jac(FE, dinv, tipRot, st, dpos0, tips, rs - e, surfFF=0.01)
        fd = (fp - fm) / (2 * h)
        assert np.abs(fd - J[:, :, k]).max() < 1e-6 * max(1.0, np.abs(J).max())


def test_warm_results_meet_the_stop_rule_at_stable_points():
    FE, dinv, tipRot, st, dpos0, poss = _setup()
    FEs, paths, nev = O.relax_tilted(FE, dinv, poss, 16, tipRot, st, dpos0, (0.5, 0.1), **KW)
    FEw, pw, itw, nevw, nfb = O.relax_tilted_warm(FE, dinv, poss, 16, tipRot, st, dpos0, (0.5, 0.1), **KW)
    tips = (poss[:, None, :3] + np.arange(16)[None, :, None] * (tipRot[2, :3] * tipRot[2, 3])).reshape(-1, 3)
    f, J = O.tilted_force_jac(FE, dinv, tipRot, st, dpos0, tips, pw[..., :3].reshape(-1, 3))
    settled = (itw.ravel() == 0) & ((f**2).sum(1) < KW["conv_f2"])
    print(f"settled by Newton {settled.mean():.3f}, fallbacks {nfb}, evaluations warm {nevw.sum()} plain {nev.sum()}")
    assert settled.mean() > 0.8
    assert all(O.stable3(j) for j in J[settled])
    d = np.linalg.norm(pw[..., :3] - paths[..., :3], axis=-1)
    assert np.percentile(d, 99) < 0.03
    assert nevw.sum() < 0.5 * nev.sum()


def test_residual_gate_selects_plain_fire_per_column():
    """The step-0 residual gate in relaxStrokesTilted semantics (as
    orc_relax_strokes_warm's, tests/test_oracle_cpu.py): every estimate gated
    equals warm starts without a step-0 Newton run bit for bit (one more
    evaluation at step 0 and one more fallback per column), none gated equals
    the ungated run, and a threshold between two columns' residuals splits
    the columns by their own residuals (the force from orc_tilted_force_jac)."""
    FE, dinv, tipRot, st, dpos0, poss = _setup()
    args = (FE, dinv, poss, 16, tipRot, st, dpos0, (0.5, 0.1))
    tips0 = poss[:, :3]
    rng = np.random.default_rng(5)
    starts = tips0 + tipRot[:, :3].T @ dpos0[:3] + rng.uniform(-1, 1, (len(poss), 3)) * np.array([0.4, 0.4, 0.2])
    ref = O.relax_tilted_warm(*args, newton_max0=0, surfFF=0.01, **KW)
    off = O.relax_tilted_warm(*args, starts0=starts, surfFF=0.01, **KW)
    every = O.relax_tilted_warm(*args, starts0=starts, res_max=1e-12, surfFF=0.01, **KW)
    none = O.relax_tilted_warm(*args, starts0=starts, res_max=1e6, surfFF=0.01, **KW)
    for a, b in zip(every[:3], ref[:3]):
        assert np.array_equal(a, b)
    assert (every[3][:, 0] == ref[3][:, 0] + 1).all() and np.array_equal(every[3][:, 1:], ref[3][:, 1:])
    assert every[4] == ref[4] + len(poss)
    for a, b in zip(none, off):
        assert np.array_equal(a, b)
    f, _ = O.tilted_force_jac(FE, dinv, tipRot, st, dpos0, tips0, starts, surfFF=0.01)
    res = np.sort(np.linalg.norm(f, axis=1))
    k = len(res) // 2
    thr = 0.5 * (res[k - 1] + res[k])
    assert res[k - 1] < thr < res[k]
    mid = O.relax_tilted_warm(*args, starts0=starts, res_max=thr, surfFF=0.01, **KW)
    gated = np.linalg.norm(f, axis=1) > thr
    assert gated.sum() == len(res) - k
    assert np.array_equal(mid[1][gated], ref[1][gated]) and np.array_equal(mid[1][~gated], off[1][~gated])
