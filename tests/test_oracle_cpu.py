"""Pins of the FP64 oracle (CPU only).

Each oracle function is checked against something that does not retype its
formula: the compiled reference C++ (golden fixtures from
tests/golden/make_golden.py and, where built, oracle/_ref live), the
reference's own golden df-weight vectors (tests/test_common.py:10-24), and
closed forms / invariants: F = -grad E, the LJ minimum at R_ij with depth
eps_ij, Coulomb's law, exact trilinear reproduction of affine fields and
periodicity, the analytic tip equilibrium, 1-D force balance by bisection,
and Giessibl's linear-force limit of the df weights."""

import os

import numpy as np
import pytest

import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ------------------------------------------------------------ golden fixtures


def test_grid_golden_bit_exact():
    g = np.load(os.path.join(GOLDEN, "grid.npz"))
    shp = tuple(g["n"])
    F = np.zeros(shp + (3,))
    E = np.zeros(shp)
    O.lj_ff(g["n"], g["cell"], g["Rs"], g["cLJs"], F=F, E=E)
    assert np.array_equal(F, g["F_lj"]) and np.array_equal(E, g["E_lj"])
    for kind in (0, 1, 2):
        F[:] = 0
        E[:] = 0
        O.coulomb_ff(g["n"], g["cell"], g["Rs"], g["kQQs"], kind=kind, F=F, E=E)
        assert np.array_equal(F, g[f"F_coul{kind}"]), kind
        assert np.array_equal(E, g[f"E_coul{kind}"]), kind


def _golden_tip(r):
    return O.tip_params(float(r["tip_lRadial"]), float(r["tip_kRadial"]), r["tip_rPP0"], r["tip_kSpring"])


@pytest.mark.parametrize("ps", [1, 0, -1])
def test_relax_golden_bit_exact(ps):
    r = np.load(os.path.join(GOLDEN, "relax.npz"))
    rs, fs, it, tot = O.relax_strokes(r["n"], r["cell"], r["FF"], _golden_tip(r), r["tips"], probeStart=ps)
    assert np.array_equal(rs, r[f"rs_ps{ps}"])
    assert np.array_equal(fs, r[f"fs_ps{ps}"])


def test_leapfrog_golden_bit_exact():
    r = np.load(os.path.join(GOLDEN, "relax.npz"))
    rs, fs, it, tot = O.relax_strokes(r["n"], r["cell"], r["FF"], _golden_tip(r), r["tips"], relax=(300, 1e-8, 0.1, 0.1, 1.1, 0.5, 0.99), relaxAlg=0)
    assert np.array_equal(rs, r["rs_alg0"]) and np.array_equal(fs, r["fs_alg0"])


def test_stroke_itrsum_golden():
    r = np.load(os.path.join(GOLDEN, "relax.npz"))
    tips = np.ascontiguousarray(r["tips"][2, 1])
    rs = np.zeros_like(tips)
    fs = np.zeros_like(tips)
    n = np.ascontiguousarray(r["n"], dtype=np.int32)
    tot = O.lib().orc_relax_stroke(
        n.ctypes.data_as(O._ip), O._ptr(np.ascontiguousarray(r["cell"])), O._ptr(r["FF"]), O._ptr(_golden_tip(r)),
        O._ptr(np.array(O.DEFAULT_RELAX, dtype=np.float64)), 1, 1, len(tips), O._ptr(tips), O._ptr(rs), O._ptr(fs), None, None,
    )
    assert tot == int(r["stroke_itrsum"])
    assert np.array_equal(rs, r["stroke_rs"])


@pytest.mark.skipif(not O.have_reference(), reason="oracle/_ref not built (reference sources absent)")
def test_live_reference_random_case():
    rng = np.random.default_rng(99)
    n = np.array([9, 12, 10], dtype=np.int32)
    cell = np.array([[0.9, 0.1, 0.0], [-0.1, 1.2, 0.05], [0.0, 0.0, 1.0]])
    Rs = rng.uniform(0, 1, (7, 3))
    cLJs = np.column_stack([rng.uniform(5, 30, 7), rng.uniform(1e3, 3e4, 7)])
    ref = O.Reference()
    F = np.zeros(tuple(n) + (3,))
    ref.set_grid(n, cell, F)
    ref.lj(Rs, cLJs)
    assert np.array_equal(F, O.lj_ff(n, cell, Rs, cLJs))


# ------------------------------------------------------------ closed forms


def _single_atom_grid(c6, c12, kq, model="lj", n=(24, 24, 24), h=0.1, atom=(1.23, 1.07, 1.31)):
    cell = np.diag([n[0] * h, n[1] * h, n[2] * h])
    Rs = np.array([atom])
    E = np.zeros(n)
    if model == "lj":
        F = O.lj_ff(n, cell, Rs, np.array([[c6, c12]]), E=E)
    else:
        F = O.coulomb_ff(n, cell, Rs, np.array([kq]), E=E)
    x = np.arange(n[0]) * h
    P = np.stack(np.meshgrid(x, np.arange(n[1]) * h, np.arange(n[2]) * h, indexing="ij"), axis=-1)
    return F, E, P - np.array(atom), h


def test_lj_force_is_minus_energy_gradient():
    F, E, D, h = _single_atom_grid(23.6, 2.4e4, 0)
    r = np.linalg.norm(D, axis=-1)
    # central differences of E along x, where the point and its neighbours are > 2 A away
    gx = -(E[2:, :, :] - E[:-2, :, :]) / (2 * h)
    sel = r[1:-1] > 2.0
    fx = F[1:-1, :, :, 0]
    # O(h^2) truncation of the central difference: relative agreement ~1e-2 at 0.1 A
    assert np.abs(gx[sel] - fx[sel]).max() <= 2e-2 * np.abs(fx[sel]).max()


def test_lj_minimum_at_rij_with_depth_eps():
    # c6 = 2 eps R^6, c12 = eps R^12 (common.py:851-857): E has its minimum -eps at r = R
    eps, R = 0.0058, 3.5692
    c6, c12 = 2 * eps * R**6, eps * R**12
    n = np.array([1, 1, 1], dtype=np.int32)
    for r, want_sign in ((R * 0.97, +1), (R * 1.03, -1)):
        cell = np.eye(3)
        F = np.zeros((1, 1, 1, 3))
        E = np.zeros((1, 1, 1))
        O.lj_ff(n, cell, np.array([[-r, 0.0, 0.0]]), np.array([[c6, c12]]), F=F, E=E)
        assert np.sign(F[0, 0, 0, 0]) == want_sign  # repulsive inside R, attractive outside
    F = np.zeros((1, 1, 1, 3))
    E = np.zeros((1, 1, 1))
    O.lj_ff(n, np.eye(3), np.array([[-R, 0.0, 0.0]]), np.array([[c6, c12]]), F=F, E=E)
    # R2SAFE = 1e-8 moves the zero by ~1e-9 A: |F| ~ 4e-11 eV/A there
    assert abs(F[0, 0, 0, 0]) < 1e-9 and abs(E[0, 0, 0] + eps) < 1e-12


def test_coulomb_law():
    F, E, D, h = _single_atom_grid(0, 0, 4.32, model="coul")
    r = np.sqrt((D**2).sum(-1) + (1e-8))
    sel = r > 0.5
    assert np.allclose(E[sel] * r[sel], 4.32, rtol=1e-7)
    Fn = np.linalg.norm(F, axis=-1)
    assert np.allclose(Fn[sel] * r[sel] ** 2, 4.32, rtol=1e-6)
    # repulsive for positive kqq: F parallel to d
    assert (np.einsum("...i,...i->...", F, D)[sel] > 0).all()


def test_limit_vec_field_keeps_direction():
    rng = np.random.default_rng(1)
    F = rng.normal(0, 10, (5, 4, 3, 3))
    G = O.limit_vec_field(F.copy(), 10.0)
    nF, nG = np.linalg.norm(F, axis=-1), np.linalg.norm(G, axis=-1)
    assert np.allclose(np.minimum(nF, 10.0), nG, rtol=1e-14)
    assert np.allclose(G / nG[..., None], F / nF[..., None], rtol=1e-13)


def test_interpolation_exact_for_affine_fields_and_periodic():
    rng = np.random.default_rng(5)
    n = np.array([7, 6, 9], dtype=np.int32)
    cell = np.array([[0.7, 0.0, 0.0], [0.1, 0.6, 0.0], [0.0, 0.05, 0.9]])
    dcell, dinv = O.grid_shape(n, cell)
    assert np.allclose(dcell @ dinv.T, np.eye(3), atol=1e-14)  # (dCell)(dCell^-T)^T = I
    A = rng.normal(size=(3, 3))
    b = rng.normal(size=3)
    idx = np.stack(np.meshgrid(*[np.arange(k) for k in n], indexing="ij"), axis=-1)
    P = idx @ dcell
    F = np.ascontiguousarray(P @ A.T + b)
    # inside the cell away from the wrapped last layer the interpolant is exact
    g = rng.uniform(0, 1, (200, 3)) * (n - 1.001)
    r = g @ dcell
    got = O.interp(n, cell, F, r)
    assert np.abs(got - (r @ A.T + b)).max() < 1e-12
    # periodic: r and r + cell rows give the same value, negative coordinates too
    r2 = rng.uniform(-3, 3, (200, 3))
    base = O.interp(n, cell, F, r2)
    for k in range(3):
        assert np.abs(O.interp(n, cell, F, r2 + cell[k]) - base).max() < 1e-12
        assert np.abs(O.interp(n, cell, F, r2 - 2 * cell[k]) - base).max() < 1e-12


def test_zero_field_relaxes_to_tip_equilibrium():
    n = np.array([10, 10, 10], dtype=np.int32)
    cell = np.eye(3) * 2.0
    F = np.zeros((10, 10, 10, 3))
    kS = np.array([0.25, 0.25, 0.0]) / -16.0217662
    tip = O.tip_params(3.0, 30 / -16.0217662, (0.2, -0.1, 0.0), kS)
    rpp0 = np.array([0.2, -0.1, -np.sqrt(9 - 0.04 - 0.01)])
    tips = np.zeros((2, 2, 5, 3))
    tips[..., 0] = 0.3 + 0.2 * np.arange(2)[:, None, None]
    tips[..., 1] = 0.4 + 0.3 * np.arange(2)[None, :, None]
    tips[..., 2] = 5.0 - 0.1 * np.arange(5)[None, None, :]
    rs, fs, it, tot = O.relax_strokes(n, cell, F, tip, tips, probeStart=1)
    assert np.abs(rs - (tips + rpp0)).max() < 1e-12  # starts at the minimum: no motion
    assert (it == 0).all()
    # probe left behind (probeStart -1): FIRE must walk to the minimum, |F| < 1e-4
    rs, fs, it, tot = O.relax_strokes(n, cell, F, tip, tips, probeStart=-1)
    kmin = 0.25 / 16.0217662
    assert np.abs(rs - (tips + rpp0)).max() < 1e-4 / kmin
    assert (it[..., 1:] > 0).all() and (it < 1000).all()


def test_on_axis_force_balance_by_bisection():
    # one atom under the tip on a grid node column: the probe stays on the axis
    # and stops where F_grid,z (piecewise linear in z) balances the springs
    h = 0.1
    n = np.array([20, 20, 60], dtype=np.int32)
    cell = np.diag([2.0, 2.0, 6.0])
    atom = np.array([[1.0, 1.0, 0.5]])
    cLJ = np.array([[23.6, 2.4e4]])
    F = O.limit_vec_field(O.lj_ff(n, cell, atom, cLJ), 10.0)
    krad = 30 / -16.0217662
    tip = O.tip_params(3.0, krad, (0.0, 0.0, 0.0), np.array([0.25, 0.25, 0.0]) / -16.0217662)
    tips = np.zeros((1, 1, 1, 3))
    tips[0, 0, 0] = (1.0, 1.0, 6.2)
    rs, fs, it, tot = O.relax_strokes(n, cell, F, tip, tips)
    col = F[10, 10, :, 2]

    def fz(z):  # periodic piecewise-linear grid force along the column + radial spring
        g = z / h
        i = int(np.floor(g))
        t = g - i
        fg = col[i % 60] * (1 - t) + col[(i + 1) % 60] * t
        d = z - 6.2
        return fg + krad * (abs(d) - 3.0) * np.sign(d)

    lo, hi = 2.0, 4.5
    assert fz(lo) * fz(hi) < 0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if fz(lo) * fz(mid) <= 0:
            hi = mid
        else:
            lo = mid
    z_star = 0.5 * (lo + hi)
    assert abs(rs[0, 0, 0, 0] - 1.0) < 1e-12 and abs(rs[0, 0, 0, 1] - 1.0) < 1e-12
    # |F| < 1e-4 at the stop; the local stiffness is > 1 eV/A^2 here
    assert abs(rs[0, 0, 0, 2] - z_star) < 1e-4


# ------------------------------------------------------------------- df


def test_df_weights_reference_golden_vectors():
    # ppafm tests/test_common.py:10-24
    assert np.allclose(O.get_df_weight(0.1), [-5.0, 5.0])
    assert np.allclose(O.get_df_weight(1.0, dz=0.2), [-0.35594622, -0.22193265, -0.05447093, 0.05447093, 0.22193265, 0.35594622])
    assert np.allclose(O.get_simple_df_weight(n=5, dz=0.2), [-0.31362841, -0.37274317, 0, 0.37274317, 0.31362841])


def test_df_weights_linear_force_limit():
    # For F = c z the Giessibl integral is exact: df = -f0/(2 k0) dF/dz, i.e.
    # the correlation gives -c/2 for every amplitude; a constant force gives 0.
    for A, dz in ((0.8, 0.1), (1.0, 0.2), (0.35, 0.1)):
        W = O.get_df_weight(A, dz)
        assert abs(W.sum()) < 1e-12
        z = np.arange(40) * dz
        conv = np.convolve(1.7 * z, W, mode="valid")
        assert np.allclose(conv, -0.85, atol=1e-12)


def test_fz2df_matches_brute_force():
    rng = np.random.default_rng(3)
    Fz = rng.normal(size=(3, 4, 25))
    W = O.get_df_weight(0.8, 0.1)
    got = O.fz2df(Fz, 0.1, 1800.0, 30300.0, amplitude=0.8)
    n = len(W)
    want = np.zeros((3, 4, 25 - n + 1))
    for j in range(25 - n + 1):
        for k in range(n):
            want[..., j] += Fz[..., j + k] * W[n - 1 - k]
    assert np.allclose(got, want * 16.0217656 * 30300.0 / 1800.0, rtol=1e-13, atol=1e-13)


def test_planted_sign_error_is_caught():
    # a plausible mistake (dropping the reversal of the kernel) fails the
    # reference golden vectors
    w = O.get_df_weight(1.0, dz=0.2)[::-1]
    assert not np.allclose(w, [-0.35594622, -0.22193265, -0.05447093, 0.05447093, 0.22193265, 0.35594622])


# ------------------------------------------------- OpenCL-mode restatements
#
# ljc_qzs_ff, relax_tilted and convolve_z restate FF.cl / relax.cl, which
# cannot run here.  They are pinned against the CPU oracle (itself bit-exact
# with the compiled reference) through identities that hold between the two
# semantics, and against closed-form equilibria.


def _rot(ax, deg):
    t = np.deg2rad(deg)
    c, s = np.cos(t), np.sin(t)
    if ax == "x":
        return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])
    if ax == "y":
        return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


def _far_mask(n, lvec, xyz, rmin):
    idx = np.stack(np.meshgrid(*[np.arange(k) for k in n], indexing="ij"), axis=-1)
    d = np.array([lvec[1] / n[0], lvec[2] / n[1], lvec[3] / n[2]])
    P = lvec[0] + idx @ d
    r = np.linalg.norm(P[..., None, :] - xyz[None, None, None, :, :], axis=-1).min(-1)
    return r > rmin


def _ocl_case():
    rng = np.random.default_rng(21)
    xyz = rng.uniform(-1.0, 1.0, (5, 3))
    q = rng.uniform(-0.3, 0.3, 5)
    cLJs = np.stack([rng.uniform(1, 3, 5), rng.uniform(50, 150, 5)], axis=1)
    n = (12, 10, 14)
    lvec = np.array([[-2.0, -1.8, -2.2], [4.0, 0.0, 0.0], [0.2, 3.6, 0.0], [0.0, 0.1, 4.4]])
    return xyz, q, cLJs, n, lvec


def test_ljc_lj_part_matches_cpu_lj():
    xyz, q, cLJs, n, lvec = _ocl_case()
    FE = O.ljc_qzs_ff(n, lvec, np.concatenate([xyz, 0 * q[:, None]], 1), cLJs, np.zeros(4), np.zeros(4))
    F = np.zeros(n + (3,))
    E = np.zeros(n)
    O.lj_ff(n, lvec[1:], xyz - lvec[0], cLJs, F=F, E=E)
    m = _far_mask(n, lvec, xyz, 1.0)
    rel = np.linalg.norm(FE[..., :3] - F, axis=-1) / np.maximum(np.linalg.norm(F, axis=-1), 1e-6)
    assert rel[m].max() < 3e-5  # R2SAFE 1e-6 vs 1e-8 only
    # energy relative to the attractive term's size (E crosses zero)
    E6 = np.zeros(n)
    O.lj_ff(n, lvec[1:], xyz - lvec[0], cLJs * np.array([1.0, 0.0]), F=np.zeros(n + (3,)), E=E6)
    assert (np.abs(FE[..., 3] - E) / (np.abs(E) + np.abs(E6)))[m].max() < 3e-5


def test_ljc_charge_stencil_equals_shifted_cpu_coulomb():
    xyz, q, cLJs, n, lvec = _ocl_case()
    Qs, QZs = np.array([-10.0, 20.0, -10.0, 0.0]), np.array([0.1, 0.0, -0.1, 0.0])
    FE = O.ljc_qzs_ff(n, lvec, np.concatenate([xyz, q[:, None]], 1), 0 * cLJs, Qs, QZs)
    Fx = np.zeros(n + (3,))
    Ex = np.zeros(n)
    for Qk, dz in zip(Qs, QZs):
        if Qk == 0:
            continue
        F = np.zeros(n + (3,))
        E = np.zeros(n)
        # the field at p + dz*z from atom R equals the field at p from R - dz*z
        O.coulomb_ff(n, lvec[1:], xyz - lvec[0] - np.array([0, 0, dz]), q * 14.399645, kind=0, F=F, E=E)
        Fx += Qk * (14.399644 / 14.399645) * F
        Ex += Qk * (14.399644 / 14.399645) * E
    m = _far_mask(n, lvec, xyz, 1.0)
    scale = np.abs(Fx).max()
    assert np.abs(FE[..., :3] - Fx)[m].max() < 3e-5 * scale
    assert np.abs(FE[..., 3] - Ex)[m].max() < 3e-5 * np.abs(Ex).max()


def _tilted(FE, dinv, start, nz, R, stiffness, dpos0, surf=0.0, step=-0.1):
    tipRot = np.zeros((3, 4))
    tipRot[:, :3] = R
    tipRot[2, 3] = step
    poss = np.array([[start[0], start[1], start[2], 0.0]])
    return O.relax_tilted(FE, dinv, poss, nz, tipRot, np.asarray(stiffness, float), np.asarray(dpos0, float), (0.5, 0.1),
                          surfFF=surf, max_iters=4000)


def test_tilted_surface_force_equilibrium_on_tip_axis():
    # no grid force, a constant force along the tip axis C: the probe stays on
    # the axis where the radial spring balances it, |d| = L + surf / krad
    FE = np.zeros((8, 8, 8, 4))
    dinv = np.eye(3) * 10.0
    R = _rot("y", 10.0) @ _rot("z", 25.0)
    L, krad, klat, surf = 3.0, 30 / -16.0217662, 0.25 / -16.0217662, 0.02
    tp0 = np.array([0.3, 0.2, 5.0])
    FEs, paths, nev = _tilted(FE, dinv, tp0, 4, R, (klat, klat, 0.0, krad), (0.0, 0.0, -L, L), surf=surf)
    C = R[2]
    for s in range(4):
        tp = tp0 + s * (-0.1) * C
        expect = tp - (L + surf / krad) * C
        assert np.abs(paths[0, s, :3] - expect).max() < 1e-3, s
    assert nev[0, 0] > 1  # the first step has to move


def test_tilted_constant_field_equilibrium_by_energy_gradient():
    from scipy.optimize import fsolve

    F0 = np.array([0.02, -0.015, 0.03])
    FE = np.zeros((8, 8, 8, 4))
    FE[..., :3] = F0
    dinv = np.eye(3) * 10.0
    R = _rot("x", 10.0) @ _rot("z", 20.0)
    L, krad, klat = 3.0, 30 / -16.0217662, 1.0 / -16.0217662
    k = np.array([klat, klat, 0.0])
    d0 = np.array([0.1, -0.05, -np.sqrt(L**2 - 0.0125)])

    def energy(dq):  # spring energy in the tip frame (k < 0 restores)
        return -0.5 * np.sum(k * (dq - d0) ** 2) - 0.5 * krad * (np.linalg.norm(dq) - L) ** 2

    def residual(dq, h=1e-6):
        g = np.array([(energy(dq + h * e) - energy(dq - h * e)) / (2 * h) for e in np.eye(3)])
        return R @ F0 - g

    dq = fsolve(residual, d0, xtol=1e-13)
    tp0 = np.array([0.2, 0.1, 5.0])
    FEs, paths, nev = _tilted(FE, dinv, tp0, 3, R, (klat, klat, 0.0, krad), (d0[0], d0[1], d0[2], L))
    for s in range(3):
        tp = tp0 + s * (-0.1) * R[2]
        assert np.abs(paths[0, s, :3] - (tp + R.T @ dq)).max() < 3e-3, s
        # the stored force is the grid force in the tip frame
        assert np.allclose(FEs[0, s, :3], R @ F0, atol=1e-12)


def test_tilted_affine_field_sampling_layout():
    # springs off, an affine grid force -K (r - r0) on a sheared cell: every
    # stroke step ends at r0, which pins dinv and the x-fastest layout
    n = np.array([9, 8, 10])
    cell = np.array([[1.8, 0.0, 0.0], [0.3, 1.6, 0.0], [0.0, 0.2, 2.0]])
    dcell = cell / n[:, None]
    dinv = np.linalg.inv(dcell).T
    K = np.diag([0.5, 0.4, 0.6])
    r0 = np.array([0.9, 0.8, 1.0])
    idx = np.stack(np.meshgrid(*[np.arange(v) for v in n], indexing="ij"), axis=-1)
    P = idx @ dcell
    FE = np.zeros(tuple(n) + (4,))
    FE[..., :3] = -(P - r0) @ K.T
    R = _rot("y", 10.0)
    FEs, paths, nev = _tilted(FE, dinv, r0 + np.array([0.1, -0.1, 3.1]), 3, R, (0, 0, 0, 0), (0.0, 0.0, -3.0, 3.0), step=-0.05)
    assert np.abs(paths[0, :, :3] - r0).max() < 5e-4


def test_convolve_z_brute_force():
    rng = np.random.default_rng(4)
    F = rng.normal(size=(3, 4, 17))
    w = rng.normal(size=6)
    got = O.convolve_z(F, w)
    ref = np.zeros((3, 4, 12))
    for j in range(12):
        for kk in range(6):
            ref[..., j] += F[..., j + kk] * w[kk]
    assert np.allclose(got, ref, atol=1e-13)


# ------------------------------------------------------- warm-start network


def test_warmstart_backward_matches_finite_differences():
    import warmstart as W

    rng = np.random.default_rng(3)
    theta = W.init(1) + rng.normal(0, 0.05, W.NPARAM)
    P = rng.normal(size=(7, 5, 5, 3, 3))
    T = rng.normal(size=(7, 3))
    L, g = W.loss_grad(theta, P, T)
    idx = rng.choice(W.NPARAM, 60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros(W.NPARAM)
        e[i] = h
        fd = (W.loss_grad(theta + e, P, T)[0] - W.loss_grad(theta - e, P, T)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_warmstart_layers_are_valid_convolutions():
    # conv1 output at (ox, oy) only sees patch cells ox..ox+2, oy..oy+2
    import warmstart as W

    rng = np.random.default_rng(4)
    theta = W.init(2)
    P = rng.normal(size=(1, 5, 5, 3, 3))
    out, c = W.forward(theta, P)
    P2 = P.copy()
    P2[0, 4, 4] += 1.0  # only the (2, 2) window contains (4, 4)
    _, c2 = W.forward(theta, P2)
    d = np.abs(c2["Z1"] - c["Z1"]).max(-1)[0]
    assert (d.reshape(3, 3) > 0).sum() == 1 and d[8] > 0


def test_bf16_rounding():
    import warmstart as W

    x = np.array([1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-9, -2.5, 3.0e-3])
    r = W.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0  # ties to even
    assert r[2] == 1.0 + 2**-7
    assert r[3] == -2.5
    assert abs(r[4] - 3.0e-3) <= 3.0e-3 * 2**-8


# ------------------------------------------------- warm starts (Newton + FIRE)


def _bump_field(n=(32, 32, 40), h=0.25, A=0.2, sx=1.0, sz=2.0, c=(4.0, 4.0, 2.0)):
    """F = -grad E on grid nodes for a repulsive Gaussian bump E =
    A exp(-((x-cx)^2 + (y-cy)^2)/sx^2 - (z-cz)^2/sz^2): on its axis the lateral
    curvature is -2A/sx^2 (times the z factor), far beyond the lateral spring,
    so the on-axis equilibrium of a probe above it is a saddle."""
    g = np.stack(np.meshgrid(*[np.arange(k) * h for k in n], indexing="ij"), -1)
    d = g - np.array(c)
    E = A * np.exp(-(d[..., 0] ** 2 + d[..., 1] ** 2) / sx ** 2 - d[..., 2] ** 2 / sz ** 2)
    F = np.stack([2 * d[..., 0] / sx ** 2 * E, 2 * d[..., 1] / sx ** 2 * E, 2 * d[..., 2] / sz ** 2 * E], -1)
    cell = np.diag([n[0] * h, n[1] * h, n[2] * h])
    return np.ascontiguousarray(F), cell


def _tip():
    from cases import tip_params

    return tip_params()


def test_warm_jacobian_matches_finite_differences():
    from cases import case

    cs = case(0, small=True)
    n, cell, F = cs["cfg"]["n"], cs["lvec"][1:], cs["FF"]
    rng = np.random.default_rng(5)
    h = 0.1
    # points inside cells (fractions 0.2..0.8) so that +-eps stays in one cell
    nodes = rng.integers(10, 50, size=(40, 3))
    r = (nodes + rng.uniform(0.2, 0.8, size=(40, 3))) * h
    tips = r + np.array([0.0, 0.0, 3.0]) + rng.normal(0, 0.3, size=(40, 3))
    f, J = O.force_jac(n, cell, F, _tip(), tips, r)
    eps = 1e-6
    Jfd = np.zeros_like(J)
    for a in range(3):
        dr = np.zeros(3)
        dr[a] = eps
        fp, _ = O.force_jac(n, cell, F, _tip(), tips, r + dr)
        fm, _ = O.force_jac(n, cell, F, _tip(), tips, r - dr)
        Jfd[:, :, a] = (fp - fm) / (2 * eps)
    scale = np.abs(J).max(axis=(1, 2), keepdims=True)
    assert np.abs(J - Jfd).max() / scale.min() < 1e-6
    # the force part is the relaxation's own total force (pp_force): a zero
    # field with the probe at tip + rPP0 gives F = 0
    z = np.zeros_like(F)
    f0, _ = O.force_jac(n, cell, z, _tip(), tips, tips + np.array([0.0, 0.0, -3.0]))
    assert np.abs(f0).max() < 1e-12


def test_stable3_matches_eigenvalues():
    rng = np.random.default_rng(1)
    hits = 0
    for _ in range(4000):
        J = rng.normal(size=(3, 3)) - rng.uniform(0, 2) * np.eye(3)
        re = np.linalg.eigvals(J).real.max()
        if abs(re) < 1e-6:
            continue
        assert O.stable3(J) == (re < 0)
        hits += re < 0
    assert 200 < hits < 3800  # both outcomes exercised


def test_warm_without_newton_is_plain_fire():
    from cases import case
    from ppafm_amd import molgen as M

    cs = case(0, small=True)
    n, cell, F = cs["cfg"]["n"], cs["lvec"][1:], cs["FF"]
    tips = (M.scan_tips(cs["cfg"]) - cs["lvec"][0])[:6, :6]
    rs, fs, it, _ = O.relax_strokes(n, cell, F, _tip(), tips)
    rw, fw, iw, nev, nfb = O.relax_strokes_warm(n, cell, F, _tip(), tips, newton_max=0, newton_max0=0, retry=0)
    assert np.array_equal(rs, rw) and np.array_equal(fs, fw) and np.array_equal(it, iw) and nfb == 0


def test_warm_results_are_stable_minima():
    from cases import case
    from ppafm_amd import molgen as M

    cs = case(0, small=True)
    n, cell, F = cs["cfg"]["n"], cs["lvec"][1:], cs["FF"]
    tips = (M.scan_tips(cs["cfg"]) - cs["lvec"][0])[::3, ::3]
    rs, _, it, _ = O.relax_strokes(n, cell, F, _tip(), tips)
    rw, _, iw, nev, nfb = O.relax_strokes_warm(n, cell, F, _tip(), tips)
    npos = rw.size // 3
    # Newton falls back to FIRE only where the continuation start is already
    # in a non-convex region (deep in the approach of this rough case)
    assert nfb < 0.1 * npos
    f, J = O.force_jac(n, cell, F, _tip(), tips.reshape(-1, 3), rw.reshape(-1, 3))
    fn = np.linalg.norm(f, axis=1)
    # the positions a Newton iterate or FIRE's first evaluation at a Newton
    # zero ended (no FIRE move): the reference's stop rule holds there, at a
    # stable point (FIRE itself, in the few fallbacks, may stop at any point
    # that meets the rule, as the reference's does)
    conv = fn**2 < 1e-8
    assert conv.mean() > 0.9
    newton_ended = conv & (iw.reshape(-1) == 0)
    assert all(O.stable3(Jk) for Jk in J[newton_ended])
    assert newton_ended.sum() > 0.9 * conv.sum()
    # the sample after FIRE's stop (|F|^2 < 1e-8 before its last move)
    assert np.percentile(fn, 99) < 1e-3
    # plain FIRE stops within its stop-rule scatter of the same minima
    d = np.linalg.norm(rw - rs, axis=-1)
    assert np.percentile(d, 99) < 0.03
    # with several times fewer force evaluations (FIRE: iterations + 1 each)
    assert nev.sum() < 0.3 * (it.sum() + npos)


def test_residual_gate_selects_plain_fire_per_column():
    """The step-0 residual gate: where the total force at a column's
    estimate is larger than res_max, step 0 is plain FIRE from the plain
    start.  A gate every estimate fails (res_max tiny) reproduces warm starts
    without a step-0 Newton run bit for bit, with one more evaluation per
    column (the gate's) and one more fallback per column; a gate none fails
    (res_max huge) reproduces the ungated run; a threshold between two
    columns' residuals splits the columns into exactly those two runs, by
    their own residuals (the force at the estimates from orc_force_jac)."""
    from cases import case
    from ppafm_amd import molgen as M

    cs = case(0, small=True)
    n, cell, F = cs["cfg"]["n"], cs["lvec"][1:], cs["FF"]
    tips = (M.scan_tips(cs["cfg"]) - cs["lvec"][0])[::3, ::3]
    nx, ny, nz, _ = tips.shape
    rs, _, _, _ = O.relax_strokes(n, cell, F, _tip(), tips)
    rng = np.random.default_rng(11)
    # estimates: step 0's relaxed positions moved by up to 0.4 A laterally and 0.2 A along z
    starts = rs[:, :, 0] + rng.uniform(-1, 1, (nx, ny, 3)) * np.array([0.4, 0.4, 0.2])
    ref = O.relax_strokes_warm(n, cell, F, _tip(), tips, newton_max0=0)
    off = O.relax_strokes_warm(n, cell, F, _tip(), tips, starts0=starts)
    every = O.relax_strokes_warm(n, cell, F, _tip(), tips, starts0=starts, res_max=1e-12)
    none = O.relax_strokes_warm(n, cell, F, _tip(), tips, starts0=starts, res_max=1e6)
    for a, b in zip(every[:3], ref[:3]):
        assert np.array_equal(a, b)
    assert (every[3][:, :, 0] == ref[3][:, :, 0] + 1).all() and np.array_equal(every[3][:, :, 1:], ref[3][:, :, 1:])
    assert every[4] == ref[4] + nx * ny
    for a, b in zip(none, off):
        assert np.array_equal(a, b)
    f, _ = O.force_jac(n, cell, F, _tip(), tips[:, :, 0].reshape(-1, 3), starts.reshape(-1, 3))
    res = np.sort(np.linalg.norm(f, axis=1))
    k = len(res) // 2
    thr = 0.5 * (res[k - 1] + res[k])
    assert res[k - 1] < thr < res[k]
    mid = O.relax_strokes_warm(n, cell, F, _tip(), tips, starts0=starts, res_max=thr)
    gated = (np.linalg.norm(f, axis=1) > thr).reshape(nx, ny)
    assert gated.sum() == len(res) - k
    assert np.array_equal(mid[0][gated], ref[0][gated]) and np.array_equal(mid[0][~gated], off[0][~gated])
    assert np.array_equal(mid[2][gated], ref[2][gated]) and np.array_equal(mid[2][~gated], off[2][~gated])


def test_warm_rejects_saddle_on_a_repulsive_bump():
    F, cell = _bump_field()
    n = F.shape[:3]
    tip = _tip()
    # strokes just off the bump's axis: the lateral instability pushes the
    # probe sideways; Newton from the near-axis start heads for the on-axis
    # saddle, which the stability test rejects
    nz = 12
    tz = 2.0 + 3.0 + 2.0 - 0.1 * np.arange(nz)
    off = [0.002, 0.01, 0.05]
    tips = np.zeros((len(off), 1, nz, 3))
    for a, dx in enumerate(off):
        tips[a, 0, :, 0] = 4.0 + dx
        tips[a, 0, :, 1] = 4.0
        tips[a, 0, :, 2] = tz
    rs, _, _, _ = O.relax_strokes(n, cell, F, tip, tips)
    rw, _, _, _, nfb = O.relax_strokes_warm(n, cell, F, tip, tips, newton_max0=0)
    f, J = O.force_jac(n, cell, F, tip, tips[0, 0, :1], np.array([[4.0, 4.0, rs[0, 0, 0, 2]]]))
    assert not O.stable3(J[0])  # the on-axis point is a saddle
    assert nfb > 0
    # the probe leaves the axis as plain FIRE does, and ends where FIRE ends
    assert (np.abs(rw[..., 0] - 4.0)[:, :, -1] > 0.1).all()
    assert np.abs(rw - rs).max() < 0.03


# ------------------------------------------------------------ stiffness


@pytest.mark.parametrize("tag", ["05", "02"])
def test_stiffness_golden_bit_exact(tag):
    r = np.load(os.path.join(GOLDEN, "relax.npz"))
    s = np.load(os.path.join(GOLDEN, "stiff.npz"))
    ev, vecs, nnan = O.stiffness(r["n"], r["cell"], r["FF"], _golden_tip(r), s["tips"], s["rpps"], which=3,
                                 ddisp=int(tag) / 100)
    assert nnan == 0
    assert np.array_equal(ev, s[f"evals_{tag}"])
    for k in range(3):
        assert np.array_equal(vecs[k], s[f"evec{k + 1}_{tag}"]), k


@pytest.mark.skipif(not O.have_reference(), reason="oracle/_ref not built (reference sources absent)")
def test_stiffness_live_reference_random_case():
    rng = np.random.default_rng(7)
    n = np.array([14, 12, 30], dtype=np.int32)
    cell = np.array([[1.4, 0.0, 0.0], [0.2, 1.2, 0.0], [0.0, 0.0, 3.0]])
    Rs = np.column_stack([rng.uniform(0.2, 1.4, 4), rng.uniform(0.2, 1.2, 4), np.full(4, 0.4)])
    cLJs = np.tile([23.6, 2.4e4], (4, 1)) * rng.uniform(0.8, 1.2, (4, 1))
    F = O.limit_vec_field(O.lj_ff(n, cell, Rs, cLJs), 10.0)
    tip = O.tip_params(1.2, 25 / -O.eVA_Nm, (0.1, 0.0, 0.0), np.array([0.4, 0.3, 0.0]) / -O.eVA_Nm)
    tips = np.column_stack([rng.uniform(0, 1.5, 40), rng.uniform(0, 1.3, 40), rng.uniform(2.3, 2.9, 40)])
    rpps = tips + np.array([0.1, 0.0, -1.2]) + rng.uniform(-0.2, 0.2, (40, 3))
    ref = O.Reference()
    ref.set_grid(n, cell, F)
    ref.set_tip(1.2, 25 / -O.eVA_Nm, np.array([0.1, 0.0, 0.0]), np.array([0.4, 0.3, 0.0]) / -O.eVA_Nm)
    for which in (0, 2, 3):
        evr, vr = ref.stiffness(tips, rpps, which=which, ddisp=0.03)
        ev, vecs, nnan = O.stiffness(n, cell, F, tip, tips, rpps, which=which, ddisp=0.03)
        assert np.array_equal(ev, evr)
        for k in range(which):
            assert np.array_equal(vecs[k], vr[k]), (which, k)


def _affine_grid(M, F0, n=(20, 20, 20), h=0.1):
    cell = np.diag([n[0] * h, n[1] * h, n[2] * h])
    idx = np.stack(np.meshgrid(*[np.arange(k) * h for k in n], indexing="ij"), axis=-1)
    return cell, np.ascontiguousarray(F0 + idx @ M.T)


def test_stiffness_affine_field_closed_form():
    """Field F0 + M r (interpolated exactly away from the wrap) with a linear
    tip (kRadial = 0, springs kS): the dynamical matrix is exactly
    -(M + M^T)/2 - diag(kS) whatever ddisp, so its eigen-decomposition is
    pinned by numpy's symmetric eigensolver."""
    rng = np.random.default_rng(11)
    M = rng.normal(0, 0.5, (3, 3))
    F0 = rng.normal(0, 0.2, 3)
    cell, F = _affine_grid(M, F0)
    ks = np.array([-0.4, -0.25, -0.1])
    tip = O.tip_params(1.0, 0.0, (0.05, -0.02, -0.9), ks)
    rpps = rng.uniform(0.6, 1.3, (25, 3))
    tips = rpps + rng.uniform(-0.5, 0.5, (25, 3))
    S = -(M + M.T) / 2 - np.diag(ks)
    w, V = np.linalg.eigh(S)
    for dd in (0.05, 0.01):
        ev, vecs, nnan, D = O.stiffness((20, 20, 20), cell, F, tip, tips, rpps, which=3, ddisp=dd, want_dynmat=True)
        assert nnan == 0
        assert np.abs(D - S).max() < 1e-12
        # the reference's decimal 1/3 and sqrt(3) cost ~1e-11 relative
        assert np.abs(ev - w[None, :]).max() < 1e-9 * np.abs(w).max()
        for k in range(3):
            assert np.abs(np.abs(vecs[k] @ V[:, k]) - 1).max() < 1e-9
            assert np.abs(np.linalg.norm(vecs[k], axis=1) - 1).max() < 1e-12


# ------------------------------------------------- further force-field models


def _models():
    return np.load(os.path.join(GOLDEN, "models.npz")), np.load(os.path.join(GOLDEN, "grid.npz"))


GRID_MODELS = [("vdw", "vdw", "cLJs", 180.0), ("vdwre1", "vdw_r2", "REs", 0.5), ("vdwre2", "vdw_r4", "REs", 0.5),
               ("vdwre3", "vdw_invr4", "REs", 0.03), ("vdwre4", "vdw_invr8", "REs", 0.01), ("morse", "morse", "REs", -1.6),
               ("d3", "d3", "d3_coeffs", 0.0), ("gauss", "gauss", "cRAs", 0.0), ("slater", "slater", "cRAs", 0.0),
               ("r4", "r4", "cRAs", 0.0),
               # after evalRadialFF replaced the damping constants (they persist)
               ("vdwre1_after", "vdw_r2", "REs", 0.7), ("vdwre2_after", "vdw_r4", "REs", 0.9),
               ("vdwre3_after", "vdw_invr4", "REs", 0.05), ("vdwre4_after", "vdw_invr8", "REs", 0.02),
               ("vdw_after", "vdw", "cLJs", 90.0)]


@pytest.mark.parametrize("tag,name,coef,par", GRID_MODELS)
def test_model_grids_golden_bit_exact(tag, name, coef, par):
    m, g = _models()
    coefs = g["cLJs"] if coef == "cLJs" else m[coef]
    F = np.zeros(tuple(g["n"]) + (3,))
    E = np.zeros(tuple(g["n"]))
    O.model_ff(name, g["n"], g["cell"], g["Rs"], coefs, par, F=F, E=E)
    assert np.array_equal(F, m["F_" + tag]) and np.array_equal(E, m["E_" + tag])


def test_radial_golden_bit_exact():
    m, _ = _models()
    cf = {-3: (3.4, 0.004), -2: (23.6, 2.4e4), -1: (23.6,), 0: (23.6,), 1: (3.4, 0.004), 2: (3.4, 0.004), 3: (3.4, 0.004), 4: (3.4, 0.004)}
    for kind in range(-3, 5):
        E, F = O.radial(kind, m["radial_rs"], cf[kind])
        assert np.array_equal(E, m[f"radial_E{kind}"]) and np.array_equal(F, m[f"radial_F{kind}"]), kind
    for kind, ad in ((0, 90.0), (1, 0.7), (2, 0.9), (3, 0.05), (4, 0.02)):
        E, F = O.radial(kind, m["radial_rs"], cf[kind], adamp=ad)
        assert np.array_equal(E, m[f"radialc_E{kind}"]) and np.array_equal(F, m[f"radialc_F{kind}"]), kind


def test_d3_coeffs_golden():
    import sys

    sys.path.insert(0, GOLDEN)
    import make_golden as MG

    m, _ = _models()
    c = O.d3_coeffs(m["d3_Rs"], m["d3_iZs"], MG.d3_tables(), m["d3_params"], 8)
    assert np.abs(c - m["d3_coeffs"]).max() <= 1e-14 * np.abs(m["d3_coeffs"]).max()


def test_splines_golden_and_cubic_reproduction():
    m, _ = _models()
    assert np.array_equal(O.spline_uniform(1.0, 0.2, m["spline_ydys_u"], m["spline_xs_u"]), m["spline_u"])
    assert np.array_equal(O.spline_nonuniform(m["spline_kx"], m["spline_ydys"], m["spline_xs_n"]), m["spline_n"])
    # a cubic Hermite interpolant reproduces any cubic from its values and slopes
    p = np.poly1d([0.3, -1.1, 0.7, 2.0])
    kx = np.array([0.2, 0.5, 1.1, 1.4, 2.3, 3.0])
    ydys = np.column_stack([p(kx), p.deriv()(kx)])
    xs = np.linspace(0.25, 2.95, 37)
    assert np.abs(O.spline_nonuniform(kx, ydys, xs) - p(xs)).max() < 1e-12
    ku = 0.4 + 0.3 * np.arange(9)
    assert np.abs(O.spline_uniform(0.4, 0.3, np.column_stack([p(ku), p.deriv()(ku)]), xs + 0.3) - p(xs + 0.3)).max() < 1e-12


def test_spline_tip_relax_and_stiffness_golden():
    m, _ = _models()
    r = np.load(os.path.join(GOLDEN, "relax.npz"))
    sp = (m["tipspline_xs"], m["tipspline_ydys"])
    rs, fs = O.relax_strokes_spline(r["n"], r["cell"], r["FF"], _golden_tip(r), r["tips"], sp)
    assert np.array_equal(rs, m["spline_rs"]) and np.array_equal(fs, m["spline_fs"])
    ev, vecs, nnan = O.stiffness(r["n"], r["cell"], r["FF"], _golden_tip(r), r["tips"].reshape(-1, 3), rs.reshape(-1, 3),
                                 which=3, spline=sp)
    assert np.array_equal(ev, m["spline_evals"])
    for k in range(3):
        assert np.array_equal(vecs[k], m[f"spline_evec{k + 1}"])


def test_linear_spline_tip_equals_harmonic_spring():
    """A spline of k (x - L) (values and slopes) is the harmonic radial spring:
    the relaxation agrees with the spring's to rounding."""
    r = np.load(os.path.join(GOLDEN, "relax.npz"))
    tip = _golden_tip(r)
    xs = np.linspace(0.3, 3.0, 9)
    kR, L = float(r["tip_kRadial"]), float(r["tip_lRadial"])
    sp = (xs, np.column_stack([kR * (xs - L), np.full_like(xs, kR)]))
    rs, fs = O.relax_strokes_spline(r["n"], r["cell"], r["FF"], tip, r["tips"], sp)
    assert np.abs(rs - r["rs_ps1"]).max() < 1e-6


@pytest.mark.parametrize("name,coef", [("morse", (3.1, 0.006)), ("vdw_r2", (3.1, 0.006)), ("vdw_r4", (3.1, 0.006)),
                                       ("vdw_invr4", (3.1, 0.006)), ("vdw_invr8", (3.1, 0.006)), ("d3", (20.0, 300.0, 900.0, 2.0e4)),
                                       ("lj_re", (3.1, 0.006)), ("inv_r6", (12.0,))])
def test_radial_force_is_minus_energy_gradient(name, coef):
    kind = {v: k for k, v in O.RADIAL_KIND.items()}.get(name)
    rs = np.linspace(1.2, 7.0, 300)
    h = 1e-5
    if kind is None:  # d3 / morse: through the grid builder on a line of points
        par = -1.6 if name == "morse" else 0.0
        E = lambda x: _line_energy(name, x, coef, par)  # noqa: E731
        F = _line_force(name, rs, coef, par)
    else:
        E = lambda x: O.radial(kind, x, coef)[0]  # noqa: E731
        F = O.radial(kind, rs, coef)[1]
    g = -(E(rs + h) - E(rs - h)) / (2 * h)
    assert np.abs(g - F).max() <= 1e-6 * max(np.abs(F).max(), 1e-8)


def _line(name, xs, coef, par):
    # one atom at the origin of a 1-point-wide grid whose points lie at x
    out = []
    for x in np.atleast_1d(xs):
        F = np.zeros((1, 1, 1, 3))
        E = np.zeros((1, 1, 1))
        O.model_ff(name, (1, 1, 1), np.eye(3) * 50, np.array([[-x, 0.0, 0.0]]), np.array([coef]), par, F=F, E=E)
        out.append((E[0, 0, 0], F[0, 0, 0, 0]))
    return np.array(out)


def _line_energy(name, xs, coef, par):
    return _line(name, xs, coef, par)[:, 0]


def _line_force(name, xs, coef, par):
    return _line(name, xs, coef, par)[:, 1]


def test_morse_minimum_and_lj_re_identity():
    # Morse: minimum -E0 at r = R0 (the 1e-8 softening shifts it by ~1e-9)
    r = np.array([3.1])
    e = _line_energy("morse", r, (3.1, 0.006), -1.6)
    assert abs(e[0] + 0.006) < 1e-12
    # LJ in (R, E) equals LJ with c6 = 2 E R^6, c12 = E R^12
    rs = np.linspace(2.0, 6.0, 50)
    R, E0 = 3.1, 0.006
    e1, f1 = O.radial(-3, rs, (R, E0))
    e2, f2 = O.radial(-2, rs, (2 * E0 * R ** 6, E0 * R ** 12))
    assert np.abs(e1 - e2).max() < 1e-13 and np.abs(f1 - f2).max() < 1e-13


@pytest.mark.skipif(not O.have_reference(), reason="oracle/_ref not built (reference sources absent)")
def test_model_grids_live_reference_random_case():
    rng = np.random.default_rng(5)
    n = np.array([7, 8, 9], dtype=np.int32)
    cell = np.array([[1.0, 0.1, 0.0], [0.0, 1.1, 0.1], [0.05, 0.0, 1.2]])
    Rs = rng.uniform(-0.5, 1.5, (6, 3))
    REs = np.column_stack([rng.uniform(2.5, 3.5, 6), rng.uniform(0.01, 0.1, 6)])
    ref = O.Reference()
    for name, ref_name, extra, par in (("vdw_r4", "getVdWFF_RE", (2, 0.8), 0.8), ("morse", "getMorseFF", (-1.3,), -1.3)):
        F, E = np.zeros(tuple(n) + (3,)), np.zeros(tuple(n))
        ref.set_grid(n, cell, F, E)
        ref.grid_model(ref_name, Rs, REs, *extra)
        F2, E2 = np.zeros_like(F), np.zeros_like(E)
        O.model_ff(name, n, cell, Rs, REs, par, F=F2, E=E2)
        assert np.array_equal(F, F2) and np.array_equal(E, E2), name
