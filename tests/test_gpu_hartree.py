"""GPU parity of the Hartree and full-density-based force fields
(ppafm/ocl/field.py DataGrid family, FFTCrossCorrelation, calc_force_hartree,
run_gradPotentialGrid + run_evalLJC_Hartree, calc_force_fdbm; FF.cl kernels)
against the FP64 numpy oracle (oracle/hartree_oracle.py, pinned by
tests/test_hartree_oracle_cpu.py).

The reference runs these in FP32 OpenCL with reikna FFTs and cannot run
here; the GPU path computes in FP64, so the grid pieces are compared to the
FP64 oracle at 1e-10 of each array's scale, the force-field grids at the grid
bar (1e-5 relative, 1e-3 eV/A floor), positions at 1e-3 A and df at 1e-3 Hz.
The Hartree potentials and densities are synthetic (sums of Gaussians on
periodic grids): the reference's example data needs a download.
"""

import numpy as np
import pytest

import hartree_oracle as H
import oracle as O
from ppafm_amd import common as C
from ppafm_amd import device as D
from ppafm_amd import molgen as M
from ppafm_amd.ocl import field as FFcl
from ppafm_amd.ocl.AFMulator import AFMulator

pytestmark = pytest.mark.gpu


def _lv(o, cell):
    return np.array([o, *cell], dtype=np.float64)


def _close(a, b, tol=1e-10):
    scale = max(np.abs(b).max(), 1e-30)
    err = np.abs(np.asarray(a) - b).max() / scale
    return err, err < tol


def test_interp_at_matches_oracle():
    print("case: 20x18x24 skew potential grid -> 24x22x30 grid, rotated 0.4 rad about z and tilted", flush=True)
    rng = np.random.default_rng(0)
    lv_in = _lv([-1.0, 0.5, 0.2], [[6.0, 0.3, 0.0], [0.4, 5.5, 0.0], [0.0, 0.2, 7.0]])
    buf = rng.standard_normal((20, 18, 24)).astype(np.float32)
    g = FFcl.HartreePotential(buf, lv_in)
    lv_out = _lv([0.3, -0.7, 1.1], np.diag([7.0, 6.5, 8.0]))
    a, b = 0.4, 0.1
    R = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]]) @ np.array([[1, 0, 0], [0, np.cos(b), -np.sin(b)], [0, np.sin(b), np.cos(b)]])
    rc = np.array([2.0, 1.0, 3.0])
    out = g.interp_at(lv_out, (24, 22, 30), rot=R, rot_center=rc)
    assert isinstance(out, FFcl.HartreePotential) and out.shape == (24, 22, 30)
    ref = H.interp_at(buf.astype(np.float64), lv_in, lv_out, (24, 22, 30), rot=R, rot_center=rc)
    err, ok = _close(out.array, ref)
    print(f"interp_at max rel err {err:.3g}")
    assert ok


def test_interp_tip_at_matches_oracle():
    print("case: 30^3 tip grid (6 A cell) onto 40x36x44 grid of another step", flush=True)
    rng = np.random.default_rng(1)
    lv_in = _lv([0, 0, 0], np.diag([6.0, 6.0, 6.0]))
    buf = rng.standard_normal((30, 30, 30)).astype(np.float32)
    t = FFcl.TipDensity(buf, lv_in)
    lv_out = _lv([3.0, 2.0, 1.0], [[8.0, 0, 0], [0.5, 7.2, 0], [0, 0, 8.8]])
    n = (40, 36, 44)
    out = t.interp_at(lv_out, n)
    ref = H.interp_tip_at(buf.astype(np.float64), lv_in, lv_out, n)
    # points within 1e-9 of the zero-padding boundary may round either way
    idx = [np.arange(m) for m in n]
    for a in range(3):
        idx[a] = np.where(idx[a] > n[a] // 2, idx[a] - n[a], idx[a])
    i, j, k = np.meshgrid(*idx, indexing="ij")
    s = H.steps(lv_out, n)
    p = i[..., None] * s[0] + j[..., None] * s[1] + k[..., None] * s[2]
    frac = np.abs(p @ np.linalg.inv(lv_in[1:]).T)
    keep = (np.abs(frac - 0.5) > 1e-9).all(axis=-1)
    err, ok = _close(out.array[keep], ref[keep])
    print(f"interp_tip_at max rel err {err:.3g}; {np.count_nonzero(ref)} nonzero points")
    assert ok and np.count_nonzero(ref) > 1000


@pytest.mark.parametrize("order", ["C", "F"])
def test_grad_matches_oracle(order):
    print(f"case: grad of a 16x20x12 grid, order {order}, then accumulate", flush=True)
    rng = np.random.default_rng(2)
    lv = _lv([1, 2, 3], np.diag([4.0, 5.0, 3.0]))
    buf = rng.standard_normal((16, 20, 12)).astype(np.float32)
    g = FFcl.DataGrid(buf, lv)
    scale = [-1.0, -2.0, 0.5, 3.0]
    out = g.grad(scale=scale, order=order)
    ref = H.grad(buf.astype(np.float64), np.diag(H.steps(lv, (16, 20, 12))), scale)
    got = out.array if order == "C" else out._dev.download().reshape(12, 20, 16, 4).transpose(2, 1, 0, 3)
    err, ok = _close(got, ref)
    assert ok, err
    base = rng.standard_normal((16, 20, 12, 4))
    acc = D.DeviceArray.from_host(base)
    g.grad(scale=scale, array_out=acc, order="C", accumulate=True)
    err2, ok2 = _close(acc.download(), base + ref)
    print(f"grad max rel err {err:.3g}, accumulate {err2:.3g}")
    assert ok2


def test_grad_potential_and_tip_charges():
    print("case: gradPotentialGrid 24x20x26 potential -> 18x16x20 grid, h 0.05; four tip charges", flush=True)
    rng = np.random.default_rng(3)
    lv_pot = _lv([-0.5, -0.3, 0.0], np.diag([6.0, 5.0, 6.5]))
    pot = rng.standard_normal((24, 20, 26)).astype(np.float32)
    lv = _lv([0.2, 0.1, 0.4], np.diag([4.5, 4.0, 5.0]))
    n = (18, 16, 20)
    field = D.DeviceArray(n + (4,), np.float64)
    from ppafm_amd import _lib

    gpot = FFcl.DataGrid(pot, lv_pot)  # keeps the device copy alive through the call
    _lib.check(_lib.lib().ppafm_grid_grad_potential(_lib.dptr(lv_pot.copy()), _lib.iptr(np.array(pot.shape, np.int32)),
                                                    gpot.dev.ptr, _lib.dptr(lv.copy()), _lib.iptr(np.array(n, np.int32)),
                                                    0.05, field.ptr, None), "grad_potential")
    ref = H.grad_potential(pot.astype(np.float64), lv_pot, lv, n, 0.05)
    F = field.download()
    err, ok = _close(F, ref)
    assert ok, err
    base = rng.standard_normal(n + (4,))
    grid = D.DeviceArray.from_host(base)
    Qs, QZs = np.array([-10.0, 20.0, -10.0, 0.0]), np.array([0.1, 0.0, -0.1, 0.3])
    _lib.check(_lib.lib().ppafm_grid_add_tip_charges(_lib.dptr(lv.copy()), _lib.iptr(np.array(n, np.int32)), field.ptr, _lib.dptr(Qs), _lib.dptr(QZs),
                                                     grid.ptr, None), "tip charges")
    ref2 = base + H.tip_charges(F, lv, Qs, QZs)
    err2, ok2 = _close(grid.download(), ref2)
    print(f"gradPotentialGrid max rel err {err:.3g}; tip charges {err2:.3g}")
    assert ok2


def test_power_add_mult_clamp():
    print("case: 30x30x30 grid power_positive (p 1.2), add_mult, hard and soft clamps", flush=True)
    rng = np.random.default_rng(4)
    lv = _lv([0, 0, 0], np.diag([3.0, 3.0, 3.0]))
    # densities with a positive total (the normalisation asserts that)
    a = (rng.standard_normal((30, 30, 30)) + 0.5).astype(np.float32)
    b = (rng.standard_normal((30, 30, 30)) + 0.5).astype(np.float32)
    ga, gb = FFcl.ElectronDensity(a, lv), FFcl.ElectronDensity(b, lv)
    # no positive values: the factor is total / 0 and the assertion fires
    with np.errstate(divide="ignore"), pytest.raises(AssertionError):
        FFcl.ElectronDensity(-np.abs(a), lv).power_positive(p=1.2)
    a64, b64 = a.astype(np.float64), b.astype(np.float64)
    p = ga.power_positive(p=1.2, in_place=False)
    assert isinstance(p, FFcl.ElectronDensity)
    assert _close(p.array, H.power_positive(a64, 1.2))[1]
    assert np.isclose(ga._get_normalization_factor(), a64.sum() / a64[a64 > 0].sum(), rtol=1e-12)
    m = ga.add_mult(gb, scale=-2.5, in_place=False)
    assert _close(m.array, a64 - 2.5 * b64)[1]
    h = ga.clamp(minimum=-0.5, maximum=0.7, in_place=False)
    assert np.array_equal(h.array, np.clip(a64, -0.5, 0.7))
    s = ga.clamp(maximum=0.7, clamp_type="soft", soft_clamp_width=0.2, in_place=False)
    want = (a64 - 0.7) / (1 + np.exp((a64 - 0.7) / 0.2)) + 0.7
    assert _close(s.array, want, 1e-12)[1]
    gb.power_positive(p=1.0, in_place=True)
    assert np.isclose(gb.array.sum(), b64.sum(), rtol=1e-10)


@pytest.mark.parametrize("multipole,tilt", [({"dz2": -0.1}, 0.0), ({"s": 0.3, "pz": 0.2, "dxy": 0.05, "foo": 1.0}, 0.25), (None, 0.0)])
def test_multipole_tip_density(multipole, tilt):
    print(f"case: multipole tip density {multipole} tilt {tilt} on a 40x36x48 skew grid", flush=True)
    cell = np.array([[8.0, 0, 0], [0.6, 7.5, 0], [0, 0, 9.0]])
    n = (40, 36, 48)
    t = FFcl.MultipoleTipDensity(cell, n, sigma=0.71, multipole=multipole, tilt=tilt)
    ref = H.tip_density(cell, n, sigma=0.71, multipole=multipole, tilt=tilt)
    err, ok = _close(t.array, ref, 1e-12)
    print(f"tip density max rel err {err:.3g}")
    assert ok and np.allclose(t.lvec[0], 0) and np.allclose(t.lvec[1:], cell)


@pytest.mark.parametrize("n", [(32, 24, 40), (30, 28, 26)])
def test_fft_cross_correlation(n):
    print(f"case: FFT cross-correlation on {n}", flush=True)
    rng = np.random.default_rng(5)
    lv = _lv([0, 0, 0], np.diag([5.0, 4.0, 6.0]))
    rho = FFcl.TipDensity(rng.standard_normal(n).astype(np.float32), lv)
    arr = FFcl.HartreePotential(rng.standard_normal(n).astype(np.float32), lv)
    fc = FFcl.FFTCrossCorrelation(rho)
    E = fc.correlate(arr, scale=-1.5)
    ref = H.correlate_fft(rho.array.astype(np.float64), arr.array.astype(np.float64), rho.cell_vol, scale=-1.5)
    err, ok = _close(E.array, ref, 1e-12)
    assert np.array_equal(arr.array, arr.array)  # input untouched
    E2 = fc.correlate(arr.dev, E=E, scale=2.0)
    err2, ok2 = _close(E2.array, ref * (2.0 / -1.5), 1e-12)
    print(f"correlate max rel err {err:.3g}, {err2:.3g}")
    assert ok and ok2


# ----------------------------------------------------------- AFMulator paths


def _sample_grid(xyz, charges, lv, n, width, amp):
    """A periodic sum of Gaussians on the grid (lv, n) (minimum image)."""
    P = H.points(lv, n)
    cell = lv[1:]
    inv = np.linalg.inv(cell)
    V = np.zeros(n)
    for r, q in zip(xyz, charges):
        d = P - r
        f = d @ inv
        d = (f - np.round(f)) @ cell
        V += amp * q * np.exp(-(d**2).sum(-1) / (2 * width**2))
    return V


def _afm(**kw):
    return AFMulator(scan_dim=(20, 20, 12), scan_window=((-1.5, -1.5, 7.0), (1.5, 1.5, 8.2)), df_steps=6, **kw)


def _mol():
    xyz, Zs, qs = M.molecule(14, 21)
    return xyz - xyz.mean(axis=0) + np.array([0.0, 0.0, 1.0]), Zs, qs


def _oracle_lj(afm, xyzs, Zs, pbc_lvec=None):
    lvec = afm.lvec.astype(np.float32).astype(np.float64)
    REAs = C.getAtomsREA(afm.iZPP, Zs, afm.typeParams, alphaFac=-1.0)
    cLJs = C.REA2LJ(REAs)
    qs = np.zeros(len(Zs))
    if pbc_lvec is not None:
        Zs, xyzs, qs, cLJs, REAs = C.PBCAtoms3D_np(Zs, xyzs, qs, cLJs, REAs, pbc_lvec, npbc=afm.npbc)
    x32 = np.asarray(xyzs, np.float32).astype(np.float64)
    atoms = np.concatenate([x32, np.zeros((len(x32), 1))], axis=1)
    cLJs = np.asarray(cLJs, np.float32).astype(np.float64)
    FE = O.ljc_qzs_ff(afm.nDim, lvec, atoms, cLJs, np.zeros(4), np.zeros(4))
    return FE, lvec, x32, np.asarray(REAs, np.float32).astype(np.float64), cLJs, np.asarray(Zs)


def _oracle_relax_df(afm, G, lvec):
    nx, ny, nz = afm.scan_dim
    dinv = np.linalg.inv(lvec[1:]).T * np.array(afm.nDim[:3])[:, None]
    xs = np.linspace(afm.scan_window[0][0], afm.scan_window[1][0], nx)
    ys = np.linspace(afm.scan_window[0][1], afm.scan_window[1][1], ny)
    poss = np.zeros((nx, ny, 4))
    poss[..., 0] = xs[:, None] - afm.lvec[0][0]
    poss[..., 1] = ys[None, :] - afm.lvec[0][1]
    poss[..., 2] = afm.scan_window[1][2] - afm.lvec[0][2]
    tipRot = np.zeros((3, 4))
    tipRot[:, :3] = np.eye(3)
    tipRot[2, 3] = -np.float32(afm.dz)
    t = afm.tipR0
    dpos0 = np.array([t[0], t[1], -np.sqrt(t[2] ** 2 - t[0] ** 2 - t[1] ** 2), t[2]], dtype=np.float32).astype(np.float64)
    FEs, paths, nev = O.relax_tilted(G, dinv, poss.reshape(-1, 4), nz, tipRot, afm.stiffness.astype(np.float64), dpos0,
                                     afm.relax_params[:2].astype(np.float64), max_iters=128, conv_f2=float(np.float32(1e-8)),
                                     f2safe=float(np.float32(1e-8)))
    X = O.convolve_z(FEs[..., 2].reshape(nx, ny, nz), afm.dfWeight.astype(np.float64))
    return X, paths


def _check_afm(afm, G_ref, lvec, X, label):
    G = afm._grid.download()
    d = np.linalg.norm(G[..., :3] - G_ref[..., :3], axis=-1) / np.maximum(np.linalg.norm(G_ref[..., :3], axis=-1), 1e-3)
    de = np.abs(G[..., 3] - G_ref[..., 3]) / np.maximum(np.abs(G_ref[..., 3]), 1e-3)
    Xref, paths = _oracle_relax_df(afm, G_ref, lvec)
    nx, ny, nz = afm.scan_dim
    p_gpu = afm.downloadPaths() - afm.lvec[0].astype(np.float32)
    dp = np.linalg.norm(p_gpu.reshape(-1, nz, 3) - paths[..., :3], axis=-1)
    dX = np.abs(X - Xref)
    print(f"{label}: grid rel err F {d.max():.3g} E {de.max():.3g}; |dr| max {dp.max():.3g} A; |ddf| max {dX.max():.3g} Hz "
          f"(df range {Xref.min():.3g} .. {Xref.max():.3g})")
    assert d.max() < 1e-5 and de.max() < 1e-5
    assert dp.max() < 1e-3 and dX.max() < 1e-3


def _hartree_pot(xyz, qs):
    lv_pot = _lv([-6.0, -6.0, -4.0], np.diag([12.0, 12.0, 11.0]))
    V = _sample_grid(xyz, qs, lv_pot, (60, 60, 55), 0.9, 2.0)
    return FFcl.HartreePotential(V, lv_pot), lv_pot


def test_afmulator_hartree_multipole_tip():
    print("case: AFMulator LJ+Hartree, dz2 tip density, 14-atom molecule, rotated 0.3 rad, 20x20x12 scan", flush=True)
    xyz, Zs, qs = _mol()
    pot, lv_pot = _hartree_pot(xyz, qs)
    afm = _afm(rho={"dz2": -0.1}, npbc=(0, 0, 0))
    a = 0.3
    rot = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1.0]])
    rc = xyz.mean(axis=0)
    X = afm(xyz, Zs, pot, rot=rot, rot_center=rc)
    # oracle: LJ of the rotated atoms + grad of (interp(pot) (x) tip density)
    xyz_rot = (xyz - rc) @ rot.T + rc
    FE, lvec, _, _, _, _ = _oracle_lj(afm, xyz_rot, Zs)
    n = tuple(int(v) for v in afm.nDim[:3])
    pot_i = H.interp_at(pot.array.astype(np.float64), lv_pot, lvec, n, rot=np.linalg.inv(rot), rot_center=rc)
    tip = H.tip_density(lvec[1:], n, sigma=afm.sigma, multipole={"dz2": -0.1})
    tip = H.interp_tip_at(tip, _lv([0, 0, 0], lvec[1:]), lvec, n)
    cell_vol = abs(np.linalg.det(H.steps(lvec, n)))
    E = H.correlate_fft(tip, pot_i, cell_vol)
    G_ref = FE + H.grad(E, np.diag(H.steps(lvec, n)), (-1, -1, -1, 1))
    _check_afm(afm, G_ref, lvec, X, "LJ+Hartree (dz2 tip)")


@pytest.mark.parametrize("matching", [False, True])
def test_afmulator_hartree_tip_charges(matching):
    print(f"case: AFMulator LJ+Hartree with point-charge tip (no density), potential grid matching the FF grid: {matching}", flush=True)
    xyz, Zs, qs = _mol()
    afm = _afm(npbc=(0, 0, 0))
    n = tuple(int(v) for v in afm.nDim[:3])
    lvec = afm.lvec.astype(np.float32).astype(np.float64)
    if matching:
        lv_pot = lvec.copy()
        pot = FFcl.HartreePotential(_sample_grid(xyz, qs, lv_pot, n, 0.9, 2.0), lv_pot)
    else:
        pot, lv_pot = _hartree_pot(xyz, qs)
    X = afm(xyz, Zs, pot)
    FE, lvec, _, _, _, _ = _oracle_lj(afm, xyz, Zs)
    V = pot.array.astype(np.float64)
    if matching:
        field = H.grad(V, np.diag(H.steps(lv_pot, n)), (-1, -1, -1, 1))
    else:
        field = H.grad_potential(V, lv_pot, lvec, n, FFcl.DEFAULT_FD_STEP)
    Qs = np.float32(afm.Qs).astype(np.float64)
    QZs = np.float32(afm.QZs).astype(np.float64)
    G_ref = FE + H.tip_charges(field, lvec, Qs, QZs)
    _check_afm(afm, G_ref, lvec, X, f"LJ+Hartree (tip charges, matching={matching})")
    with pytest.raises(NotImplementedError):
        afm(xyz, Zs, pot, rot=np.array([[0, -1.0, 0], [1.0, 0, 0], [0, 0, 1.0]]))


def _tables(seed=9):
    rng = np.random.default_rng(seed)
    ref_cn = rng.uniform(0, 4, (94, 5))
    ref_cn[:, 3:] = -1.0
    ref_cn[7, 2] = -1.0
    return dict(R_COV=rng.uniform(0.5, 2.0, 94), REF_CN=ref_cn, REF_C6=rng.uniform(1.0, 40.0, (94, 94, 5, 5)),
                R4R2=rng.uniform(1.0, 6.0, 94), K=np.array([16.0, 4.0 / 3.0, 4.0]))


@pytest.mark.parametrize("vdw", ["LJ", "D3"])
def test_afmulator_fdbm(vdw):
    print(f"case: AFMulator FDBM, vdW {vdw}, s tip density, B_pauli 1.2, delta density on its own 6 A grid, 14-atom molecule", flush=True)
    xyz, Zs, qs = _mol()
    pot, lv_pot = _hartree_pot(xyz, qs)
    dens = FFcl.ElectronDensity(_sample_grid(xyz, np.ones(len(Zs)), lv_pot, (60, 60, 55), 0.6, 0.5), lv_pot)
    lv_d = _lv([0, 0, 0], np.diag([6.0, 6.0, 6.0]))
    dd = _sample_grid(np.array([[0.0, 0, 0.1], [0, 0, -0.1]]), np.array([1.0, -1.0]), lv_d, (30, 30, 30), 0.5, 0.05)
    rho_delta = FFcl.TipDensity(dd, lv_d)
    params = {"s6": 1.0, "s8": 0.79, "a1": 0.43, "a2": 4.44}
    FFcl.set_d3_tables(**{k: v for k, v in _tables().items() if k != "K"}, K=_tables()["K"])
    # the Pauli term needs a density with a positive total (power_positive
    # normalises by total / positive part): an s-like tip density
    afm = _afm(rho={"s": 0.8}, rho_delta=rho_delta, B_pauli=1.2, A_pauli=12.0, fdbm_vdw_type=vdw, d3_params=params, lj_vdw_damp=2,
               npbc=(0, 0, 0))
    X = afm(xyz, Zs, pot, rho_sample=dens)
    FE, lvec, x32, REAs, cLJs, Zp = _oracle_lj(afm, xyz, Zs)
    n = tuple(int(v) for v in afm.nDim[:3])
    cell_vol = abs(np.linalg.det(H.steps(lvec, n)))
    pot_i = H.interp_at(pot.array.astype(np.float64), lv_pot, lvec, n)
    rs = H.interp_at(dens.array.astype(np.float64), lv_pot, lvec, n)
    tip = H.tip_density(lvec[1:], n, sigma=afm.sigma, multipole={"s": 0.8})
    tip = H.power_positive(tip, 1.2)
    tip = H.interp_tip_at(tip, _lv([0, 0, 0], lvec[1:]), lvec, n)
    delta = H.interp_tip_at(dd.astype(np.float32).astype(np.float64), lv_d, lvec, n)
    E = H.correlate_fft(delta, pot_i, cell_vol, scale=-1.0) + 12.0 * H.correlate_fft(tip, H.power_positive(rs, 1.2), cell_vol)
    G_ref = H.grad(E, np.diag(H.steps(lvec, n)), (-1, -1, -1, 1))
    F = np.zeros(n + (3,))
    Ev = np.zeros(n)
    rel = x32 - lvec[0]
    if vdw == "LJ":
        O.model_ff("vdw_r4", n, lvec[1:], rel, REAs[:, :2], par=float(np.float32(0.5)), F=F, E=Ev)
    else:
        t = _tables()
        coeffs = O.d3_coeffs(rel, Zp, t, (1.0, 0.79, 0.43, 4.44 * 0.5291772109217), afm.iZPP)
        O.model_ff("d3", n, lvec[1:], rel, coeffs, F=F, E=Ev)
    G_ref[..., :3] += F
    G_ref[..., 3] += Ev
    _check_afm(afm, G_ref, lvec, X, f"FDBM ({vdw})")


def test_datagrid_files_round_trip(tmp_path):
    print("case: HartreePotential xsf with atoms read by from_file; to_file of 3D and 4D grids", flush=True)
    from ppafm_amd import io

    rng = np.random.default_rng(6)
    lv = _lv([0.5, -0.5, 0.0], np.diag([5.0, 4.0, 6.0]))
    data = rng.standard_normal((10, 8, 12))
    Zs0, xyz0 = np.array([6, 8]), np.array([[1.0, 2.0, 3.0], [1.5, 2.5, 3.5]])
    path = str(tmp_path / "pot.xsf")
    io.saveXSFData(path, data, lv, head=io.primcoords2Xsf(Zs0, xyz0.T, lv))
    h, xyzs, Zs = FFcl.HartreePotential.from_file(path, scale=2.0)
    assert isinstance(h, FFcl.HartreePotential) and h.shape == (10, 8, 12)
    want = np.vectorize(lambda v: float("%10.5e" % v))(data) * 2.0
    assert np.array_equal(h.array, want.astype(np.float32)) and np.allclose(h.lvec, lv)
    assert list(Zs) == [6, 8] and np.allclose(xyzs, xyz0)
    h.to_file(str(tmp_path / "V.xsf"), clamp=1.0)
    back = io.load_scal_field(str(tmp_path / "V"))[0]
    assert np.allclose(back, np.minimum(h.array.astype(np.float64), 1.0), rtol=1e-5, atol=1e-6)
    gr = h.grad(scale=[-1.0, -1.0, -1.0, 1.0])
    gr.to_file(str(tmp_path / "E.xsf"))
    Fv = io.load_vec_field(str(tmp_path / "E"))[0]
    w = io.load_scal_field(str(tmp_path / "E_w"))[0]
    assert np.allclose(Fv, gr.array[..., :3], rtol=1e-5, atol=1e-6) and np.allclose(w, gr.array[..., 3], rtol=1e-5, atol=1e-6)


def test_tip_density_subcores():
    print("case: TipDensity.subCores, 48^3 grid of a 6 A cell, C and O inside, H near a face (periodic image)", flush=True)
    lv = _lv([0, 0, 0], np.diag([6.0, 6.0, 6.0]))
    n = (48, 48, 48)
    base = _sample_grid(np.array([[3.0, 3.0, 3.5]]), [1.0], lv, n, 1.0, 0.5)
    t = FFcl.TipDensity(base, lv)
    xyzs = np.array([[3.0, 3.0, 3.0], [3.0, 3.0, 4.1], [0.3, 3.0, 3.0]])
    Zs = np.array([6, 8, 1])
    out = t.subCores(xyzs, Zs, Rcore=0.7)
    assert isinstance(out, FFcl.TipDensity) and out.array.dtype == np.float32
    # oracle: the normalised R4 density of every periodic image within 1 A of the cell
    inds, R_img = C.findPBCAtoms3D_cutoff(xyzs.T, lv[1:], Rcut=1.0)
    val = {1: 1.0, 6: 4.0, 8: 6.0}
    amps = np.array([[-val[int(Zs[i])] / ((np.pi * 32) / 105 * 0.7**3), 0.7] for i in inds])
    E = t.array.astype(np.float64).copy()
    O.model_ff("r4", n, lv[1:], R_img.T - lv[0], amps, F=np.zeros(n + (3,)), E=E)
    err = np.abs(out.array - E.astype(np.float32)).max() / np.abs(E).max()
    dV = abs(np.linalg.det(lv[1:])) / np.prod(n)
    lost = (out.array.astype(np.float64) - t.array.astype(np.float64)).sum() * dV
    print(f"subCores max rel err vs oracle {err:.3g}; electrons removed {-lost:.6f} (valence 11); images {len(inds)}")
    assert err < 1e-6 and len(inds) == 4
    assert abs(lost + 11.0) < 11.0 * 1e-3


def test_afmulator_hartree_periodic_images_and_skew_grid():
    print("case: AFMulator LJ+Hartree with the default npbc (1,1,0): atoms replicated in the potential's cell; skew FF grid rejected", flush=True)
    xyz, Zs, qs = _mol()
    pot, lv_pot = _hartree_pot(xyz, qs)
    afm = _afm(rho={"dz2": -0.1})
    assert tuple(afm.npbc) == (1, 1, 0)
    X = afm(xyz, Zs, pot)
    assert np.allclose(afm.sample_lvec, lv_pot[1:])
    FE, lvec, _, _, _, _ = _oracle_lj(afm, xyz, Zs, pbc_lvec=lv_pot[1:])
    n = tuple(int(v) for v in afm.nDim[:3])
    pot_i = H.interp_at(pot.array.astype(np.float64), lv_pot, lvec, n)
    tip = H.interp_tip_at(H.tip_density(lvec[1:], n, sigma=afm.sigma, multipole={"dz2": -0.1}), _lv([0, 0, 0], lvec[1:]), lvec, n)
    E = H.correlate_fft(tip, pot_i, abs(np.linalg.det(H.steps(lvec, n))))
    G_ref = FE + H.grad(E, np.diag(H.steps(lvec, n)), (-1, -1, -1, 1))
    _check_afm(afm, G_ref, lvec, X, "LJ+Hartree (periodic images)")
    skew = afm.lvec.copy()
    skew[2, 0] = 0.5
    afm2 = _afm(rho={"dz2": -0.1}, lvec=skew, npbc=(0, 0, 0))
    with pytest.raises(NotImplementedError):
        afm2(xyz, Zs, pot)
