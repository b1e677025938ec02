"""GPU parity of the batched ML generators (ppafm/ml/Generator.py): every
image of a batch equals the image the reference's per-call loop produces
(Generator.py:96-113 and 474-498: one AFMulator call per sample and tip,
re-run here through the same AFMulator one call at a time), the callbacks
fire in the reference's order, and a plan is reused across batches when
only the molecules change."""

import numpy as np
import pytest

import oracle as O
from ppafm_amd import common as C
from ppafm_amd import molgen as M
from ppafm_amd.ml.Generator import GeneratorAFMtrainer, InverseAFMtrainer
from ppafm_amd.ocl.AFMulator import AFMulator

pytestmark = pytest.mark.gpu

SYM = {1: "H", 6: "C", 7: "N", 8: "O"}


def _afm():
    return AFMulator(pixPerAngstrome=10, scan_dim=(32, 32, 14), scan_window=((2.0, 2.0, 7.0), (6.0, 6.0, 8.4)), df_steps=6)


def _write_mols(tmp_path, n):
    paths = []
    for k in range(n):
        xyz, Zs, qs = M.molecule(8 + 3 * k, 100 + k)
        p = tmp_path / f"m{k}.xyz"
        with open(p, "w") as f:
            f.write(f"{len(Zs)}\ncomment\n")
            for z, r, q in zip(Zs, xyz, qs):
                f.write(f"{SYM[int(z)]} {r[0]:.6f} {r[1]:.6f} {r[2]:.6f} {q:.6f}\n")
        paths.append(str(p))
    return paths


class Recording(InverseAFMtrainer):
    def on_batch_start(self):
        self.log.append("batch")

    def on_sample_start(self):
        self.log.append(("sample", self.counter))

    def on_afm_start(self):
        self.log.append(("afm", self.counter, self.afmulator.iZPP))

    def on_afm_end(self):
        self.log.append(("end", float(np.abs(self.Xs).sum())))
        self.Xs += 1.0  # in-place edits of the image are kept, as in the reference


def test_inverse_trainer_matches_per_call_loop(tmp_path):
    print("case: 7 molecules (8-26 atoms), batches of 4, tips O and Xe-like (iZPP 8, 54), 32x32x14 scan", flush=True)
    paths = _write_mols(tmp_path, 7)
    iZPPs, Qs, QZs = [8, 54], [[-10, 20, -10, 0], [-0.1, 0, 0, 0]], [[0.1, 0, -0.1, 0], [0, 0, 0, 0]]
    gen = Recording(_afm(), [], paths, batch_size=4, iZPPs=iZPPs, Qs=Qs, QZs=QZs)
    gen.log = []
    batches = list(gen)
    assert [len(b[2]) for b in batches] == [4, 3]
    # the same steps, one AFMulator call each
    afm = _afm()
    ref_gen = InverseAFMtrainer(afm, [], paths, batch_size=4, iZPPs=iZPPs, Qs=Qs, QZs=QZs)
    maxd = 0.0
    for bi, (Xs, Ys, mols) in enumerate(batches):
        for s in range(len(mols)):
            mol = ref_gen.molecules[bi * 4 + s]
            xyzs, qs, Zs = mol[:, :3].copy(), mol[:, 3].copy(), mol[:, 4].astype(np.int32)
            ref_gen.xyzs, ref_gen.qs, ref_gen.Zs = xyzs, qs, Zs
            ref_gen.handle_positions()
            for i, (iZPP, Q, Qz) in enumerate(zip(iZPPs, Qs, QZs)):
                afm.iZPP = iZPP
                afm.setQs(Q, Qz)
                ref_gen.REAs = C.getAtomsREA(iZPP, Zs, afm.typeParams, alphaFac=-1.0)
                ref_gen.handle_distance()
                X = afm(ref_gen.xyzs, Zs, qs, REAs=ref_gen.REAs) + 1.0
                maxd = max(maxd, float(np.abs(Xs[i][s] - X).max()))
                assert np.array_equal(Xs[i][s], X), (bi, s, i)
    print(f"max |dX| batched vs per-call: {maxd:.3g}")
    # callback order: per sample, afm starts for each tip; afm ends after the batch
    first = gen.log[: gen.log.index("batch", 1)]
    assert first[0] == "batch"
    assert [e[0] for e in first[1:13]] == ["sample", "afm", "afm"] * 4
    assert [e[0] for e in first[13:]] == ["end"] * 8


class Capture(InverseAFMtrainer):
    def on_afm_start(self):
        a = self.afmulator
        self.seen.append(dict(xyzs=self.xyzs.copy(), qs=self.qs.copy(), REAs=self.REAs.copy(), Qs=list(a.Qs),
                              QZs=list(a.QZs)))


def _oracle_image(afm, st):
    """The AFMulator image of one generator step by the FP64 oracle pipeline:
    evalLJC_QZs_noPos grid, relaxStrokesTilted, convolveZ (AFMulator.py:142-179
    with the inputs in FP32, as the reference hands them to OpenCL)."""
    cLJs = C.REA2LJ(st["REAs"]).astype(np.float32).astype(np.float64)
    x32 = st["xyzs"].astype(np.float32).astype(np.float64)
    q32 = np.asarray(st["qs"], np.float32).astype(np.float64)
    lvec = afm.lvec.astype(np.float32).astype(np.float64)
    atoms = np.concatenate([x32, q32[:, None]], axis=1)
    FE = O.ljc_qzs_ff(afm.nDim, lvec, atoms, cLJs, np.float32(st["Qs"]).astype(np.float64),
                      np.float32(st["QZs"]).astype(np.float64))
    nx, ny, nz = afm.scan_dim
    dinv = np.linalg.inv(lvec[1:]).T * np.array(afm.nDim[:3])[:, None]
    poss = np.zeros((nx, ny, 4))
    poss[..., 0] = np.linspace(afm.scan_window[0][0], afm.scan_window[1][0], nx)[:, None] - afm.lvec[0][0]
    poss[..., 1] = np.linspace(afm.scan_window[0][1], afm.scan_window[1][1], ny)[None, :] - afm.lvec[0][1]
    poss[..., 2] = afm.scan_window[1][2] - afm.lvec[0][2]
    tipRot = np.zeros((3, 4))
    tipRot[:, :3] = np.eye(3)
    tipRot[2, 3] = -np.float32(afm.dz)
    t = afm.tipR0
    dpos0 = np.array([t[0], t[1], -np.sqrt(t[2] ** 2 - t[0] ** 2 - t[1] ** 2), t[2]], dtype=np.float32).astype(np.float64)
    FEs, _, _ = O.relax_tilted(FE, dinv, poss.reshape(-1, 4), nz, tipRot, afm.stiffness.astype(np.float64), dpos0,
                               afm.relax_params[:2].astype(np.float64), max_iters=128, conv_f2=float(np.float32(1e-8)),
                               f2safe=float(np.float32(1e-8)))
    return O.convolve_z(FEs[..., 2].reshape(nx, ny, nz), afm.dfWeight.astype(np.float64))


def test_inverse_trainer_images_match_oracle_pipeline(tmp_path):
    """Generated images (the batched plan of Generator.py:425-534's loop) of
    the first sample, both tips, against the FP64 oracle pipeline on the
    state the generator handed the simulator (molecule after handle_distance,
    tip charges, REAs); FP64 grid, so the bar holds at the max: 1e-3 Hz."""
    print("case: InverseAFMtrainer, 2 molecules, tips O and Xe-like, FP64 grid, 32x32x14 scan vs oracle", flush=True)
    paths = _write_mols(tmp_path, 2)
    afm = AFMulator(pixPerAngstrome=10, scan_dim=(32, 32, 14), scan_window=((2.0, 2.0, 7.0), (6.0, 6.0, 8.4)), df_steps=6,
                    grid_f64=True)
    gen = Capture(afm, [], paths, batch_size=2, iZPPs=[8, 54], Qs=[[-10, 20, -10, 0], [-0.1, 0, 0, 0]],
                  QZs=[[0.1, 0, -0.1, 0], [0, 0, 0, 0]])
    gen.seen = []
    Xs, _, mols = next(iter(gen))
    assert len(gen.seen) == 4  # 2 samples x 2 tips, sample-major
    for i in range(2):
        Xref = _oracle_image(afm, gen.seen[i])
        d = np.abs(Xs[i][0] - Xref)
        print(f"tip {i}: max |dX| {d.max():.3g} Hz (range {Xref.min():.3g} .. {Xref.max():.3g})")
        assert Xs[i][0].shape == Xref.shape
        assert d.max() < 1e-3


def _samples(n, seed=0):
    rng = np.random.default_rng(seed)
    for k in range(n):
        xyz, Zs, qs = M.molecule(10 + 2 * k, 200 + k)
        a = rng.uniform(0, 2 * np.pi)
        rot = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1.0]])
        yield dict(xyzs=xyz + rng.uniform(-3, 3, 3), Zs=Zs, qs=qs, rot=rot)


class DfSteps(GeneratorAFMtrainer):
    def on_sample_start(self):
        self.randomize_tip(0.3)


def test_generator_trainer_matches_per_call_loop():
    print("case: 6 samples (rotated, shifted), batches of 3, LJ+PC with 2 tips, random tip tilt per sample", flush=True)
    iZPPs, Qs, QZs = [8, 8], [[-10, 20, -10, 0], [-0.2, 0, 0, 0]], [[0.1, 0, -0.1, 0], [0, 0, 0, 0]]
    np.random.seed(5)
    gen = DfSteps(_afm(), [], list(_samples(6)), sim_type="LJ+PC", batch_size=3, iZPPs=iZPPs, Qs=Qs, QZs=QZs)
    out = list(gen)
    assert len(out) == 2 and out[0][0].shape[:2] == (3, 2)
    np.random.seed(5)
    afm = _afm()
    ref = DfSteps(afm, [], list(_samples(6)), sim_type="LJ+PC", batch_size=3, iZPPs=iZPPs, Qs=Qs, QZs=QZs)
    it = iter(ref)
    maxd = 0.0
    for bi, (Xs, Ys, mols, sws) in enumerate(out):
        for s in range(len(mols)):
            ref.sample_dict = ref._load_next_sample()
            d = ref.sample_dict
            c = d["xyzs"].mean(axis=0)
            ref.xyzs_rot = np.dot(d["xyzs"] - c, d["rot"].T) + c
            ref.handle_positions()
            ref.on_sample_start()
            for i, (iZPP, Q, Qz) in enumerate(zip(iZPPs, Qs, QZs)):
                afm.iZPP = iZPP
                afm.setQs(Q, Qz)
                d["REAs"] = C.getAtomsREA(iZPP, d["Zs"], afm.typeParams, alphaFac=-1.0)
                ref.handle_distance()
                afm.setScanWindow(ref.scan_window, ref.scan_dim, df_steps=ref.df_steps)
                afm.setLvec()
                X = afm(d["xyzs"], d["Zs"], d["qs"], rot=d["rot"], REAs=d["REAs"])
                assert np.allclose(sws[s][i], np.array(ref.scan_window))
                maxd = max(maxd, float(np.abs(Xs[s][i] - X).max()))
                assert np.array_equal(Xs[s][i], X), (bi, s, i)
    del it
    print(f"max |dX| batched vs per-call: {maxd:.3g}")


def test_plan_reused_across_batches(tmp_path):
    """Same window and grid, other molecules: the second batch reuses the
    first batch's plan (one cached plan per tip setting)."""
    paths = _write_mols(tmp_path, 8)
    afm = _afm()
    gen = InverseAFMtrainer(afm, [], paths, batch_size=4, iZPPs=[8], Qs=[[-10, 20, -10, 0]], QZs=[[0.1, 0, -0.1, 0]])
    b1 = next(iter(gen))
    plans = dict(afm._plans)
    b2 = next(gen)
    assert set(afm._plans) == set(plans) and all(afm._plans[k] is plans[k] for k in plans)
    X = afm(*(lambda m: (m[:, :3], m[:, 4].astype(np.int32), m[:, 3]))(gen.molecules[5]),
            REAs=C.getAtomsREA(8, gen.molecules[5][:, 4].astype(np.int32), afm.typeParams, alphaFac=-1.0))
    assert np.array_equal(b2[0][0][1], X)
    print(f"batch timings: {gen.timings}")


def _gauss_grid(xyz, w, lv, n, width, amp):
    """A periodic sum of Gaussians on the grid (lv, n)."""
    s = np.stack([lv[i + 1] / n[i] for i in range(3)])
    i, j, k = np.meshgrid(*[np.arange(m) for m in n], indexing="ij")
    P = lv[0] + i[..., None] * s[0] + j[..., None] * s[1] + k[..., None] * s[2]
    cell, inv = lv[1:], np.linalg.inv(lv[1:])
    V = np.zeros(n)
    for r, q in zip(xyz, w):
        f = (P - r) @ inv
        d = (f - np.round(f)) @ cell
        V += amp * q * np.exp(-(d**2).sum(-1) / (2 * width**2))
    return V


def _density_samples(n, fdbm):
    from ppafm_amd.ocl import field as FFcl

    lv = np.array([[-2.0, -2.0, -4.0], [12.0, 0, 0], [0, 12.0, 0], [0, 0, 11.0]])
    for k in range(n):
        xyz, Zs, qs = M.molecule(10 + 2 * k, 300 + k)
        xyz = xyz - xyz.mean(axis=0) + np.array([4.0, 4.0, 1.0])
        d = dict(xyzs=xyz, Zs=Zs, qs=FFcl.HartreePotential(_gauss_grid(xyz, qs, lv, (48, 48, 44), 0.9, 2.0), lv))
        if fdbm:
            d["rho_sample"] = FFcl.ElectronDensity(_gauss_grid(xyz, np.ones(len(Zs)), lv, (48, 48, 44), 0.6, 0.5), lv)
        yield d


@pytest.mark.parametrize("sim_type", ["LJ+Hartree", "FDBM"])
def test_generator_density_modes_match_per_call(sim_type):
    from ppafm_amd.ocl import field as FFcl

    fdbm = sim_type == "FDBM"
    print(f"case: GeneratorAFMtrainer {sim_type}, 3 samples in batches of 2, synthetic potentials and densities", flush=True)
    lv_t = np.array([[0.0, 0, 0], [6.0, 0, 0], [0, 6.0, 0], [0, 0, 6.0]])
    if fdbm:
        rho = FFcl.TipDensity(_gauss_grid(np.zeros((1, 3)), [1.0], lv_t, (30, 30, 30), 0.7, 0.3), lv_t)
        delta = FFcl.TipDensity(_gauss_grid(np.array([[0, 0, 0.1], [0, 0, -0.1]]), [1.0, -1.0], lv_t, (30, 30, 30), 0.5, 0.05), lv_t)
        rhos, rho_deltas, iZPPs, Qs, QZs = [rho], [delta], [8], None, None
    else:
        rhos, rho_deltas, iZPPs = [{"dz2": -0.1}, None], None, [8, 8]
        Qs, QZs = [[0, 0, 0, 0], [-10, 20, -10, 0]], [[0, 0, 0, 0], [0.1, 0, -0.1, 0]]

    def make():
        afm = _afm()
        afm.fdbm_vdw_type, afm.B_pauli = "LJ", 1.0
        afm.npbc = (0, 0, 0)
        return afm

    gen = GeneratorAFMtrainer(make(), [], list(_density_samples(3, fdbm)), sim_type=sim_type, batch_size=2, iZPPs=iZPPs, Qs=Qs, QZs=QZs,
                              rhos=rhos, rho_deltas=rho_deltas, density_cutoff=2.0 if fdbm else None)
    if fdbm:
        gen.set_fdbm_parameters(15.0, 1.1)
    out = list(gen)
    assert [len(b[2]) for b in out] == [2, 1]
    # the same steps, one AFMulator call each
    afm = make()
    ref = GeneratorAFMtrainer(afm, [], list(_density_samples(3, fdbm)), sim_type=sim_type, batch_size=2, iZPPs=iZPPs, Qs=Qs, QZs=QZs,
                              rhos=rhos, rho_deltas=rho_deltas, density_cutoff=2.0 if fdbm else None)
    if fdbm:
        ref.set_fdbm_parameters(15.0, 1.1)
    iter(ref)
    maxd = 0.0
    for bi, (Xs, Ys, mols, sws) in enumerate(out):
        for s in range(len(mols)):
            ref.sample_dict = ref._load_next_sample()
            d = ref.sample_dict
            c = d["xyzs"].mean(axis=0)
            ref.xyzs_rot = np.dot(d["xyzs"] - c, d["rot"].T) + c
            ref.handle_positions()
            for i, iZPP in enumerate(iZPPs):
                afm.iZPP = iZPP
                afm.setQs(ref.Qs[i], ref.QZs[i])
                afm.forcefield.rho, afm.forcefield.fft_corr = ref.rhos_tip[i], ref.ffts_tip[i]
                afm.forcefield.rho_delta, afm.forcefield.fft_corr_delta = ref.rhos_tip_delta[i], ref.ffts_tip_delta[i]
                d["REAs"] = C.getAtomsREA(iZPP, d["Zs"], afm.typeParams, alphaFac=-1.0)
                ref.handle_distance()
                afm.setScanWindow(ref.scan_window, ref.scan_dim, df_steps=ref.df_steps)
                afm.setLvec()
                X = afm(**d)
                maxd = max(maxd, float(np.abs(Xs[s][i] - X).max()))
                assert np.allclose(sws[s][i], np.array(ref.scan_window))
                assert np.array_equal(Xs[s][i], X), (bi, s, i)
            assert np.allclose(mols[s][:, 3], 0)
    print(f"max |dX| generator vs per-call: {maxd:.3g}; df range {np.min(out[0][0]):.3g} .. {np.max(out[0][0]):.3g}")
    assert np.isfinite(out[0][0]).all() and np.abs(out[0][0]).max() > 0
