"""Drop-in for ``ppafm.ocl.AFMulator.AFMulator`` (ocl/AFMulator.py:23-719),
point-charge method, running on MI355X through libppafm_hip.so.

Same constructor keywords, call signature, output shape and errors.  The
force field is the OpenCL-mode model (LJ + tip charges Qs at z offsets QZs,
evalLJC_QZs_noPos, FF.cl:366-423), the relaxation follows relaxStrokesTilted
(relax.cl:391-464: dt 0.5, dtmin 0.05, 128 steps, dt/alpha kept along the
stroke) and the df stack is convolveZ with get_simple_df_weight
(relax.cl:571-608, AFMulator.py:226-229).  Interpolation is exact software
trilinear, where the reference uses the hardware image filter.

The Hartree method (``qs`` a HartreePotential: LJ plus the potential
cross-correlated with the tip density set by setRho, or with the tip charges
Qs/QZs when no density is set, field.py:1446-1504 and 1168-1317) and the
full-density-based method (``rho_sample`` with setRhoDelta,
field.py:1506-1606) build their grids with the kernels of ppafm_amd.ocl.field
(FP64 FFT cross-correlations on hipFFT).  Tip densities are built on first
use, on the force-field grid of the time setRho / setRhoDelta was called.
The DFT-D3 vdW of the density-based method needs the D3 tables, which the
caller registers with ppafm_amd.ocl.field.set_d3_tables.
"""

import os
import warnings

import numpy as np

from .. import common
from .. import device as D
from . import field as FFcl

VALID_SIZES = np.array([16, 32, 64, 128, 192, 256, 384, 512, 768, 1024, 1536, 2048])


class _Scanner:
    """Holds the relaxation settings the reference keeps on its
    RelaxedScanner (ocl/relax.py:93-110)."""

    def __init__(self, stiffness, relax_params):
        self.stiffness = stiffness
        self.relax_params = relax_params
        self.zstep = 0.1


class _ForceField:
    """The tip-density state of ForceField_LJC (field.py:845-1000) that
    AFMulator and the ML generators read and assign: rho / fft_corr (tip
    density of the Hartree electrostatics and the Pauli term) and
    rho_delta / fft_corr_delta (the density-based electrostatics).  A density
    given to set_rho / set_rho_delta is built when first read."""

    def __init__(self):
        self._spec = {}
        self._rho = self._fft_corr = self._rho_delta = self._fft_corr_delta = None
        self.iZPP = None

    def setPP(self, Z_pp):
        self.iZPP = int(Z_pp)

    def set_rho(self, rho, lvec, n, sigma, B):
        self._rho = self._fft_corr = None
        self._spec.pop("rho", None)
        if rho is not None:
            self._spec["rho"] = (rho, np.array(lvec), tuple(int(v) for v in n[:3]), sigma, B)

    def set_rho_delta(self, rho_delta, lvec, n):
        self._rho_delta = self._fft_corr_delta = None
        self._spec.pop("rho_delta", None)
        if rho_delta is not None:
            self._spec["rho_delta"] = (rho_delta, np.array(lvec), tuple(int(v) for v in n[:3]), None, 1.0)

    def _build(self, which):
        spec = self._spec.pop(which, None)
        if spec is None:
            return
        rho, lvec, n, sigma, B = spec
        if isinstance(rho, dict):
            rho = FFcl.MultipoleTipDensity(lvec[1:], n, sigma=sigma, multipole=rho)
        if not np.allclose(B, 1.0):
            rho = rho.power_positive(p=B, in_place=False)
        # prepareBuffers (field.py:961-986): onto the force-field grid
        if not (np.allclose(rho.lvec, lvec) and np.allclose(rho.shape, n)):
            rho = rho.interp_at(lvec, n)
        fft = FFcl.FFTCrossCorrelation(rho)
        if which == "rho":
            self._rho, self._fft_corr = rho, fft
        else:
            self._rho_delta, self._fft_corr_delta = rho, fft

    def has_rho(self):
        return self._rho is not None or "rho" in self._spec

    def has_rho_delta(self):
        return self._rho_delta is not None or "rho_delta" in self._spec

    @property
    def rho(self):
        self._build("rho")
        return self._rho

    @rho.setter
    def rho(self, v):
        self._spec.pop("rho", None)
        self._rho = v

    @property
    def fft_corr(self):
        self._build("rho")
        return self._fft_corr

    @fft_corr.setter
    def fft_corr(self, v):
        self._spec.pop("rho", None)
        self._fft_corr = v

    @property
    def rho_delta(self):
        self._build("rho_delta")
        return self._rho_delta

    @rho_delta.setter
    def rho_delta(self, v):
        self._spec.pop("rho_delta", None)
        self._rho_delta = v

    @property
    def fft_corr_delta(self):
        self._build("rho_delta")
        return self._fft_corr_delta

    @fft_corr_delta.setter
    def fft_corr_delta(self, v):
        self._spec.pop("rho_delta", None)
        self._fft_corr_delta = v


class AFMulator:
    bMergeConv = False
    relaxParams = [0.5, 0.1, 0.1 * 0.2, 0.1 * 5.0]
    bSaveFF = False

    def __init__(
        self,
        pixPerAngstrome=10,
        lvec=None,
        scan_dim=(128, 128, 30),
        scan_window=((2.0, 2.0, 7.0), (18.0, 18.0, 10.0)),
        iZPP=8,
        QZs=[0.1, 0, -0.1, 0],
        Qs=[-10, 20, -10, 0],
        rho=None,
        sigma=0.71,
        rho_delta=None,
        A_pauli=18.0,
        B_pauli=1.0,
        fdbm_vdw_type="D3",
        d3_params="PBE",
        lj_vdw_damp=2,
        df_steps=10,
        tipR0=[0.0, 0.0, 3.0],
        tipStiffness=[0.25, 0.25, 0.0, 30.0],
        npbc=(1, 1, 0),
        f0Cantilever=30300,
        kCantilever=1800,
        colorscale="gray",
        minimize_memory=False,
        grid_f64=False,
    ):
        self.scanner = _Scanner(np.array(tipStiffness, dtype=np.float32) / -common.eVA_Nm, np.array(self.relaxParams, dtype=np.float32))
        self.iZPP = iZPP
        self.tipR0 = tipR0
        self.f0Cantilever = f0Cantilever
        self.kCantilever = kCantilever
        self.npbc = npbc
        self.sigma = sigma
        self.A_pauli = A_pauli
        self.B_pauli = B_pauli
        self.fdbm_vdw_type = fdbm_vdw_type
        self.d3_params = d3_params
        self.lj_vdw_damp = lj_vdw_damp
        self.sample_lvec = None
        self.colorscale = colorscale
        self.minimize_memory = minimize_memory
        # point-charge force fields: float4 grids with FP32 pairs, as the
        # reference's evalLJC_QZs_noPos (FF.cl:366-423); grid_f64=True sums
        # every pair in FP64 into a double4 grid.  The Hartree and
        # density-based force fields always use double4 grids.
        self.grid_f64 = bool(grid_f64)
        self._grid = None
        self._nDim_alloc = None
        self._out = None
        self.forcefield = _ForceField()
        self._rho = None
        self.rho_delta = None
        self._old_nDim = np.zeros(3, dtype=np.int64)
        self.setScanWindow(scan_window, scan_dim, df_steps)
        self.setLvec(lvec, pixPerAngstrome)
        self.setRho(rho, sigma, B_pauli)
        self.setRhoDelta(rho_delta)
        self.A_pauli = A_pauli
        self.setQs(Qs, QZs)
        self.typeParams = common.loadSpecies("atomtypes.ini")

    # ------------------------------------------------------------------ call
    def eval(self, xyzs, Zs, qs, rho_sample=None, sample_lvec=None, rot=np.eye(3), rot_center=None, REAs=None, X=None, plot_to_dir=None):
        self.prepareFF(xyzs, Zs, qs, rho_sample, sample_lvec, rot, rot_center, REAs)
        self.prepareScanner()
        X = self.evalAFM(X)
        if plot_to_dir:
            raise NotImplementedError("plotting is not on this path")
        return X

    def __call__(self, *args, **kwargs):
        return self.eval(*args, **kwargs)

    def eval_(self, mol):
        return self.eval(mol.xyzs, mol.Zs, mol.qs)

    # ----------------------------------------------------------------- setup
    def setLvec(self, lvec=None, pixPerAngstrome=None):
        if pixPerAngstrome is not None:
            self.pixPerAngstrome = pixPerAngstrome
        if lvec is not None:
            self.lvec = np.array(lvec, dtype=np.float64)
        else:
            self.lvec = get_lvec(self.scan_window, tipR0=self.tipR0, pixPerAngstrome=self.pixPerAngstrome)
        if hasattr(self, "nDim"):
            self._old_nDim = np.array(self.nDim[:3])
        self.nDim = common.genFFSampling(self.lvec, pixPerAngstrome=self.pixPerAngstrome)

    def _lvec_ff(self):
        """The force-field lattice as the device grid uses it (FP32 values,
        like the reference's ForceField_LJC.lvec0 / lvec)."""
        return self.lvec.astype(np.float32).astype(np.float64)

    def setScanWindow(self, scan_window=None, scan_dim=None, df_steps=None):
        if scan_window is not None:
            self.scan_window = scan_window
        if scan_dim is not None:
            self.scan_dim = tuple(scan_dim)
        if df_steps is not None:
            if df_steps <= 0 or df_steps > self.scan_dim[2]:
                raise ValueError(f"df_steps should be between 1 and scan_dim[2]({self.scan_dim[2]}), but got {df_steps}.")
            self.df_steps = df_steps
        self.dz = (self.scan_window[1][2] - self.scan_window[0][2]) / self.scan_dim[2]
        self.dfWeight = common.get_simple_df_weight(self.df_steps, dz=self.dz).astype(np.float32)
        self.dfWeight *= common.eVA_Nm * self.f0Cantilever / self.kCantilever
        self.amplitude = self.dz * len(self.dfWeight)
        self._out = None

    @property
    def stiffness(self):
        return self.scanner.stiffness

    @property
    def relax_params(self):
        return self.scanner.relax_params

    def setRho(self, rho=None, sigma=0.71, B_pauli=1.0):
        """Set the tip charge density (AFMulator.py:243-284): a multipole dict
        (MultipoleTipDensity on the force-field grid with width sigma), a
        TipDensity, or None to remove it.  With B_pauli != 1 the density is
        raised to that power (FDBM Pauli term)."""
        if rho is not None:
            if not isinstance(rho, (dict, FFcl.TipDensity)):
                raise ValueError(f"rho should of type `TipDensity`, but got `{type(rho)}`")
            self.sigma = sigma
            self.B_pauli = B_pauli
            self._rho = rho
            self.forcefield.set_rho(rho, self._lvec_ff(), self.nDim, sigma, B_pauli)
        else:
            self._rho = None
            self.forcefield.set_rho(None, None, None, None, None)

    def setBPauli(self, B_pauli=1.0):
        """Set the Pauli repulsion exponent of the FDBM (AFMulator.py:286-288)."""
        self.setRho(self._rho, sigma=self.sigma, B_pauli=B_pauli)

    def setRhoDelta(self, rho_delta=None):
        """Set the tip electron delta density of the FDBM electrostatics
        (AFMulator.py:290-308); None removes it."""
        self.rho_delta = rho_delta
        if rho_delta is not None:
            if not isinstance(rho_delta, FFcl.TipDensity):
                raise ValueError(f"rho_delta should of type `TipDensity`, but got `{type(rho_delta)}`")
            self.forcefield.set_rho_delta(rho_delta, self._lvec_ff(), self.nDim)
        else:
            self.forcefield.set_rho_delta(None, None, None)

    def setQs(self, Qs, QZs):
        if (len(Qs) != 4) or (len(QZs) != 4):
            raise ValueError("Qs and Qzs must have length 4")
        self.Qs = Qs
        self.QZs = QZs

    def setStiffness(self, tipStiffness):
        self.scanner.stiffness = np.array(tipStiffness, dtype=np.float32) / -common.eVA_Nm

    # --------------------------------------------------------------- imaging
    def _method(self, qs, rho_sample, sample_lvec):
        """The force-field method of one call and its Hartree potential
        (AFMulator.py:364-396): 'point-charge', 'hartree' or 'fdbm'."""
        if rho_sample is not None:
            if not isinstance(rho_sample, FFcl.ElectronDensity):
                raise ValueError(f"rho_sample should of type `ElectronDensity`, but got `{type(rho_sample)}`")
            if not isinstance(qs, FFcl.HartreePotential):
                raise ValueError(f"qs should be HartreePotential when rho_sample is not None, but got type `{type(qs)}`.")
            if self.rho_delta is None:
                raise ValueError("rho_delta should be set when rho_sample is not None.")
        pot = None
        if isinstance(qs, FFcl.HartreePotential):
            pot = qs
            self.sample_lvec = np.asarray(sample_lvec if sample_lvec is not None else pot.lvec[1:], dtype=np.float64)
            assert self.sample_lvec.shape == (3, 3), f"sample_lvec has shape {self.sample_lvec.shape}, but should have shape (3, 3)"
        elif isinstance(qs, FFcl.DataGrid):
            raise ValueError(f"qs should be point charges or a HartreePotential, but got `{type(qs)}`")
        elif sample_lvec is not None:
            self.sample_lvec = np.asarray(sample_lvec, dtype=np.float64)
        method = "fdbm" if rho_sample is not None else "hartree" if pot is not None else "point-charge"
        if method != "fdbm" and not np.allclose(self.B_pauli, 1.0):
            warnings.warn(f"Not using FDBM, but tip density exponent is {self.B_pauli}! This is probably not what you want to do!")
        return method, pot

    def _prepare_atoms(self, xyzs, Zs, qs, rho_sample=None, sample_lvec=None, rot=np.eye(3), rot_center=None, REAs=None, full=False):
        """Atoms of one molecule as the grid kernel takes them: (xyzq relative
        to the grid origin, cLJs, charged); with full also (REAs, Zs, rot_center)
        after the periodic images.  qs: point charges or None (the callers
        replace a Hartree potential by zeros, AFMulator.py:374-377)."""
        xyzs = np.asarray(xyzs, dtype=np.float64)
        qs = np.zeros(len(Zs)) if qs is None else np.asarray(qs, dtype=np.float64)
        if sample_lvec is not None:
            self.sample_lvec = np.asarray(sample_lvec, dtype=np.float64)
        npbc = self.npbc if self.sample_lvec is not None else (0, 0, 0)
        if rot_center is None:
            rot_center = xyzs.mean(axis=0)
        if REAs is None:
            REAs = common.getAtomsREA(self.iZPP, Zs, self.typeParams, alphaFac=-1.0)
        cLJs = common.REA2LJ(REAs)
        if sum(npbc) > 0:
            Zs, xyzs, qs, cLJs, REAs = common.PBCAtoms3D_np(Zs, xyzs, qs, cLJs, REAs, self.sample_lvec, npbc=npbc)
        # atoms are rotated about rot_center (field.py:1694-1697), kept in FP32
        # like the reference's device buffers (field.py:926-931)
        xyzs = np.asarray(xyzs, dtype=np.float64) - rot_center
        xyzs = np.dot(xyzs, np.asarray(rot).T) + rot_center
        xyzs = xyzs.astype(np.float32).astype(np.float64)
        qs = np.asarray(qs, dtype=np.float32).astype(np.float64)
        cLJs = np.asarray(cLJs, dtype=np.float32).astype(np.float64)
        lvec0 = self.lvec[0].astype(np.float32).astype(np.float64)
        xyzq = np.concatenate([xyzs - lvec0[None, :], qs[:, None]], axis=1)
        if full:
            REAs = np.asarray(REAs, dtype=np.float32).astype(np.float64)
            return xyzq, cLJs, not np.allclose(qs, 0), REAs, np.asarray(Zs), rot_center
        return xyzq, cLJs, not np.allclose(qs, 0)

    def _ff_model(self, charged):
        stencil = list(zip(np.float32(self.Qs).tolist(), np.float32(self.QZs).tolist())) if charged else []
        return D.ff_model(1, with_lj=True, stencil=stencil, with_energy=True, fmax=None)

    def prepareFF(self, xyzs, Zs, qs, rho_sample=None, sample_lvec=None, rot=np.eye(3), rot_center=None, REAs=None):
        self.check_scan_window()
        n = tuple(int(v) for v in self.nDim[:3])
        # the grid size changed: rebuild the tip densities (AFMulator.py:346-355)
        if (np.array(self._old_nDim[:3]) != np.array(n)).any():
            if self._rho is not None:
                self.setRho(self._rho, self.sigma, self.B_pauli)
                self.setRhoDelta(self.rho_delta)
            self._old_nDim = np.array(n)
        method, pot = self._method(qs, rho_sample, sample_lvec)
        if pot is not None:
            qs = None
        xyzq, cLJs, charged, REAs, Zs_pbc, rot_center = self._prepare_atoms(xyzs, Zs, qs, rho_sample, sample_lvec, rot, rot_center, REAs, full=True)
        dtype = np.float64 if (self.grid_f64 or method != "point-charge") else np.float32
        if self._grid is None or self._nDim_alloc != n or self._grid.dtype != dtype:
            if self._nDim_alloc != n or getattr(self, "_ws", None) is None:
                self._ws = FFcl.Workspace(n)
            self._grid = D.DeviceArray(n + (4,), dtype)
            self._nDim_alloc = n
        lvec_ff = self._lvec_ff()
        cell = lvec_ff[1:]
        if method == "point-charge":
            self._model = self._ff_model(charged)
            D.ff_build(n, cell, xyzq, cLJs, self._model, out=self._grid)
        elif method == "hartree":
            self._model = self._ff_model(False)
            D.ff_build(n, cell, xyzq, cLJs, self._model, out=self._grid)
            if not self.forcefield.has_rho():
                if not np.allclose(rot, np.eye(3)):
                    raise NotImplementedError("Force field calculation with rotation for Hartree potential with point charges tip density is not implemented.")
                Qs = np.float32(self.Qs).astype(np.float64)
                QZs = np.float32(self.QZs).astype(np.float64)
                FFcl.calc_force_hartree_tip_charges(self._grid, lvec_ff, n, pot, Qs, QZs, ws=self._ws)
            else:
                FFcl.calc_force_hartree(self._grid, lvec_ff, n, pot, self.forcefield.fft_corr, rot=np.linalg.inv(rot), rot_center=rot_center,
                                        ws=self._ws)
        else:
            self.forcefield.setPP(self.iZPP)
            if not self.forcefield.has_rho():
                raise ValueError("the tip density (setRho) should be set for the FDBM")
            FFcl.calc_force_fdbm(self._grid, lvec_ff, n, pot, rho_sample, self.forcefield.fft_corr, self.forcefield.fft_corr_delta,
                                 A=self.A_pauli, B=self.B_pauli, rot=np.linalg.inv(rot), rot_center=rot_center, ws=self._ws)
            xyz_rel = xyzq[:, :3]
            if self.fdbm_vdw_type == "D3":
                coeffs = FFcl.d3_coeffs(xyz_rel, Zs_pbc, self.iZPP, self.d3_params)
                FFcl.add_d3(self._grid, n, cell, xyz_rel, coeffs)
            elif self.fdbm_vdw_type == "LJ":
                coefs = cLJs if self.lj_vdw_damp <= 0 else REAs[:, :2]
                FFcl.add_vdw(self._grid, n, cell, xyz_rel, coefs, self.lj_vdw_damp)
            else:
                raise ValueError(f"Invalid vdw type `{self.fdbm_vdw_type}`")
        self._method_last = method

    def batch(self, molecules, tuning=None):
        """Evaluate many molecules in one batch plan (one captured HIP graph).

        molecules: sequence of (xyzs, Zs, qs) or dicts with those keys plus
        the optional ``rot``, ``rot_center``, ``REAs`` of ``__call__``.
        Returns the list of df images X, each as ``__call__`` returns it.
        No counterpart in the reference (one __call__ per molecule,
        AFMulator.py:142-179); the numbers equal the per-molecule calls."""
        reqs = []
        for m in molecules:
            kw = dict(m) if isinstance(m, dict) else dict(zip(("xyzs", "Zs", "qs"), m))
            reqs.append(self.request(**kw))
        return self.run_requests(reqs, tuning=tuning)

    def request(self, xyzs, Zs, qs=None, rho_sample=None, sample_lvec=None, rot=np.eye(3), rot_center=None, REAs=None,
                X=None, plot_to_dir=None):
        """Everything one ``eval`` with the current settings would compute,
        captured for a later batched evaluation (run_requests): the molecule's
        atoms in the grid frame, the grid, the scan, the tip, the relaxation
        and the df weights.  Later changes to the settings do not affect it.
        Takes the arguments of ``eval``; X (an output array) is not used."""
        if plot_to_dir:
            raise NotImplementedError("plotting is not on this path")
        if rho_sample is not None or isinstance(qs, FFcl.DataGrid):
            # Hartree / density-based grids: evaluated now, one call each
            return dict(image=self.eval(xyzs, Zs, qs, rho_sample, sample_lvec, rot, rot_center, REAs).copy())
        self.check_scan_window()
        xyzq, cLJs, charged = self._prepare_atoms(xyzs, Zs, qs, rho_sample, sample_lvec, rot, rot_center, REAs)
        self.prepareScanner()
        n = tuple(int(v) for v in self.nDim[:3])
        cell = self.lvec[1:].astype(np.float32).astype(np.float64)
        return dict(xyzq=xyzq, cLJs=cLJs, charged=charged, Qs=tuple(np.float32(self.Qs).tolist()),
                    QZs=tuple(np.float32(self.QZs).tolist()), n=n, cell=cell, scan=self._scan, tip=self._tip, rp=self._rp,
                    dfw=self.dfWeight.astype(np.float64), out_shape=(self.scan_dim[0], self.scan_dim[1],
                                                                     self.scan_dim[2] - len(self.dfWeight) + 1))

    _PLAN_CACHE = 4  # batch plans kept per AFMulator (grids and graphs stay allocated)

    def run_requests(self, reqs, tuning=None):
        """Evaluate requests from ``request`` in as few batch plans as their
        settings allow: requests with the same tip charges, tip, relaxation and
        df weights share a plan (one captured graph), whatever their grids,
        scans and molecules.  Plans are cached and reused when the grids and
        scans recur (the molecules may change).  Returns the df images in
        request order."""
        from ..batch import BatchScan

        groups = {}
        out = [None] * len(reqs)
        for i, r in enumerate(reqs):
            if "image" in r:
                out[i] = r["image"]
                continue
            key = (r["Qs"], r["QZs"], bytes(r["tip"]), bytes(r["rp"]), r["dfw"].tobytes())
            groups.setdefault(key, []).append(i)
        cache = self.__dict__.setdefault("_plans", {})
        for key, idx in groups.items():
            charged = any(reqs[i]["charged"] for i in idx)
            r0 = reqs[idx[0]]
            stencil = list(zip(r0["Qs"], r0["QZs"])) if charged else []
            model = D.ff_model(1, with_lj=True, stencil=stencil, with_energy=True, fmax=None)
            packs = [D.pack_atoms(reqs[i]["xyzq"], reqs[i]["cLJs"], model.coulomb_const) for i in idx]
            geo = tuple((reqs[i]["n"], reqs[i]["cell"].tobytes(), bytes(reqs[i]["scan"])) for i in idx)
            pkey = (key, charged, geo, repr(tuning), self.grid_f64)
            b = cache.get(pkey)
            if b is None:
                items = [dict(n=reqs[i]["n"], cell=reqs[i]["cell"], natoms=len(p), scan=reqs[i]["scan"]) for i, p in zip(idx, packs)]
                b = BatchScan(items, model, r0["rp"], r0["tip"], df_weights=r0["dfw"], df_scale=1.0, grid_f64=self.grid_f64, tuning=tuning)
                if len(cache) >= self._PLAN_CACHE:
                    cache.pop(next(iter(cache)))
                cache[pkey] = b
            b.set_atoms(packs)
            b.run()
            b.synchronize()
            for k, i in enumerate(idx):
                out[i] = b.df_raw(k)
        return out

    def prepareScanner(self):
        self.pos0 = np.array([0, 0, self.scan_window[1][2]]) - self.lvec[0]
        rot = np.eye(3)
        zstep = self.dz
        tipR0 = self.tipR0
        d0 = np.array([tipR0[0], tipR0[1], -np.sqrt(tipR0[2] ** 2 - tipR0[0] ** 2 - tipR0[1] ** 2)], dtype=np.float32)
        st = self.stiffness
        self._tip = D.tip_desc(d0.astype(np.float64), st[:3].astype(np.float64), float(st[3]), float(np.float32(tipR0[2])), rot=rot)
        nx, ny, nz = self.scan_dim
        start = self.scan_window[0][:2]
        end = self.scan_window[1][:2]
        p0 = np.array([self.pos0[0] + start[0], self.pos0[1] + start[1], self.pos0[2]])
        self._scan = D.scan_desc(
            nx, ny, nz, p0=p0,
            da=((end[0] - start[0]) / max(nx - 1, 1), 0, 0),
            db=(0, (end[1] - start[1]) / max(ny - 1, 1), 0),
            dstep=(0, 0, -np.float32(zstep)),
        )
        rp = self.relax_params
        self._rp = D.relax_params(mode=1, alg=1, max_iters=128, conv_f2=float(np.float32(1e-8)), dt=float(rp[0]), damping=float(rp[1]), dtmin_frac=float(np.float32(0.1)), f2safe=float(np.float32(1e-8)))

    def evalAFM(self, X=None):
        nx, ny, nz = self.scan_dim
        if self._out is None or self._out[0].shape != (nx, ny, nz, 4):
            self._out = (D.DeviceArray((nx, ny, nz, 4), np.float32), D.DeviceArray((nx, ny, nz, 3), np.float32))
        force, paths = self._out
        cell = self.lvec[1:].astype(np.float32).astype(np.float64)
        D.relax(tuple(int(v) for v in self.nDim[:3]), cell, self._grid, self._rp, self._tip, self._scan, pos=paths, force=force)
        self._paths = paths
        nw = len(self.dfWeight)
        nzout = nz - nw + 1
        conv = D.convolve_z_dev(force, nx * ny, nz, self.dfWeight.astype(np.float64), scale=1.0, in_stride=nz * 4, in_step=4, in_offset=2)
        Xd = conv.download().reshape(nx, ny, nzout)
        if X is None:
            X = Xd.copy()
        else:
            if X.shape != tuple(self.scan_dim[:2]) + (nzout,):
                raise ValueError(
                    f"Expected an array of shape {tuple(self.scan_dim[:2]) + (nzout,)} for storing AFM image, but got an array of shape {X.shape} instead."
                )
            X[:, :, :] = Xd
        return X

    # ------------------------------------------------------------ accessors
    def downloadFF(self):
        """Force field (nx, ny, nz, 4) float32 (Fx, Fy, Fz, E)."""
        return self._grid.download().astype(np.float32)

    def downloadPaths(self):
        """Probe positions (nx, ny, nz, 3) in the molecule frame (relax.py:282-305)."""
        return self._paths.download() + self.lvec[0].astype(np.float32)

    # ------------------------------------------------------------ params I/O
    @classmethod
    def from_params(cls, file_path="./params.ini"):
        parameters, sample_lvec = _get_params(file_path)
        afmulator = cls(**parameters)
        afmulator.sample_lvec = sample_lvec
        return afmulator

    def load_params(self, file_path="./params.ini"):
        parameters, sample_lvec = _get_params(file_path)
        if parameters["tipStiffness"] is not None:
            self.scanner.stiffness = np.array(parameters["tipStiffness"], dtype=np.float32) / -common.eVA_Nm
        self.iZPP = parameters["iZPP"]
        self.tipR0 = parameters["tipR0"]
        self.f0Cantilever = parameters["f0Cantilever"]
        self.kCantilever = parameters["kCantilever"]
        self.npbc = parameters["npbc"]
        self.A_pauli = parameters["A_pauli"]
        self.setScanWindow(parameters["scan_window"], parameters["scan_dim"], parameters["df_steps"])
        self.setLvec(parameters["lvec"], parameters["pixPerAngstrome"])
        self.setRho(parameters["rho"], parameters["sigma"], parameters["B_pauli"])
        self.sample_lvec = sample_lvec
        self.colorscale = parameters["colorscale"]

    def save_params(self, file_path="./params.ini"):
        k = self.stiffness * -common.eVA_Nm
        nDim = self.nDim
        scan_step = (
            (self.scan_window[1][0] - self.scan_window[0][0]) / (self.scan_dim[0] - 1),
            (self.scan_window[1][1] - self.scan_window[0][1]) / (self.scan_dim[1] - 1),
            (self.scan_window[1][2] - self.scan_window[0][2]) / self.scan_dim[2],
        )
        with open(file_path, "w") as f:
            f.write(f"probeType {self.iZPP}\n")
            if isinstance(self._rho, dict) and len(self._rho) > 0:
                if len(self._rho) > 1:
                    warnings.warn("More than one tip multipole type specified. Only writing the first one into params.ini.")
                tip = list(self._rho)[0]
                f.write(f"tip {tip}\n")
                f.write(f"charge {self._rho[tip]}\n")
                f.write(f"sigma {self.sigma}\n")
            f.write(f"stiffness {k[0]} {k[1]} {k[3]}\n")
            f.write(f"r0Probe {self.tipR0[0]} {self.tipR0[1]} {self.tipR0[2]}\n")
            f.write(f"PBC {(np.array(self.npbc) > 0).any()}\n")
            f.write(f"nPBC {self.npbc[0]} {self.npbc[1]} {self.npbc[2]}\n")
            if self.sample_lvec is not None:
                for name, row in zip(("gridA", "gridB", "gridC"), self.sample_lvec):
                    f.write(f"{name} {row[0]} {row[1]} {row[2]}\n")
            for name, row in zip(("FFgrid0", "FFgridA", "FFgridB", "FFgridC"), self.lvec):
                f.write(f"{name} {row[0]} {row[1]} {row[2]}\n")
            f.write(f"gridN {nDim[0]} {nDim[1]} {nDim[2]}\n")
            f.write(f"scanMin {self.scan_window[0][0]} {self.scan_window[0][1]} {self.scan_window[0][2]}\n")
            f.write(f"scanMax {self.scan_window[1][0]} {self.scan_window[1][1]} {self.scan_window[1][2]}\n")
            f.write(f"scanStep {scan_step[0]} {scan_step[1]} {scan_step[2]}\n")
            f.write(f"kCantilever {self.kCantilever}\n")
            f.write(f"f0Cantilever {self.f0Cantilever}\n")
            f.write(f"Amplitude {self.amplitude}\n")
            f.write(f"Apauli {self.A_pauli}\n")
            f.write(f"Bpauli {self.B_pauli}\n")

    def check_scan_window(self):
        for i, dim in enumerate("xyz"):
            if self.npbc[i]:
                continue
            lvec_start = self.lvec[0, i]
            lvec_end = lvec_start + self.lvec[i + 1, i]
            scan_start = self.scan_window[0][i] - 0.3
            scan_end = self.scan_window[1][i] + 0.3
            if i == 2:
                scan_start -= self.tipR0[2]
                scan_end -= self.tipR0[2]
            if (scan_start < lvec_start) or (scan_end > lvec_end):
                warnings.warn(f"The edge of the scan window in {dim} dimension is very close or extends over the boundary of the force-field grid which is not periodic in {dim} dimension.")


def _get_params(file_path):
    parameters = common.PpafmParameters.from_file(file_path)
    lvec = np.array([parameters.FFgrid0, parameters.FFgridA, parameters.FFgridB, parameters.FFgridC])
    if (lvec < 0).any():
        lvec = None
    sample_lvec = np.array([parameters.gridA, parameters.gridB, parameters.gridC])
    if (np.array(parameters.gridN) == 0).any() or lvec is None:
        pixPerAngstrome = 10
    else:
        rx, ry, rz = (round(parameters.gridN[i] / np.linalg.norm(lvec[i + 1])) for i in range(3))
        pixPerAngstrome = rx if np.allclose([rx, ry], rz) else int(np.max([rx, ry, rz]))
    scan_window = (parameters.scanMin, parameters.scanMax)
    scan_dim = (
        round((scan_window[1][0] - scan_window[0][0]) / parameters.scanStep[0]) + 1,
        round((scan_window[1][1] - scan_window[0][1]) / parameters.scanStep[1]) + 1,
        round((scan_window[1][2] - scan_window[0][2]) / parameters.scanStep[2]),
    )
    from ..io import ELEMENT_Z

    iZPP = parameters.probeType
    iZPP = ELEMENT_Z[iZPP] if iZPP in ELEMENT_Z else int(iZPP)
    tipStiffness = parameters.stiffness
    if (np.array(tipStiffness) < 0).any():
        tipStiffness = [0.25, 0.25, 0.0, 30.0]
    else:
        tipStiffness = np.insert(tipStiffness, 2, 0.0)
    afmulator_params = {
        "lvec": lvec,
        "pixPerAngstrome": pixPerAngstrome,
        "scan_dim": scan_dim,
        "scan_window": scan_window,
        "iZPP": iZPP,
        "rho": {parameters.tip: parameters.charge},
        "sigma": parameters.sigma,
        "A_pauli": parameters.Apauli,
        "B_pauli": parameters.Bpauli,
        "df_steps": round(parameters.Amplitude / parameters.scanStep[2]),
        "tipR0": parameters.r0Probe,
        "tipStiffness": tipStiffness,
        "npbc": tuple(int(v) * int(parameters.PBC) for v in parameters.nPBC),
        "f0Cantilever": parameters.f0Cantilever,
        "kCantilever": parameters.kCantilever,
        "colorscale": parameters.colorscale,
    }
    return afmulator_params, sample_lvec


def get_lvec(scan_window, pad=(3.0, 3.0, 5.0), tipR0=(0.0, 0.0, 3.0), pixPerAngstrome=10):
    pad = np.array(pad)
    tipR0 = np.array(tipR0)
    center = (np.array(scan_window[0]) + np.array(scan_window[1])) / 2
    box_size = (np.array(scan_window[1]) - np.array(scan_window[0])) + 2 * pad
    nDim = (pixPerAngstrome * box_size).round().astype(np.int32)
    nDim = np.array([VALID_SIZES[VALID_SIZES >= d][0] for d in nDim])
    box_size = nDim / pixPerAngstrome
    origin = center - box_size / 2 - tipR0
    return np.array([origin, [box_size[0], 0, 0], [0, box_size[1], 0], [0, 0, box_size[2]]])
