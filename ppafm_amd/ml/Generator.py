"""Drop-in for ``ppafm.ml.Generator`` (ml/Generator.py:14-713): the training
data generators, with each batch's AFM images computed in batched GPU calls.

The reference evaluates one ``AFMulator`` call per molecule and tip inside
its sample loop (Generator.py:96-113, 474-498), each a blocking grid build,
relaxation and download.  Here the loop runs the same callbacks in the same
order and, where the reference evaluates, records a request
(``AFMulator.request``: the molecule and every setting at that moment);
after the loop all requests of the batch are evaluated by
``AFMulator.run_requests`` -- one captured HIP graph per group of requests
that share tip charges, tip, relaxation and df weights, reused from batch to
batch while the grids and scans recur.  The images equal the per-call ones.

Differences that follow from deferring the evaluation:
  * ``InverseAFMtrainer.on_afm_end`` runs after the batch is evaluated, once
    per sample and tip in the reference's order, with ``self.xyzs``,
    ``self.qs``, ``self.Zs``, ``self.REAs`` and ``self.Xs`` set to that
    sample's values (changes to ``self.Xs`` in place are kept, as there);
    settings it changes take effect from the next batch, not the next
    sample.
  * The Hartree and density-based types ("lj+hartree", "fdbm") run one
    AFMulator call per sample and tip when the request is recorded, on the
    GPU grids of ppafm_amd.ocl.field; each tip density keeps its own FFT
    correlator.
AuxMaps are not part of this package: ``aux_maps`` takes any callables with
the reference's AuxMap call signatures.
"""

import random
import time

import numpy as np

from .. import common as PPU
from .. import io
from ..ocl import field as FFcl


def _xyzq(xyzs, qs):
    return np.concatenate([xyzs, np.asarray(qs)[:, None]], axis=1)


class InverseAFMtrainer:
    """Batches (Xs, Ys, mols) from xyz files (Generator.py:14-278): Xs one
    (n_batch, sx, sy, sz) array per tip, Ys one (n_batch, sx, sy) array per
    aux map, mols (n_atoms, 5) arrays [x, y, z, q, Z]."""

    def __init__(self, afmulator, aux_maps, paths, batch_size=30, distAbove=5.3, iZPPs=[8], Qs=[[-10, 20, -10, 0]],
                 QZs=[[0.1, 0, -0.1, 0]]):
        assert len(iZPPs) == len(Qs) and len(Qs) == len(QZs)
        self.afmulator = afmulator
        self.aux_maps = aux_maps
        self.paths = paths
        self.batch_size = batch_size
        self.distAbove = distAbove
        self.distAboveActive = distAbove
        self.iZPPs = iZPPs
        self.Qs = Qs
        self.QZs = QZs
        self.read_xyzs()
        self.counter = 0
        self.timings = {}

    def __iter__(self):
        self.counter = 0
        return self

    def __len__(self):
        return int(np.floor(len(self.molecules) / self.batch_size))

    def __next__(self):
        if self.counter >= len(self.molecules):
            raise StopIteration
        self.on_batch_start()
        t0 = time.perf_counter()
        mols = []
        Ys = [[] for _ in self.aux_maps]
        batch_size = min(self.batch_size, len(self.molecules) - self.counter)
        pending = []  # (tip index, request, sample state)
        for _ in range(batch_size):
            mol = self.molecules[self.counter]
            mols.append(mol)
            self.xyzs = mol[:, :3]
            self.qs = mol[:, 3]
            self.Zs = mol[:, 4].astype(np.int32)
            self.handle_positions()
            self.on_sample_start()
            for i, (iZPP, Q, Qz) in enumerate(zip(self.iZPPs, self.Qs, self.QZs)):
                self.afmulator.iZPP = iZPP
                self.afmulator.setQs(Q, Qz)
                self.REAs = PPU.getAtomsREA(self.afmulator.iZPP, self.Zs, self.afmulator.typeParams, alphaFac=-1.0)
                self.handle_distance()
                self.on_afm_start()
                req = self.afmulator.request(self.xyzs, self.Zs, self.qs, REAs=self.REAs)
                state = (self.xyzs.copy(), self.qs.copy(), self.Zs.copy(), self.REAs.copy())
                pending.append((i, req, state))
            for j, aux_map in enumerate(self.aux_maps):
                Ys[j].append(aux_map(_xyzq(self.xyzs, self.qs), self.Zs))
            self.counter += 1
        t1 = time.perf_counter()
        images = self.afmulator.run_requests([p[1] for p in pending])
        t2 = time.perf_counter()
        Xs = [[] for _ in self.iZPPs]
        for (i, _, state), X in zip(pending, images):
            self.xyzs, self.qs, self.Zs, self.REAs = state
            Xs[i].append(X)
            self.Xs = X
            self.on_afm_end()
        Xs = [np.stack(x, axis=0) for x in Xs]
        Ys = [np.stack(y, axis=0) for y in Ys]
        self.timings = dict(prepare_s=t1 - t0, afm_s=t2 - t1, total_s=time.perf_counter() - t0)
        return Xs, Ys, mols

    def read_xyzs(self):
        self.molecules = []
        for path in self.paths:
            xyzs, Zs, qs, _ = io.loadXYZ(path)
            self.molecules.append(np.concatenate([xyzs, qs[:, None], Zs[:, None]], axis=1))

    def handle_positions(self):
        """Centre the molecule in the scan window laterally (Generator.py:165-171)."""
        sw = self.afmulator.scan_window
        centre = np.array([sw[1][0] + sw[0][0], sw[1][1] + sw[0][1]]) / 2
        self.xyzs[:, :2] += centre - self.xyzs[:, :2].mean(axis=0)

    def handle_distance(self):
        """Put the molecule's top, widened by vdW radii, distAboveActive below
        the scan window top (Generator.py:173-184)."""
        r_pp = self.afmulator.typeParams[self.afmulator.iZPP - 1][0]
        r_vdw = self.REAs[:, 0] - r_pp
        zs = self.xyzs[:, 2]
        imax = np.argmax(zs + r_vdw)
        total = self.distAboveActive + r_vdw[imax] + r_pp - (zs.max() - zs[imax])
        self.xyzs[:, 2] += (self.afmulator.scan_window[1][2] - total) - zs.max()

    def shuffle_molecules(self):
        random.shuffle(self.molecules)

    def augment_with_rotations(self, rotations):
        """One molecule per rotation matrix (Generator.py:192-206)."""
        mols, self.molecules = self.molecules, []
        for mol in mols:
            for xyzs in rotate(mol[:, :3], rotations):
                self.molecules.append(np.concatenate([xyzs, mol[:, 3:4], mol[:, 4:5]], axis=1))

    def augment_with_rotations_entropy(self, rotations, n_best_rotations=30):
        """The n_best_rotations rotations with the flattest tops (Generator.py:208-224)."""
        mols, self.molecules = self.molecules, []
        for mol in mols:
            rots = sortRotationsByEntropy(mol[:, :3], rotations)[:n_best_rotations]
            for xyzs in rotate(mol[:, :3], rots):
                self.molecules.append(np.concatenate([xyzs, mol[:, 3:4], mol[:, 4:5]], axis=1))

    def randomize_tip(self, max_tilt=0.5):
        self.afmulator.tipR0[:2] = np.array(getRandomUniformDisk()) * max_tilt

    def randomize_distance(self, delta=0.25):
        self.distAboveActive = np.random.uniform(self.distAbove - delta, self.distAbove + delta)

    def randomize_mol_parameters(self, rndQmax=0.0, rndRmax=0.0, rndEmax=0.0, rndAlphaMax=0.0):
        """Random perturbations of the current molecule's charges and REA
        parameters (Generator.py:244-258)."""
        n = len(self.qs)
        if rndQmax > 0:
            self.qs[:] += rndQmax * (np.random.rand(n) - 0.5)
        if rndRmax > 0:
            self.REAs[:, 0] += rndRmax * (np.random.rand(n) - 0.5)
        if rndEmax > 0:
            self.REAs[:, 1] *= 1 + rndEmax * (np.random.rand(n) - 0.5)
        if rndAlphaMax > 0:
            self.REAs[:, 2] *= 1 + rndAlphaMax * (np.random.rand(n) - 0.5)

    def on_batch_start(self):
        """Called at the start of each batch."""

    def on_sample_start(self):
        """Called before the first AFM image of each sample."""

    def on_afm_start(self):
        """Called before each AFM image is requested."""

    def on_afm_end(self):
        """Called for each AFM image once the batch is evaluated."""


class GeneratorAFMtrainer:
    """Batches (Xs, Ys, mols, sws) from a generator of sample dicts
    (Generator.py:281-684): Xs (batch, n_tip, nx, ny, nz), Ys (batch,
    n_auxmap, nx, ny), mols (n_atoms, 5) arrays, sws (batch, n_tip, 2, 3)
    scan windows.  Callbacks: on_batch_start, then per sample
    on_sample_start and per tip on_afm_start, in the reference's order."""

    def __init__(self, afmulator, aux_maps, sample_generator, sim_type="LJ", batch_size=30, distAbove=5.3, iZPPs=[8],
                 Qs=None, QZs=None, rhos=None, rho_deltas=None, ignore_elements=[], density_cutoff=None):
        self.afmulator = afmulator
        self.aux_maps = aux_maps
        self.sample_generator = sample_generator
        self.sim_type = sim_type.lower()
        if self.sim_type not in ("lj", "lj+pc", "lj+hartree", "fdbm"):
            raise ValueError(f"Unknown simulation type `{self.sim_type}`")
        self.batch_size = batch_size
        self.distAbove = distAbove
        self.distAboveActive = distAbove
        self.iZPPs = iZPPs
        self.ignore_elements = ignore_elements
        self.density_cutoff = density_cutoff if self.sim_type == "fdbm" else None
        self._prepare_tip_buffers(rhos, rho_deltas, self.density_cutoff)
        if Qs is None or QZs is None:
            if self.sim_type == "lj+pc":
                raise ValueError("Both Qs and QZs should be specified when using point-charge electrostatics.")
            self.Qs = self.QZs = [[0, 0, 0, 0] for _ in range(len(iZPPs))]
        else:
            assert len(Qs) == len(QZs) == len(iZPPs), "Inconsistent lengths for tip charge arrays."
            self.Qs, self.QZs = Qs, QZs
        sw = self.afmulator.scan_window
        self.scan_window = sw
        self.scan_size = (sw[1][0] - sw[0][0], sw[1][1] - sw[0][1], sw[1][2] - sw[0][2])
        self.scan_dim = self.afmulator.scan_dim
        self.df_steps = self.afmulator.df_steps
        self.z_size = self.scan_dim[2] - self.df_steps + 1
        self.timings = {}

    def _prepare_tip_buffers(self, rhos=None, rho_deltas=None, density_cutoff=None):
        """One tip density (and FDBM delta density) with its FFT correlator
        per tip (Generator.py:389-418).  Each tip keeps its own correlator
        (the reference's shares one when the grids have the same shape)."""
        if rhos is None:
            self.rhos_tip = self.rhos_tip_delta = self.ffts_tip = self.ffts_tip_delta = [None] * len(self.iZPPs)
            return
        if len(rhos) != len(self.iZPPs):
            raise ValueError(f"The length of rhos ({len(rhos)}) does not match the length of iZPPs ({len(self.iZPPs)})")
        if rho_deltas is not None and len(rhos) != len(rho_deltas):
            raise ValueError(f"The length of rhos ({len(rhos)}) does not match the length of rho_deltas ({len(rho_deltas)})")
        self.rhos_tip, self.ffts_tip, self._rhos_original = [], [], []
        for rho in rhos:
            if density_cutoff is not None:
                rho = rho.clamp(maximum=density_cutoff, in_place=False)
            self.afmulator.setRho(rho, sigma=self.afmulator.sigma, B_pauli=self.afmulator.B_pauli)
            self.rhos_tip.append(self.afmulator.forcefield.rho)
            self.ffts_tip.append(self.afmulator.forcefield.fft_corr)
            self._rhos_original.append(rho)
        if rho_deltas is not None:
            self.rhos_tip_delta, self.ffts_tip_delta = [], []
            for rho_delta in rho_deltas:
                self.afmulator.setRhoDelta(rho_delta)
                self.rhos_tip_delta.append(self.afmulator.forcefield.rho_delta)
                self.ffts_tip_delta.append(self.afmulator.forcefield.fft_corr_delta)
        else:
            self.rhos_tip_delta = self.ffts_tip_delta = [None] * len(self.iZPPs)

    def __iter__(self):
        self.sample_dict = {}
        self.sample_iterator = iter(self.sample_generator)
        self.iteration_done = False
        return self

    def __len__(self):
        if not hasattr(self.sample_generator, "__len__"):
            raise RuntimeError("Cannot infer the number of batches because sample generator does not have length attribute.")
        return int(np.floor(len(self.sample_generator) / self.batch_size))

    def __next__(self):
        """Point-charge samples are recorded and evaluated after the loop in
        batch plans; Hartree and density-based samples are evaluated when
        recorded (one AFMulator call each, as in the reference)."""
        if self.iteration_done:
            raise StopIteration
        self.on_batch_start()
        t0 = time.perf_counter()
        mols, Ys, sws, reqs = [], [], [], []
        for _ in range(self.batch_size):
            try:
                self.sample_dict = self._load_next_sample()
            except StopIteration:
                self.sample_dict = None
                self.iteration_done = True
                break
            rot = self.sample_dict["rot"]
            xyzs = self.sample_dict["xyzs"]
            Zs = self.sample_dict["Zs"]
            pot = self.sample_dict["qs"] if isinstance(self.sample_dict["qs"], FFcl.HartreePotential) else None
            qs = np.zeros(len(Zs)) if pot is not None else self.sample_dict["qs"]
            centre = xyzs.mean(axis=0)
            self.xyzs_rot = np.dot(xyzs - centre, rot.T) + centre
            mols.append(np.concatenate([self.xyzs_rot, qs[:, None], Zs[:, None]], axis=1))
            self.handle_positions()
            self.on_sample_start()
            sws_, reqs_ = [], []
            for iZPP, rho_tip, rho_tip_delta, fft_tip, fft_tip_delta, Q, Qz in zip(
                self.iZPPs, self.rhos_tip, self.rhos_tip_delta, self.ffts_tip, self.ffts_tip_delta, self.Qs, self.QZs
            ):
                self.afmulator.iZPP = iZPP
                self.afmulator.setQs(Q, Qz)
                self.afmulator.forcefield.rho = rho_tip
                self.afmulator.forcefield.fft_corr = fft_tip
                self.afmulator.forcefield.rho_delta = rho_tip_delta
                self.afmulator.forcefield.fft_corr_delta = fft_tip_delta
                self.sample_dict["REAs"] = PPU.getAtomsREA(self.afmulator.iZPP, self.sample_dict["Zs"], self.afmulator.typeParams,
                                                           alphaFac=-1.0)
                self.handle_distance()
                self.afmulator.setScanWindow(self.scan_window, self.scan_dim, df_steps=self.df_steps)
                self.afmulator.setLvec()
                self.on_afm_start()
                kw = {k: v for k, v in self.sample_dict.items() if k != "rho_sample" or v is not None}
                reqs_.append(self.afmulator.request(**kw))
                sws_.append(np.array(self.scan_window))
            xyzqs = _xyzq(self.sample_dict["xyzs"], qs)
            Ys.append([aux_map(xyzqs, self.sample_dict["Zs"], pot, self.sample_dict["rot"]) for aux_map in self.aux_maps])
            sws.append(sws_)
            reqs.append(reqs_)
        if not mols:
            raise StopIteration
        t1 = time.perf_counter()
        flat = self.afmulator.run_requests([r for rr in reqs for r in rr])
        t2 = time.perf_counter()
        ntip = len(self.iZPPs)
        Xs = np.array([flat[s * ntip:(s + 1) * ntip] for s in range(len(mols))])
        self.timings = dict(prepare_s=t1 - t0, afm_s=t2 - t1, total_s=time.perf_counter() - t0)
        return Xs, np.array(Ys), mols, np.array(sws)

    def _load_next_sample(self):
        """The next sample dict, checked for the simulation type
        (Generator.py:536-565)."""
        d = next(self.sample_iterator)
        if self.sim_type == "lj":
            d["qs"] = np.zeros((len(d["Zs"]),), dtype=np.float32)
            d["rho_sample"] = None
        elif ("qs" not in d) or (d["qs"] is None):
            raise ValueError("Sample dict does not contain qs, which is required for a simulation with electrostatics")
        elif self.sim_type == "lj+pc":
            if not isinstance(d["qs"], np.ndarray):
                raise ValueError("qs in sample dict is not a numpy array, which is required for a LJ+PC simulation.")
            d["rho_sample"] = None
        elif self.sim_type == "lj+hartree":
            if not isinstance(d["qs"], FFcl.HartreePotential):
                raise ValueError("qs in sample dict is not a HartreePotential, which is required for a LJ+Hartree simulation.")
            d["rho_sample"] = None
        elif ("rho_sample" not in d) or (d["rho_sample"] is None):
            raise ValueError("Sample dict does not contain rho_sample, which is required for an FDBM simulation.")
        if "rot" not in d:
            d["rot"] = np.eye(3)
        if self.density_cutoff is not None:
            d["rho_sample"] = d["rho_sample"].clamp(maximum=self.density_cutoff, in_place=False)
        return d

    def handle_positions(self):
        """Centre the scan window laterally on the molecule (Generator.py:576-594)."""
        ss, sw = self.scan_size, self.scan_window
        xyzs = self.xyzs_rot
        if self.ignore_elements:
            keep = ~np.isin(self.sample_dict["Zs"], self.ignore_elements)
            xyzs = xyzs[keep]
        c = xyzs[:, :2].mean(axis=0)
        sw = ((c[0] - ss[0] / 2, c[1] - ss[1] / 2, sw[0][2]), (c[0] + ss[0] / 2, c[1] + ss[1] / 2, sw[1][2]))
        self.scan_window = sw
        for aux_map in self.aux_maps:
            if hasattr(aux_map, "scan_window"):
                aux_map.scan_window = ((sw[0][0], sw[0][1]), (sw[1][0], sw[1][1]))

    def handle_distance(self):
        """Scan window bottom distAboveActive above the molecule's vdW top
        (Generator.py:596-615)."""
        r_pp = self.afmulator.typeParams[self.afmulator.iZPP - 1][0]
        r_vdw = self.sample_dict["REAs"][:, 0] - r_pp
        zs = self.xyzs_rot[:, 2]
        if self.ignore_elements:
            keep = ~np.isin(self.sample_dict["Zs"], self.ignore_elements)
            zs, r_vdw = zs[keep], r_vdw[keep]
        imax = np.argmax(zs + r_vdw)
        z_min = zs.max() + self.distAboveActive + r_vdw[imax] + r_pp - (zs.max() - zs[imax])
        sw = self.scan_window
        self.scan_window = ((sw[0][0], sw[0][1], z_min), (sw[1][0], sw[1][1], z_min + self.scan_size[2]))

    def set_fdbm_parameters(self, A_pauli, B_pauli):
        """Pauli prefactor and exponent of an FDBM simulation: the tip
        densities are raised to the new power (Generator.py:617-635).  Does
        nothing for the other types."""
        if self.sim_type != "fdbm":
            return
        self.afmulator.A_pauli = A_pauli
        self.rhos_tip, self.ffts_tip = [], []
        for rho in self._rhos_original:
            self.afmulator.setRho(rho, sigma=self.afmulator.sigma, B_pauli=B_pauli)
            self.rhos_tip.append(self.afmulator.forcefield.rho)
            self.ffts_tip.append(self.afmulator.forcefield.fft_corr)

    def randomize_df_steps(self, minimum=4, maximum=20):
        """Random df steps with the output depth and z step kept (Generator.py:637-657)."""
        self.df_steps = np.random.randint(minimum, maximum + 1)
        self.scan_dim = (self.scan_dim[0], self.scan_dim[1], self.z_size + self.df_steps - 1)
        self.scan_size = (self.scan_size[0], self.scan_size[1], self.afmulator.dz * self.scan_dim[2])

    def randomize_tip(self, max_tilt=0.5):
        self.afmulator.tipR0[:2] = max_tilt * np.array(getRandomUniformDisk())

    def randomize_distance(self, delta=0.25):
        self.distAboveActive = np.random.uniform(self.distAbove - delta, self.distAbove + delta)

    def on_batch_start(self):
        """Called at the start of each batch."""

    def on_sample_start(self):
        """Called after a sample is loaded."""

    def on_afm_start(self):
        """Called before each AFM image is requested."""


def sortRotationsByEntropy(xyzs, rotations):
    """Rotations ordered by decreasing top entropy of the molecule along
    each rotation's z row (Generator.py:687-695)."""
    scored = [(PPU.maxAlongDirEntropy(xyzs, np.asarray(rot[2]).ravel())[2], k) for k, rot in enumerate(rotations)]
    scored.sort(key=lambda t: -t[0])
    return [rotations[k] for _, k in scored]


def rotate(xyzs, rotations):
    return [np.dot(xyzs, np.asarray(rot).T) for rot in rotations]


def getRandomUniformDisk():
    """A point uniform on the unit disk (Generator.py:705-713)."""
    r = np.random.rand(2)
    rad = np.sqrt(r[0])
    ang = r[1] * 2.0 * np.pi
    return rad * np.cos(ang), rad * np.sin(ang)
