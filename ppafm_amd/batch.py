"""Batched relaxed scans: many molecules per call.

A :class:`BatchScan` owns one batch plan of libppafm_hip.so
(``ppafm_batch_*`` in include/ppafm_hip.h): every grid, relaxation output and
df stack of the batch lives back to back in device memory, and one call of
:meth:`BatchScan.run` replays a captured HIP graph that builds all grids in
one launch, relaxes every pixel column of every molecule in one persistent
launch and converts the force stacks to df.

The reference has no batched entry point (one ``AFMulator.__call__`` per
molecule, ppafm/ocl/AFMulator.py:142-179, and the ML generators loop over it,
ppafm/ml/Generator.py:474-498).  Semantics per molecule are those of
:class:`ppafm_amd.scan.PlainScan` (ProbeParticle, ``BatchScan.plain``) or of
``AFMulator`` (OpenCL, ``AFMulator.batch``); a batch gives the same numbers
as the per-molecule calls.

Autotuning: :meth:`BatchScan.tune` times the persistent relaxation's launch
shapes (pixel tile, waves per workgroup, workgroups per CU) on a sample of the
batch and keeps the fastest; :func:`tuned` caches winners per device and
problem class in a JSON file (``PPAFM_TUNING_CACHE`` or
``~/.cache/ppafm_amd/tuning.json``).
"""

import ctypes
import json
import os
import pathlib

import numpy as np

from . import _lib
from . import common as PPU
from . import device as D
from ._lib import HipError, check, lib
from .scan import plain_setup

# Default step-0 residual gate of the learned warm start (eV/A): an estimate
# where the total force is larger than this is not used; the column's step 0
# is plain FIRE from the reference's start.
RESIDUAL_MAX = 0.5

BUF_GRIDS, BUF_POS, BUF_FORCE, BUF_ITERS, BUF_DF, BUF_COUNTERS, BUF_ATOMS = range(7)
BUF_STATUS = 8


def _scan_copy(dst, src):
    dst.nx, dst.ny, dst.nz = src.nx, src.ny, src.nz
    for k in range(3):
        dst.p0[k], dst.da[k], dst.db[k], dst.dstep[k] = src.p0[k], src.da[k], src.db[k], src.dstep[k]


def device_info():
    name = ctypes.create_string_buffer(128)
    ncu = ctypes.c_int()
    check(lib().ppafm_device_info(name, 128, ctypes.byref(ncu)), "device_info")
    return name.value.decode(errors="replace"), int(ncu.value)


class BatchScan:
    """A batch plan: items share the force-field model, relaxation parameters,
    tip and df weights; each item has its own grid (n, cell), atom count and
    scan.

    items: sequence of dicts with ``n`` (3 ints), ``cell`` (3x3, rows a, b, c),
    ``natoms`` and ``scan`` (a ``ppafm_scan`` descriptor in the grid frame).
    tuning: None (defaults), a dict (tile_y, waves_per_block, blocks_per_cu)
    or "auto" (cached winner for this device and problem class, tuned on the
    first uploaded atoms and cached on a miss).
    """

    def __init__(self, items, model, rp, tip, df_weights=None, df_scale=1.0, grid_f64=True, keep_iters=False,
                 tuning=None, stream=None, tune_sample=32):
        if len(items) < 1:
            raise ValueError("a batch needs at least one item")
        self.nitems = len(items)
        self.items = list(items)
        self.model, self.rp, self.tip = model, rp, tip
        self.grid_f64 = bool(grid_f64)
        self.w = np.zeros(0) if df_weights is None else np.ascontiguousarray(df_weights, dtype=np.float64)
        if len(self.w) > 64:
            raise ValueError("at most 64 df weights")
        arr = (_lib.BatchItem * self.nitems)()
        for k, it in enumerate(self.items):
            for a in range(3):
                arr[k].n[a] = int(it["n"][a])
            c = np.asarray(it["cell"], dtype=np.float64).reshape(9)
            for a in range(9):
                arr[k].cell[a] = float(c[a])
            arr[k].natoms = int(it["natoms"])
            _scan_copy(arr[k].scan, it["scan"])
        self._arr = arr
        tun = None
        if isinstance(tuning, dict):
            tun = _lib.BatchTuning(**tuning)
        wp = self.w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if len(self.w) else None
        h = lib().ppafm_batch_create(self.nitems, arr, ctypes.byref(model), ctypes.byref(rp), ctypes.byref(tip),
                                     int(self.grid_f64), len(self.w), wp, float(df_scale), int(bool(keep_iters)),
                                     ctypes.byref(tun) if tun is not None else None, stream)
        if not h:
            raise HipError("ppafm_batch_create failed: " + lib().ppafm_last_error().decode(errors="replace"))
        self._h = h
        self._lib = lib()
        self.stream = lib().ppafm_batch_stream(h)
        n1 = self.nitems + 1
        self.grid_off = np.zeros(n1, np.int64)
        self.pos_off = np.zeros(n1, np.int64)
        self.df_off = np.zeros(n1, np.int64)
        check(lib().ppafm_batch_offsets(h, self.grid_off.ctypes.data, self.pos_off.ctypes.data, self.df_off.ctypes.data), "batch_offsets")
        self.natoms = np.array([int(it["natoms"]) for it in self.items])
        self.tune_sample = tune_sample
        self._auto = tuning == "auto"
        self._warm_key = ""

    # ------------------------------------------------------------ plain mode
    @classmethod
    def plain(cls, lvecs, ns, scans, tipR0=(0.0, 0.0, 3.0), k_lat=0.25, k_rad=30.0, iZPP=8, probe_q=-0.1, tip="s",
              amplitude=0.8, k0=1800.0, f0=30300.0, Fmax=10.0, relax=None, molecules=None, natoms=None,
              grid_f64=True, keep_iters=False, tuning=None, stream=None, species=None, fp64_radius=None):
        """ProbeParticle-semantics batch (PlainScan per item): lvecs, ns and
        scans are per-item grid lattices (4,3), grid sizes and scan dicts.
        molecules: (xyzs, Zs, qs) per item (uploaded), or natoms per item and
        set_molecules later."""
        if molecules is not None:
            natoms = [len(m[1]) for m in molecules]
        if natoms is None:
            raise ValueError("give molecules or natoms")
        items = []
        shared = None
        for k, (lvec, n, scan) in enumerate(zip(lvecs, ns, scans)):
            model, rp, tip_d, scan_d, w, df_scale = plain_setup(
                lvec, scan, tipR0=tipR0, k_lat=k_lat, k_rad=k_rad, probe_q=probe_q, tip=tip, amplitude=amplitude,
                k0=k0, f0=f0, Fmax=Fmax, relax=relax, fp64_radius=fp64_radius)
            if shared is None:
                shared = (model, rp, tip_d, w, df_scale)
            elif len(w) != len(shared[3]) or not np.array_equal(w, shared[3]):
                raise ValueError("all items of a batch need the same df weights (same dz)")
            items.append(dict(n=n, cell=np.asarray(lvec)[1:], natoms=int(natoms[k]), scan=scan_d))
        model, rp, tip_d, w, df_scale = shared
        b = cls(items, model, rp, tip_d, df_weights=w, df_scale=df_scale, grid_f64=grid_f64, keep_iters=keep_iters,
                tuning=tuning, stream=stream)
        b.lvecs = [np.asarray(l, dtype=np.float64) for l in lvecs]
        b.iZPP = iZPP
        b.species = species or PPU.loadSpecies()
        b.scans = [dict(s) for s in scans]
        if molecules is not None:
            b.set_molecules(molecules)
        return b

    # ------------------------------------------------------------ tilted mode
    @classmethod
    def tilted(cls, lvecs, ns, rot=np.eye(3), pos0s=None, start=(-5.0, -5.0), end=(5.0, 5.0), scan_n=(32, 32), nz=20,
               zstep=0.1, tipR0=(0.0, 0.0, 3.0), k_lat=0.25, k_rad=30.0, relax_params=None, Qs=(-0.1, 0.0, 0.0, 0.0),
               QZs=(0.0, 0.0, 0.0, 0.0), iZPP=8, df_steps=10, k0=1800.0, f0=30300.0, surf_ff=0.0, molecules=None,
               natoms=None, grid_f64=False, keep_iters=False, tuning=None, stream=None, species=None, tune_sample=32):
        """Batch of tilted scans with the OpenCL semantics of
        RelaxedScanner.run_relaxStrokesTilted (relax.cl:391-464) on the
        evalLJC_QZs_noPos grid (FF.cl:366-423): per item a grid lattice (4,3)
        and size, and pos0 (scan origin in world coordinates; default the
        grid origin + (0, 0, 12)); shared tip frame ``rot`` (rows A, B, C),
        scan window start/end along A and B, nz steps of zstep along -C,
        tip stiffness (k_lat, k_lat, 0, k_rad) N/m, the Qs/QZs charge stencil
        and df from get_simple_df_weight(df_steps, dz=zstep) as AFMulator
        (AFMulator.py:125-133).  LJ parameters come from the REA table
        (getAtomsREA with alphaFac -1, REA2LJ), as AFMulator takes them.
        Grids are float4 with FP32 pairs (the reference's precision) unless
        grid_f64."""
        from .ocl.relax import DEFAULT_relax_params, rotTip

        if molecules is not None:
            natoms = [len(m[1]) for m in molecules]
        if natoms is None:
            raise ValueError("give molecules or natoms")
        rot = np.asarray(rot, dtype=np.float64)
        dTip, tipRot, dpos0Tip, _ = rotTip(rot, zstep, (tipR0[0], tipR0[1], tipR0[2]))
        st = np.array([k_lat, k_lat, 0.0, k_rad], dtype=np.float32) / -PPU.eVA_Nm
        tip_d = D.tip_desc(np.asarray(dpos0Tip[:3], np.float64), st[:3].astype(np.float64), float(st[3]),
                           float(np.float32(dpos0Tip[3])), rot=np.asarray(tipRot, np.float64)[:, :3], surf_ff=float(surf_ff))
        rp_ = np.asarray(DEFAULT_relax_params if relax_params is None else relax_params, dtype=np.float32)
        rp = D.relax_params(mode=1, alg=1, max_iters=128, conv_f2=float(np.float32(1e-8)), dt=float(rp_[0]),
                            damping=float(rp_[1]), dtmin_frac=float(np.float32(0.1)), f2safe=float(np.float32(1e-8)))
        model = D.ff_model(1, with_lj=True, stencil=list(zip(np.float32(Qs).tolist(), np.float32(QZs).tolist())),
                           with_energy=True, fmax=None)
        w = PPU.get_simple_df_weight(df_steps, dz=zstep) * PPU.eVA_Nm * f0 / k0
        a = np.asarray(tipRot, np.float64)[0, :3]
        b_ = np.asarray(tipRot, np.float64)[1, :3]
        nx, ny = scan_n
        items = []
        for k, (lvec, n) in enumerate(zip(lvecs, ns)):
            lvec = np.asarray(lvec, dtype=np.float64)
            p = np.asarray(pos0s[k] if pos0s is not None else lvec[0] + np.array([0.0, 0.0, 12.0]), np.float64) - lvec[0]
            scan_d = D.scan_desc(nx, ny, nz, p0=p + start[0] * a + start[1] * b_,
                                 da=a * ((end[0] - start[0]) / max(nx - 1, 1)), db=b_ * ((end[1] - start[1]) / max(ny - 1, 1)),
                                 dstep=np.asarray(dTip[:3], np.float64))
            items.append(dict(n=n, cell=lvec[1:], natoms=int(natoms[k]), scan=scan_d))
        b = cls(items, model, rp, tip_d, df_weights=w, df_scale=1.0, grid_f64=grid_f64, keep_iters=keep_iters,
                tuning=tuning, stream=stream, tune_sample=tune_sample)
        b.lvecs = [np.asarray(l, dtype=np.float64) for l in lvecs]
        b.iZPP = iZPP
        b.species = species or PPU.loadSpecies()
        b._rea = True
        if molecules is not None:
            b.set_molecules(molecules)
        return b

    def set_molecules(self, molecules):
        """molecules: sequence of (xyzs, Zs, qs) in item order (atom counts
        may differ from the plan's, see set_atoms).  LJ parameters: the CPU table
        (getAtomsLJ) for plain batches, the REA table for tilted ones."""
        packs = []
        for k, (xyzs, Zs, qs) in enumerate(molecules):
            xyzs = np.asarray(xyzs, dtype=np.float64)
            if getattr(self, "_rea", False):
                cLJs = PPU.REA2LJ(PPU.getAtomsREA(self.iZPP, Zs, self.species, alphaFac=-1.0))
            else:
                cLJs = PPU.getAtomsLJ(self.iZPP, Zs, self.species)
            xyzq = np.concatenate([xyzs - self.lvecs[k][0][None, :], np.asarray(qs, dtype=np.float64)[:, None]], axis=1)
            packs.append(D.pack_atoms(xyzq, cLJs, self.model.coulomb_const))
        self.set_atoms(packs)

    def set_atoms(self, packs):
        """packs: per item (natoms, 6) ppafm_atom records (D.pack_atoms).  The
        atom counts may differ from the plan's: the plan keeps its grids and
        scans and takes the new molecules."""
        if len(packs) != self.nitems:
            raise ValueError(f"expected {self.nitems} molecules, got {len(packs)}")
        counts = np.array([len(np.asarray(p).reshape(-1, 6)) for p in packs], dtype=np.int32)
        allp = np.ascontiguousarray(np.concatenate([np.asarray(p, dtype=np.float64).reshape(-1, 6) for p in packs], axis=0))
        ptr = allp.ctypes.data if len(allp) else None
        if np.array_equal(counts, self.natoms):
            check(lib().ppafm_batch_set_atoms(self._h, ptr), "batch_set_atoms")
        else:
            check(lib().ppafm_batch_set_atoms_n(self._h, counts.ctypes.data, ptr), "batch_set_atoms_n")
            self.natoms = counts.astype(self.natoms.dtype)
            for k, it in enumerate(self.items):
                it["natoms"] = int(counts[k])
        if self._auto:
            self._auto = False
            self.set_tuning(tuned(self, nsample=self.tune_sample))

    # ------------------------------------------------------------ running
    def run(self):
        check(lib().ppafm_batch_run(self._h), "batch_run")

    def run_stages(self, stages):
        check(lib().ppafm_batch_run_stages(self._h, int(stages)), "batch_run_stages")

    def synchronize(self):
        D.synchronize(self.stream)

    def tune(self, nsample=32, reps=2):
        best = _lib.BatchTuning()
        ms = ctypes.c_double()
        check(lib().ppafm_batch_tune(self._h, int(nsample), int(reps), ctypes.byref(best), ctypes.byref(ms)), "batch_tune")
        return best.as_dict(), float(ms.value)

    @property
    def tuning(self):
        t = _lib.BatchTuning()
        check(lib().ppafm_batch_get_tuning(self._h, ctypes.byref(t)), "batch_get_tuning")
        return t.as_dict()

    def set_tuning(self, tuning):
        t = _lib.BatchTuning(**tuning)
        check(lib().ppafm_batch_set_tuning(self._h, ctypes.byref(t)), "batch_set_tuning")

    def problem_key(self):
        """Problem class of the autotuner cache."""
        it = self.items[0]
        s = it["scan"]
        return f"mode{self.rp.mode}-{'f64' if self.grid_f64 else 'f32'}-scan{s.nx}x{s.ny}x{s.nz}{self._warm_key}-v2"

    # ------------------------------------------------------------ outputs
    def _download(self, which, off_elems, shape, dtype):
        dtype = np.dtype(dtype)
        out = np.empty(shape, dtype)
        base = lib().ppafm_batch_buffer(self._h, which)
        if base is None:
            raise HipError("buffer not allocated")
        check(lib().ppafm_memcpy_d2h(out.ctypes.data, base + int(off_elems), out.nbytes, self.stream), "download")
        return out

    def _scan_shape(self, k):
        s = self.items[k]["scan"]
        return s.nx, s.ny, s.nz

    def grid(self, k):
        n = tuple(int(v) for v in self.items[k]["n"])
        esz = 32 if self.grid_f64 else 16
        return self._download(BUF_GRIDS, self.grid_off[k] * esz, n + (4,), np.float64 if self.grid_f64 else np.float32)

    def positions_raw(self, k):
        """Relaxed positions (nx, ny, nz, 3) float32, grid frame, stroke order."""
        return self._download(BUF_POS, self.pos_off[k] * 12, self._scan_shape(k) + (3,), np.float32)

    def positions(self, k):
        """Relaxed positions (nx, ny, nz, 3) in the molecule frame (plain mode)."""
        return self.positions_raw(k).astype(np.float64) + self.lvecs[k][0][None, None, None, :]

    def forces(self, k):
        return self._download(BUF_FORCE, self.pos_off[k] * 16, self._scan_shape(k) + (4,), np.float32)

    def iterations(self, k):
        return self._download(BUF_ITERS, self.pos_off[k] * 4, self._scan_shape(k), np.int32)

    def status(self, k):
        """Per-position status (nx, ny, nz) uint8 of item k (plans made with
        keep_iters): 0 converged by the stop rule, bit 0 = the iteration cap
        ended the step, bit 1 = non-finite position or grid force."""
        return self._download(BUF_STATUS, self.pos_off[k], self._scan_shape(k), np.uint8)

    def df_raw(self, k):
        """df (nx, ny, nz - nw + 1) in stroke order (index 0 = top); empty
        along z when the scan has fewer steps than the df weights."""
        nx, ny, nz = self._scan_shape(k)
        nzout = nz - len(self.w) + 1
        if nzout < 1:
            return np.zeros((nx, ny, 0), np.float32)
        return self._download(BUF_DF, self.df_off[k] * 4, (nx, ny, nzout), np.float32)

    def df_host(self, k):
        """df with z ascending, like PlainScan.df_host / Fz2df (plain mode)."""
        return self.df_raw(k)[:, :, ::-1].copy()

    def counts(self):
        """(force evaluations, positions) of the last run."""
        c = self.counters()
        return c[0], c[1]

    def counters(self):
        """(force evaluations, positions, warm-start fallbacks) of the last run."""
        c = np.empty(3, np.uint64)
        check(lib().ppafm_memcpy_d2h(c.ctypes.data, lib().ppafm_batch_buffer(self._h, BUF_COUNTERS), 24, self.stream), "download")
        return int(c[0]), int(c[1]), int(c[2])

    def lane_use(self):
        """Force evaluations of the last relaxation over its lane slots (64 per
        loop pass of each wave): the fraction of SIMD lanes that did work while
        a wave waited for its slowest column."""
        c = np.empty(4, np.uint64)
        check(lib().ppafm_memcpy_d2h(c.ctypes.data, lib().ppafm_batch_buffer(self._h, BUF_COUNTERS), 32, self.stream), "download")
        return float(c[0]) / float(c[3]) if c[3] else float("nan")

    # ------------------------------------------------------ learned warm start
    def set_warmstart(self, net=None, enable=True, net_passes=1, newton_max=16, newton_max0=12, newton_retry=0,
                      newton_tol=1e-3, newton_cap=0.3, extrap_max=1.0, residual_max=RESIDUAL_MAX):
        """Warm starts (``enable`` False: plain FIRE).  Every step first runs
        up to ``newton_max`` Newton iterations on the interpolated field from
        the continuation start (step 0: ``newton_max0`` from the network's
        estimate after ``net_passes`` passes when ``net`` is given, else from
        the plain start), each step at most ``newton_cap`` A and only where
        the force Jacobian is stable; a step below ``newton_tol`` A means
        converged and FIRE starts there, otherwise FIRE starts from the
        continuation start (counters()[2]).  Every ``newton_retry`` FIRE
        iterations Newton is tried again from FIRE's position.  FIRE ends
        every step by the plan's stop rule.  ``extrap_max`` (A, 0 = off):
        steps from the third on start their Newton run from the continuation
        moved by the last change of the probe's offset from the tip, where
        that change is at most ``extrap_max`` long (fallbacks still start
        FIRE from the continuation itself).  ``residual_max`` (eV/A, 0 =
        off): with a network and ``newton_max0`` > 0, the total force at
        each column's estimate is evaluated first, and where its magnitude
        exceeds ``residual_max`` that column's step 0 is plain FIRE from the
        reference's start instead (a fallback step in counters()[2])."""
        p = _lib.WarmParams(int(bool(enable)), int(net_passes), int(newton_max), int(newton_max0), int(newton_retry),
                            float(newton_tol), float(newton_cap), float(extrap_max), float(residual_max))
        self._ws = net if enable else None
        check(lib().ppafm_batch_set_warmstart(self._h, ctypes.byref(p), net._h if (net is not None and enable) else None),
              "batch_set_warmstart")
        self._warm_key = "" if not enable else ("-warm-net" if net is not None else "-warm")

    def set_plain(self):
        """Plain FIRE from the continuation start at every step."""
        self.set_warmstart(None, enable=False)

    def polish(self, conv_f2=1e-18, max_iters=4000):
        """FIRE from the last run's relaxed positions to |F|^2 < conv_f2:
        device arrays (positions, 3) float32 and the iteration counts."""
        n = self.total_positions()
        out = D.DeviceArray((n, 3), np.float32)
        it = D.DeviceArray((n,), np.int32)
        check(lib().ppafm_batch_polish(self._h, float(conv_f2), int(max_iters), out.ptr, it.ptr), "batch_polish")
        return out, it

    def ws_samples(self, truth, sel, jitter=0.0, seed=0, out=None, offset=0):
        """Training pairs at the global position indices ``sel`` (host int64
        array): device arrays patches (n, 232) bf16 and targets (n, 3); the
        patch centre is moved by up to ``jitter`` A (half along z), by a hash
        of the position index and the seed.  ``out`` = (P, T): write into
        those device arrays from row ``offset`` instead."""
        sel = np.ascontiguousarray(sel, dtype=np.int64)
        if len(sel) and (sel.min() < 0 or sel.max() >= self.total_positions()):
            raise ValueError("sample index out of range")
        sd = D.DeviceArray.from_host(sel if len(sel) else np.zeros(1, np.int64))
        if out is None:
            P = D.DeviceArray((max(len(sel), 1), 232), np.uint16)
            T = D.DeviceArray((max(len(sel), 1), 3), np.float32)
            offset = 0
        else:
            P, T = out
            if offset < 0 or offset + len(sel) > P.shape[0] or P.shape[0] != T.shape[0]:
                raise ValueError("samples do not fit the output arrays")
        check(lib().ppafm_batch_ws_samples(self._h, truth.ptr, len(sel), sd.ptr, P.ptr + 464 * int(offset),
                                           T.ptr + 12 * int(offset), float(jitter), int(seed) & 0xFFFFFFFF),
              "batch_ws_samples")
        self.synchronize()
        return P, T

    def predictions(self, k):
        """Step-0 estimates of the network (nx, ny, 3) of item k (grid frame)."""
        nx, ny, nz = self._scan_shape(k)
        col = sum(int(np.prod(self._scan_shape(q)[:2])) for q in range(k) if self._scan_shape(q)[2] > 0)
        return self._download(7, col * 12, (nx, ny, 3), np.float32)

    def total_positions(self):
        return int(self.pos_off[-1])

    def total_grid_points(self):
        return int(self.grid_off[-1])

    def close(self):
        if getattr(self, "_h", None):
            self._lib.ppafm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ tuning cache


def cache_path():
    p = os.environ.get("PPAFM_TUNING_CACHE")
    return pathlib.Path(p) if p else pathlib.Path.home() / ".cache" / "ppafm_amd" / "tuning.json"


def _load_cache(path):
    try:
        with open(path) as f:
            d = json.load(f)
        return d if isinstance(d, dict) else {}
    except (OSError, ValueError):
        return {}


def tuned(batch, nsample=32, reps=2, path=None):
    """Cached autotuner winner for this device and the batch's problem class;
    tunes on the batch (and stores the winner) on a miss."""
    path = pathlib.Path(path) if path else cache_path()
    name, ncu = device_info()
    key = f"{name}-{ncu}cu-{batch.problem_key()}"
    cache = _load_cache(path)
    hit = cache.get(key)
    if isinstance(hit, dict) and {"tile_y", "waves_per_block", "blocks_per_cu"} <= set(hit):
        return {k: int(hit.get(k, 0)) for k in ("tile_y", "waves_per_block", "blocks_per_cu", "refill")}
    best, ms = batch.tune(nsample=nsample, reps=reps)
    cache[key] = dict(best, ms=ms, nsample=int(min(nsample, batch.nitems)))
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        tmp = path.with_suffix(".tmp")
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass
    return best
