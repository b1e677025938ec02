"""Device-resident relaxed scan with the ProbeParticle (CPU) semantics.

One object owns the FP32 force-field grid, the relaxation outputs and the df
stack on the GPU and runs the three stages of the path back to back on one
stream:

  1. grid build   FF = clamp(LJ) + q * clamp(El)  (HighLevel.py:143-145,260-387)
  2. relaxation   relaxTipStrokes_omp semantics    (ProbeParticle.cpp:633-701)
  3. Fz -> df     Fz2df with get_df_weight         (common.py:333-430)

Positions are relative to the grid origin inside; the public accessors return
them in the molecule frame (origin added back, HighLevel.py:170-172).
"""

import numpy as np

from . import common as PPU
from . import device as D


def plain_setup(lvec, scan, tipR0=(0.0, 0.0, 3.0), k_lat=0.25, k_rad=30.0, probe_q=-0.1, tip="s", amplitude=0.8,
                k0=1800.0, f0=30300.0, Fmax=10.0, relax=None, fp64_radius=None):
    """Descriptors of one ProbeParticle-semantics scan: (ff model, relax
    params, tip, scan, df weights, df scale).  Positions are relative to the
    grid origin lvec[0].  fp64_radius: see device.ff_model (float4 grids)."""
    lvec = np.asarray(lvec, dtype=np.float64)
    model = D.ff_model(0, with_lj=True, stencil=D.PROBE_STENCILS[tip], probe_q=probe_q, fmax=Fmax, fp64_radius=fp64_radius)
    rp = D.relax_params(mode=0, **(relax or {}))
    kS = np.array([k_lat, k_lat, 0.0]) / -PPU.eVA_Nm
    tip_d = D.cpu_tip(tipR0[2], k_rad / -PPU.eVA_Nm, (tipR0[0], tipR0[1]), kS)
    nx, ny = scan["scan_n"]
    nz = scan["nz"]
    x0, x1 = scan["scan_x"]
    y0, y1 = scan["scan_y"]
    o = lvec[0]
    scan_d = D.scan_desc(
        nx, ny, nz,
        p0=(x0 - o[0], y0 - o[1], scan["z_top"] - o[2]),
        da=((x1 - x0) / max(nx - 1, 1), 0, 0),
        db=(0, (y1 - y0) / max(ny - 1, 1), 0),
        dstep=(0, 0, -scan["dz"]),
    )
    # Fz2df: correlating the stroke-ordered (z descending) Fz with the
    # weights as returned equals np.convolve on the ascending stack.
    w = PPU.get_df_weight(amplitude, dz=scan["dz"])
    df_scale = 16.0217656 * f0 / k0
    return model, rp, tip_d, scan_d, w, df_scale


class PlainScan:
    """Relaxed AFM scan of one molecule, ProbeParticle semantics, on the GPU.

    lvec: (4,3) grid origin and cell; n: grid points; scan: dict with
    ``scan_n`` (nx, ny), ``scan_x``/``scan_y`` apex ranges, ``z_top``, ``nz``,
    ``dz`` (apex descends by dz per step).  Tip: tipR0 (x0, y0, L), k_lat,
    k_rad in N/m; probe element iZPP and charge probe_q (monopole unless
    ``tip`` says "pz"/"dz2").  df: amplitude A (Fz2df ``amplitude``), k0, f0.

    grid_f64 (default) stores the grid as double4 with every pair summed in
    FP64, which keeps the relaxed positions within 1e-3 A of the FP64
    reference everywhere; grid_f64=False stores float4 (half the bytes per
    gather), whose rounding the FIRE stop rule turns into >1e-3 A stopping
    differences on a few ill-conditioned strokes (about 1 in 10^5-10^6).
    A float4 grid sums pairs closer than fp64_radius (default 4 A) in FP64;
    fp64_radius=0 sums every pair in FP32 (the FP32 grid kernel).
    """

    def __init__(self, lvec, n, scan, tipR0=(0.0, 0.0, 3.0), k_lat=0.25, k_rad=30.0, iZPP=8, probe_q=-0.1, tip="s",
                 amplitude=0.8, k0=1800.0, f0=30300.0, Fmax=10.0, relax=None, species=None, stream=None,
                 grid_f64=True, fp64_radius=None):
        self.lvec = np.asarray(lvec, dtype=np.float64)
        self.n = tuple(int(v) for v in n)
        self.scan = dict(scan)
        self.stream = stream
        self.iZPP = iZPP
        self.species = species or PPU.loadSpecies()
        self.model, self.rp, self.tip, self.scan_d, self.w, self.df_scale = plain_setup(
            self.lvec, self.scan, tipR0=tipR0, k_lat=k_lat, k_rad=k_rad, probe_q=probe_q, tip=tip, amplitude=amplitude,
            k0=k0, f0=f0, Fmax=Fmax, relax=relax, fp64_radius=fp64_radius)
        nx, ny = self.scan["scan_n"]
        nz = self.scan["nz"]
        self.grid_f64 = bool(grid_f64)
        self.grid = D.DeviceArray(self.n + (4,), np.float64 if self.grid_f64 else np.float32)
        self.pos = D.DeviceArray((nx, ny, nz, 3), np.float32)
        self.force = D.DeviceArray((nx, ny, nz, 4), np.float32)
        self.iters = D.DeviceArray((nx, ny, nz), np.int32)
        self.status_d = D.DeviceArray((nx, ny, nz), np.uint8)
        self.counters = D.DeviceArray((2,), np.uint64)
        self.counters.zero()
        self.nzout = nz - len(self.w) + 1
        self.df = D.DeviceArray((nx * ny, max(self.nzout, 1)), np.float32)
        self.atoms = None
        self.natoms = 0

    # ---------------------------------------------------------------- inputs
    def set_molecule(self, xyzs, Zs, qs):
        xyzs = np.asarray(xyzs, dtype=np.float64)
        cLJs = PPU.getAtomsLJ(self.iZPP, Zs, self.species)
        xyzq = np.concatenate([xyzs - self.lvec[0][None, :], np.asarray(qs, dtype=np.float64)[:, None]], axis=1)
        packed = D.pack_atoms(xyzq, cLJs, self.model.coulomb_const)
        self.natoms = len(packed)
        self.atoms = D.DeviceArray.from_host(packed if len(packed) else np.zeros((1, 6)))

    # ---------------------------------------------------------------- stages
    def build_grid(self):
        import ctypes

        from ._lib import check, lib

        g = D.grid_desc(self.n, self.lvec[1:], f64=self.grid_f64)
        check(lib().ppafm_ff_build_dev(ctypes.byref(g), ctypes.byref(self.model), self.natoms, self.atoms.ptr, self.grid.ptr, self.stream), "ff_build")

    def relax(self, counters=True):
        D.relax(self.n, self.lvec[1:], self.grid, self.rp, self.tip, self.scan_d, pos=self.pos, force=self.force, iters=self.iters,
                counters=self.counters if counters else None, stream=self.stream, status=self.status_d)

    def to_df(self):
        if self.nzout < 1:
            return
        nx, ny = self.scan["scan_n"]
        nz = self.scan["nz"]
        D.convolve_z_dev(self.force, nx * ny, nz, self.w, scale=self.df_scale, in_stride=nz * 4, in_step=4, in_offset=2, out=self.df, stream=self.stream)

    def run(self):
        self.build_grid()
        self.relax()
        self.to_df()

    def save_grid(self, fname, data_format="xsf", head=None, nthreads=0):
        """Write the device grid's force as a vector field the reference's
        tools read (io.save_vec_field layout): fname_x/_y/_z.xsf formatted by
        the library straight from the device copy, or fname.npz."""
        from . import io

        head = io.XSF_HEAD_DEFAULT if head is None else head
        if data_format == "xsf":
            from ._lib import check, lib

            for k, c in enumerate("xyz"):
                path = f"{fname}_{c}.xsf"
                io.write_xsf_header(path, head, self.n, self.lvec)
                check(lib().ppafm_xsf_write_data_dev(path.encode(), 1, self.grid.ptr, int(self.grid_f64), k, *self.n,
                                                     int(nthreads)), "save_grid")
        elif data_format == "npy":
            io.save_vec_field(fname, np.ascontiguousarray(self.grid_host()[..., :3], dtype=np.float64), self.lvec,
                              data_format="npy")
        else:
            raise ValueError(f"unknown data_format {data_format!r}")

    def set_grid(self, FF, E=None):
        """Use a force field from elsewhere -- (nx,ny,nz,3), e.g. from
        io.load_vec_field -- as the device grid instead of building one
        (run the relax() and to_df() stages after it, not run())."""
        FF = np.asarray(FF)
        if FF.shape != self.n + (3,):
            raise ValueError(f"FF must have shape {self.n + (3,)}, got {FF.shape}")
        g = np.zeros(self.n + (4,), dtype=np.float64 if self.grid_f64 else np.float32)
        g[..., :3] = FF
        if E is not None:
            g[..., 3] = E
        self.grid.upload(g)

    def stiffness(self, which=0, ddisp=0.05):
        """Eigenvalues (nx,ny,nz,3) of the probe's dynamical matrix at the
        relaxed positions and the first ``which`` eigenvector arrays
        (core.stiffnessMatrix / ProbeParticle.cpp:703-750), on the GPU grid."""
        ev, vecs = D.stiffness(self.n, self.lvec[1:], self.grid, self.tip, self.scan_d, self.pos, which=which,
                               ddisp=ddisp, stream=self.stream)
        return ev.download(), [v.download() if v is not None else None for v in vecs]

    def reset_counters(self):
        self.counters.zero(self.stream)

    # --------------------------------------------------------------- outputs
    def grid_host(self):
        return self.grid.download()

    def positions(self):
        """Relaxed probe positions (nx, ny, nz, 3) in stroke order, molecule frame."""
        return self.pos.download().astype(np.float64) + self.lvec[0][None, None, None, :]

    def forces(self):
        return self.force.download()

    def iterations(self):
        return self.iters.download()

    def status(self):
        """Per-position status (nx, ny, nz) uint8: 0 converged by the stop
        rule, bit 0 = the iteration cap ended the step, bit 1 = non-finite
        position or grid force."""
        return self.status_d.download()

    def df_host(self):
        """df (nx, ny, nzout) with z ascending, like Fz2df on the reversed stack."""
        nx, ny = self.scan["scan_n"]
        return self.df.download().reshape(nx, ny, self.nzout)[:, :, ::-1].copy()

    def counts(self):
        c = self.counters.download()
        return int(c[0]), int(c[1])
