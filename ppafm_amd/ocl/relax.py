"""Drop-in for ``ppafm.ocl.relax`` (ocl/relax.py:17-367): the relaxed
scanner of the GPU path, running on MI355X through libppafm_hip.so.

``RelaxedScanner.run_relaxStrokesTilted`` follows relaxStrokesTilted
(relax.cl:391-464): FIRE with dt 0.5 / dtmin 0.05, at most 128 steps per tip
position, dt and alpha kept along the stroke, the tip springs and the
surface force acting in the tip frame given by ``setScanRot``, output force
rotated into the tip frame and the probe path.  ``run`` follows the untilted
relaxStrokes (relax.cl:316-363): the same FIRE with the springs in the world
frame.  The force field is sampled by exact software trilinear interpolation
(the reference uses the hardware image filter).

Buffers live on the GPU between calls.  Force fields are (nx, ny, nz, 4)
arrays indexed FE[ix, iy, iz] = (Fx, Fy, Fz, E), as ``AFMulator.downloadFF``
returns them; the scan lattice is (scan_dim[0], scan_dim[1]) pixels of
scan_dim[2] tip positions.
"""

import numpy as np

from .. import device as D

DEFAULT_stiffness = np.array([-0.03, -0.03, -0.03, -1.0], dtype=np.float32)  # relax.py:18
DEFAULT_relax_params = np.array([0.5, 0.1, 0.02, 0.5], dtype=np.float32)  # relax.py:20


def mat3x3to4f(M):
    out = []
    for k in range(3):
        a = np.zeros(4, dtype=np.float32)
        a[:3] = M[k]
        out.append(a)
    return tuple(out)


def getInvCell(lvec):
    """Rows of (cell^-1)^T: position -> fractional coordinates (relax.py:43-47)."""
    cell = np.asarray(lvec)[1:4, 0:3]
    return mat3x3to4f(np.transpose(np.linalg.inv(cell)))


def preparePossRot(scan_dim, pos0, avec, bvec, start=(-5.0, -5.0), end=(5.0, 5.0)):
    xs = np.linspace(start[0], end[0], scan_dim[0])
    ys = np.linspace(start[1], end[1], scan_dim[1])
    As, Bs = np.meshgrid(xs, ys, indexing="ij")
    poss = np.zeros(As.shape + (4,), dtype=np.float32)
    for c in range(3):
        poss[:, :, c] = pos0[c] + As * avec[c] + Bs * bvec[c]
    return poss


def rotTip(rot, zstep, tipR0=(0.0, 0.0, 4.0)):
    """Tip step, tip frame rows (A, B, C with C.w = -zstep), equilibrium
    offset in the tip frame and in the world frame (relax.py:72-87)."""
    rot = np.asarray(rot, dtype=np.float64)
    dTip = np.zeros(4, dtype=np.float32)
    dTip[:3] = rot[2] * -zstep
    tipRot = mat3x3to4f(rot)
    tipRot[2][3] = -zstep
    dpos0Tip = np.zeros(4, dtype=np.float32)
    dpos0Tip[0] = tipR0[0]
    dpos0Tip[1] = tipR0[1]
    dpos0Tip[2] = -np.sqrt(tipR0[2] ** 2 - tipR0[0] ** 2 - tipR0[1] ** 2)
    dpos0Tip[3] = tipR0[2]
    dpos0 = np.zeros(4, dtype=np.float32)
    dpos0[:3] = np.dot(rot.transpose(), dpos0Tip[:3])
    dpos0[3] = tipR0[2]
    return dTip, tipRot, dpos0Tip, dpos0


def _f64(v):
    return np.asarray(v, dtype=np.float32).astype(np.float64)


class RelaxedScanner:
    """relax.py:93-367 on the GPU (no OpenCL context needed)."""

    def __init__(self):
        self.stiffness = DEFAULT_stiffness.copy()
        self.relax_params = DEFAULT_relax_params.copy()
        self.zstep = 0.1
        self.start = (-5.0, -5.0)
        self.end = (5.0, 5.0)
        self.tipR0 = 4.0
        self.surfFF = np.zeros(4, dtype=np.float32)
        self.lvec = None
        self.scan_dim = None
        self._grid = None
        self._n = None
        self._force = None
        self._paths = None
        self._rot = np.eye(3)
        self.pos0 = None

    # -------------------------------------------------------------- buffers
    def prepareBuffers(self, FEin_np=None, lvec=None, FEin_shape=None, scan_dim=None, nDimConv=None, nDimConvOut=None, **_):
        if lvec is not None:
            self.lvec = np.asarray(lvec, dtype=np.float64)
            self.invCell = getInvCell(self.lvec)
        if FEin_np is not None:
            self.updateBuffers(FEin=FEin_np)
        elif FEin_shape is not None:
            self._alloc_grid(tuple(int(v) for v in FEin_shape[:3]))
        if scan_dim is not None:
            self.scan_dim = tuple(int(v) for v in scan_dim)
            nx, ny, nz = self.scan_dim
            self._force = D.DeviceArray((nx, ny, nz, 4), np.float32)
            self._paths = D.DeviceArray((nx, ny, nz, 3), np.float32)
            if nDimConv is not None:
                self.nDimConv = int(nDimConv)
                self.nDimConvOut = int(nDimConvOut)

    def _alloc_grid(self, n):
        if self._grid is None or self._n != n:
            self._grid = D.DeviceArray(n + (4,), np.float64)
            self._n = n

    def releaseBuffers(self):
        for a in (self._grid, self._force, self._paths):
            if a is not None:
                a.free()
        self._grid = self._force = self._paths = None

    def updateBuffers(self, FEin=None, lvec=None, WZconv=None):
        if lvec is not None:
            self.lvec = np.asarray(lvec, dtype=np.float64)
            self.invCell = getInvCell(self.lvec)
        if FEin is not None:
            FEin = np.asarray(FEin)
            if FEin.ndim != 4 or FEin.shape[3] != 4:
                raise ValueError(f"FEin must be (nx, ny, nz, 4), got {FEin.shape}")
            self._alloc_grid(tuple(int(v) for v in FEin.shape[:3]))
            self._grid.upload(np.ascontiguousarray(FEin, dtype=np.float64))
        if WZconv is not None:
            self.WZconv = np.asarray(WZconv, dtype=np.float64)

    # ---------------------------------------------------------------- scan
    def preparePosBasis(self, start=(-5.0, -5.0), end=(5.0, 5.0)):
        self.start = start
        self.end = end

    def preparePossRot(self, pos0, avec, bvec):
        return preparePossRot(self.scan_dim, pos0, avec, bvec, self.start, self.end)

    def setScanRot(self, pos0, rot=None, zstep=None, tipR0=(0.0, 0.0, 4.0)):
        if rot is None:
            rot = np.eye(3)
        if zstep:
            self.zstep = zstep
        self._rot = np.asarray(rot, dtype=np.float64)
        self.pos0 = np.asarray(pos0, dtype=np.float64)
        self.tipR0 = tipR0
        self.dTip, self.tipRot, self.dpos0Tip, self.dpos0 = rotTip(self._rot, self.zstep, tipR0)
        return self.preparePossRot(pos0, self._rot[0], self._rot[1])

    def _scan(self, nz):
        nx, ny = self.scan_dim[:2]
        a, b = _f64(self.tipRot[0][:3]), _f64(self.tipRot[1][:3])
        s0, s1 = self.start, self.end
        da = a * ((s1[0] - s0[0]) / max(nx - 1, 1))
        db = b * ((s1[1] - s0[1]) / max(ny - 1, 1))
        p0 = self.pos0 + s0[0] * a + s0[1] * b
        return D.scan_desc(nx, ny, nz, p0=p0, da=da, db=db, dstep=_f64(self.dTip[:3]))

    def _relax_params(self):
        rp = self.relax_params
        return D.relax_params(mode=1, alg=1, max_iters=128, conv_f2=float(np.float32(1e-8)), dt=float(rp[0]), damping=float(rp[1]),
                              dtmin_frac=float(np.float32(0.1)), f2safe=float(np.float32(1e-8)))

    def _run(self, tip, nz, FEout, FEin, lvec):
        self.updateBuffers(FEin=FEin, lvec=lvec)
        if nz is None:
            nz = self.scan_dim[2]
        if self._grid is None or self.lvec is None or self.pos0 is None:
            raise RuntimeError("set the force field (prepareBuffers), lvec and the scan (setScanRot) first")
        nx, ny = self.scan_dim[:2]
        if self._force is None or self._force.shape[2] != nz:
            self._force = D.DeviceArray((nx, ny, nz, 4), np.float32)
            self._paths = D.DeviceArray((nx, ny, nz, 3), np.float32)
        cell = _f64(self.lvec[1:4, :3])
        D.relax(self._n, cell, self._grid, self._relax_params(), tip, self._scan(nz), pos=self._paths, force=self._force)
        self._nz = nz
        if FEout is not None:
            if FEout.shape != (nx, ny, nz, 4):
                raise ValueError(f"FEout must have shape {(nx, ny, nz, 4)}")
            FEout[...] = self._force.download()
            return FEout
        return self._force.download()

    def run_relaxStrokesTilted(self, FEout=None, FEin=None, lvec=None, nz=None, bCopy=True, bFinish=True):
        """Tilted strokes (relax.cl:391-464); returns FE (nx, ny, nz, 4) in the tip frame."""
        st = _f64(self.stiffness)
        tip = D.tip_desc(_f64(self.dpos0Tip[:3]), st[:3], float(st[3]), float(np.float32(self.dpos0Tip[3])), rot=_f64(np.array(self.tipRot)[:, :3]),
                         surf_ff=float(np.float32(self.surfFF[0])))
        out = self._run(tip, nz, FEout, FEin, lvec)
        return out if bCopy else None

    def run(self, FEout=None, FEin=None, lvec=None, nz=None):
        """Untilted strokes (relax.cl:316-363): springs in the world frame with
        the world-frame equilibrium offset dpos0."""
        st = _f64(self.stiffness)
        tip = D.tip_desc(_f64(self.dpos0[:3]), st[:3], float(st[3]), float(np.float32(self.dpos0[3])), rot=np.eye(3))
        return self._run(tip, nz, FEout, FEin, lvec)

    def run_convolveZ(self, FEconv=None, nz=None):
        """convolveZ (relax.cl:571-608) of the last force stack with WZconv;
        returns (nx, ny, nzout, 4) like the reference's FEconv."""
        nx, ny = self.scan_dim[:2]
        nzin = self._nz
        w = self.WZconv
        nzout = nzin - len(w) + 1
        out = np.empty((nx, ny, nzout, 4), dtype=np.float32)
        for c in range(4):
            d = D.convolve_z_dev(self._force, nx * ny, nzin, w, scale=1.0, in_stride=nzin * 4, in_step=4, in_offset=c)
            out[..., c] = d.download().reshape(nx, ny, nzout)
        if FEconv is not None:
            FEconv[...] = out
            return FEconv
        return out

    def downloadPaths(self):
        """Probe positions (nx, ny, nz, 3) with the grid origin added (relax.py:282-305)."""
        return self._paths.download() + np.asarray(self.lvec[0], dtype=np.float32)
