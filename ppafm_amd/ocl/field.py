"""Drop-in for the grid side of ``ppafm.ocl.field`` (ocl/field.py:115-843):
``DataGrid`` and its aliases ``HartreePotential``, ``ElectronDensity``,
``TipDensity``, the multipole tip density ``MultipoleTipDensity`` and the
FFT cross-correlation ``FFTCrossCorrelation``, on MI355X through
libppafm_hip.so.  ``calc_force_hartree``, ``calc_force_hartree_tip_charges``
and ``calc_force_fdbm`` assemble the force fields of
``ForceField_LJC.calc_force_hartree`` (field.py:1446-1504),
``run_gradPotentialGrid`` + ``run_evalLJC_Hartree`` (field.py:1168-1317) and
``calc_force_fdbm`` (field.py:1506-1606) on a device grid.

Same names, arguments and errors as the reference, with these differences:
device arrays are ``ppafm_amd.device.DeviceArray`` in FP64 (the reference
keeps pyopencl FP32 buffers), so grids computed on the device come back as
float64 host arrays; grids built from host arrays are rounded to float32 on
construction exactly as the reference does.
"""

import os
import warnings

import numpy as np

from .. import _lib, io
from .. import device as D
from .._lib import check, lib

DEFAULT_FD_STEP = 0.05  # field.py:27

_KINDS = {"s": 0, "px": 1, "py": 2, "pz": 3, "dz2": 4, "dx2": 5, "dy2": 6, "dxy": 7, "dxz": 8, "dyz": 9}


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _lv(lvec):
    return _f64(np.asarray(lvec, dtype=np.float64).reshape(4, 3))


def _n3(shape):
    return np.ascontiguousarray(np.asarray(shape[:3], dtype=np.int32))


class DataGrid:
    """Values on a 3D grid (3D, or 4D with components last), kept on the
    host, the device or both (field.py:115-575).

    Arguments:
        array: np.ndarray (stored as C-contiguous float32, as the reference
            does) or a float64 DeviceArray holding at least prod(shape) values.
        lvec: array-like (4, 3): origin, then the three cell vectors.
        shape: grid shape, required for a DeviceArray.
    """

    def __init__(self, array, lvec, shape=None, ctx=None):
        if isinstance(array, np.ndarray):
            if array.dtype != np.float32 or not array.flags["C_CONTIGUOUS"]:
                array = np.ascontiguousarray(array, dtype=np.float32)
            self.shape = tuple(array.shape)
            self._array = array
            self._dev = None
            self.nbytes = 0
        elif isinstance(array, D.DeviceArray):
            if shape is None:
                raise ValueError("The shape of the grid has to be specified when the array is a DeviceArray.")
            if array.dtype != np.float64:
                raise ValueError(f"device grids are float64, got {array.dtype}")
            nbytes = 8 * int(np.prod(shape))
            assert array.nbytes >= nbytes, f"shape {shape} does not fit into the buffer of size {array.nbytes}"
            self.shape = tuple(int(s) for s in shape)
            self._array = None
            self._dev = array
            self.nbytes = nbytes
        else:
            raise ValueError(f"Invalid type `{type(array)}` for array.")
        if len(self.shape) not in [3, 4]:
            raise ValueError(f"Dimension of array should be 3 or 4, but got {len(self.shape)}")
        self.lvec = np.array(lvec)
        self.origin = self.lvec[0]
        assert self.lvec.shape == (4, 3), f"lvec should have shape (4, 3), but has shape {self.lvec.shape}"
        self.ctx = ctx

    @property
    def step(self):
        """Array of vectors pointing single steps along the grid for each lattice vector."""
        return np.stack([self.lvec[i + 1] / self.shape[i] for i in range(3)])

    @property
    def cell_vol(self):
        """The volume of a grid cell in angstrom^3."""
        a, b, c = self.step
        return abs(np.cross(a, b).dot(c))

    @property
    def array(self):
        """Host array; copied from the device (as float64) if only there."""
        if self._array is None:
            self._array = self._dev.download().reshape(-1)[: int(np.prod(self.shape))].reshape(self.shape)
        return self._array

    @property
    def dev(self):
        """Device array (float64 DeviceArray); uploaded from the host array on first use."""
        if self._dev is None:
            self._dev = D.DeviceArray.from_host(_f64(self._array).reshape(self.shape))
            self.nbytes += 8 * int(np.prod(self.shape))
        return self._dev

    # the reference's name for the device buffer
    cl_array = dev

    @property
    def npts(self):
        return int(np.prod(self.shape[:3]))

    def update_array(self, array, lvec):
        """Replace the values (and lvec); an existing device copy is updated
        (reallocated when the size changes)."""
        if array.dtype != np.float32 or not array.flags["C_CONTIGUOUS"]:
            array = np.ascontiguousarray(array, dtype=np.float32)
        if self._dev is not None:
            if array.size * 8 != self._dev.nbytes:
                self.nbytes += 8 * (array.size - int(np.prod(self.shape)))
                self._dev.free()
                self._dev = D.DeviceArray(array.shape, np.float64)
            self._dev.upload(_f64(array).reshape(self._dev.shape))
        self._array = array
        self.lvec = np.array(lvec)
        self.origin = self.lvec[0]
        self.shape = tuple(array.shape)

    def release(self, keep_on_host=True):
        """Release the device copy (copied to the host first when it is the only copy and keep_on_host)."""
        if self._dev is not None:
            if keep_on_host:
                self.array
            self._dev.free()
            self._dev = None
            self.nbytes -= 8 * int(np.prod(self.shape))

    @classmethod
    def from_file(cls, file_path, scale=1.0):
        """Load grid data and atoms from a .cube or .xsf file (field.py:218-254):
        returns (grid, xyzs (N,3), Zs (N,))."""
        file_path = str(file_path)
        if file_path.endswith(".cube"):
            data, lvec, _, _ = io.loadCUBE(file_path, xyz_order=True)
            Zs, x, y, z, _ = io.loadAtomsCUBE(file_path)
        elif file_path.endswith(".xsf"):
            data, lvec, _, _ = io.loadXSFData(file_path, xyz_order=True)
            try:
                (Zs, x, y, z, _), _, _ = io.loadXSFGeom(file_path)
            except ValueError:
                warnings.warn(f"Could not read geometry from {file_path} in DataGrid.from_file.")
                Zs = np.zeros(1)
                x = y = z = np.zeros((1, 1))
        else:
            raise ValueError(f"Unsupported file format in file `{file_path}`")
        if not np.allclose(scale, 1.0):
            data *= scale
        data = cls(data, lvec)
        xyzs = np.stack([x, y, z], axis=1)
        return data, xyzs, np.array(Zs)

    def to_file(self, file_path, clamp=None):
        """Save as .xsf or .npy (field.py:256-283); a 4D grid is written as a
        vector field plus its "_w" component."""
        file_head, ext = os.path.splitext(file_path)
        if ext not in [".xsf", ".npy"]:
            raise ValueError(f"Unsupported file extension `{ext}` for saving data grid.")
        ext = ext[1:]
        array = np.array(self.array, dtype=np.float64)
        if len(self.shape) == 3:
            if clamp:
                array[array > clamp] = clamp
            io.save_scal_field(file_head, array, self.lvec, data_format=ext)
        if len(self.shape) == 4:
            assert self.shape[3] == 4, "Wrong number of components"
            vec = np.ascontiguousarray(array[..., :3])
            if clamp:
                io.limit_vec_field(vec, Fmax=clamp)
                w = array[..., 3]
                w[w > clamp] = clamp
            io.save_vec_field(file_head, vec, self.lvec, data_format=ext)
            io.save_scal_field(file_head + "_w", np.ascontiguousarray(array[..., 3]), self.lvec, data_format=ext)

    def _out_grid(self, in_place):
        if in_place:
            self._array = None  # the host copy is stale after the operation
            return self
        # the result has the caller's class; a MultipoleTipDensity result is a
        # TipDensity (its constructor builds a density, it does not wrap one:
        # the reference raises TypeError here, field.py:289-297)
        cls = TipDensity if isinstance(self, MultipoleTipDensity) else type(self)
        return cls(D.DeviceArray(self.shape, np.float64), lvec=self.lvec, shape=self.shape, ctx=self.ctx)

    def clamp(self, minimum=-np.inf, maximum=np.inf, clamp_type="hard", soft_clamp_width=1.0, in_place=True, local_size=(32,), queue=None):
        """Clamp values to [minimum, maximum], hard or with a sigmoid (field.py:285-340)."""
        if clamp_type not in ("hard", "soft"):
            raise ValueError(f"Unsupported clamp type `{clamp_type}`")
        src = self.dev
        out = self._out_grid(in_place)
        n = int(np.prod(self.shape))
        check(lib().ppafm_grid_clamp(n, src.ptr, out.dev.ptr, float(minimum), float(maximum), int(clamp_type == "soft"), float(soft_clamp_width), None), "clamp")
        return out

    def add_mult(self, array, scale=1.0, in_place=True, local_size=(32,), queue=None):
        """self + scale * array (field.py:342-375)."""
        a, b = self.dev, array.dev
        out = self._out_grid(in_place)
        n = int(np.prod(self.shape))
        check(lib().ppafm_grid_add_mult(n, a.ptr, b.ptr, float(scale), out.dev.ptr, None), "add_mult")
        return out

    def power_positive(self, p=1.2, normalize=True, in_place=True, local_size=(32,), queue=None):
        """max(0, x)^p, scaled by (sum / positive sum)^p when normalize
        (field.py:377-445)."""
        src = self.dev
        if normalize:
            scale = self._get_normalization_factor()
            assert scale > 0, "Normalizing scaling factor should be positive."
            scale = scale**p
        else:
            scale = 1.0
        out = self._out_grid(in_place)
        n = int(np.prod(self.shape))
        check(lib().ppafm_grid_power(n, src.ptr, out.dev.ptr, float(p), float(scale), None), "power_positive")
        return out

    def _get_normalization_factor(self, queue=None):
        sums = np.zeros(2)
        check(lib().ppafm_grid_sums(int(np.prod(self.shape)), self.dev.ptr, _lib.dptr(sums), None), "grid sums")
        return sums[0] / sums[1]

    def grad(self, scale=1.0, array_out=None, order="C", local_size=(32,), queue=None, accumulate=False):
        """Periodic centred-difference gradient (field.py:471-526): a 4D grid
        (..., 4) = scale * (dV/dx, dV/dy, dV/dz, V), C or Fortran order.
        accumulate=True adds into array_out (no counterpart in the reference)."""
        if len(self.shape) == 4 and self.shape[3] > 1:
            raise RuntimeError("Can only take gradient of a datagrid with a single component.")
        if isinstance(scale, float):
            scale = scale * np.ones(4)
        else:
            scale = np.array(scale, dtype=np.float64)
            if len(scale) != 4:
                raise ValueError(f"Scale should have length 4, but got {len(scale)}.")
        if order not in ["C", "F"]:
            raise ValueError(f"Unknown data order `{order}`")
        shape_out = tuple(self.shape[:3]) + (4,)
        if array_out is None:
            if accumulate:
                raise ValueError("accumulate needs array_out")
            array_out = D.DeviceArray(shape_out, np.float64)
        else:
            assert array_out.nbytes == 4 * 8 * self.npts, f"array size does not match ({array_out.nbytes} != {4 * 8 * self.npts})"
        step = _f64(np.diag(self.step))
        check(lib().ppafm_grid_grad(_lib.iptr(_n3(self.shape)), _lib.dptr(step), self.dev.ptr, _lib.dptr(_f64(scale)), int(order == "C"),
                                    int(bool(accumulate)), array_out.ptr, None), "grad")
        return DataGrid(array_out, lvec=self.lvec, shape=shape_out, ctx=self.ctx)

    def interp_at(self, lvec_new, shape_new, array_out=None, rot=np.eye(3), rot_center=np.zeros(3), local_size=(32,), queue=None):
        """Periodic trilinear interpolation onto another grid whose points are
        rotated by rot about rot_center (field.py:528-575)."""
        if len(self.shape) == 4 and self.shape[3] > 1:
            raise NotImplementedError("Interpolation for 4D grids is not implemented.")
        n_new = _n3(shape_new)
        size_new = 8 * int(np.prod(n_new))
        if array_out is None:
            array_out = D.DeviceArray(tuple(int(v) for v in n_new), np.float64)
        else:
            assert array_out.nbytes == size_new, f"array size does not match ({array_out.nbytes} != {size_new})"
        lv_new = _lv(lvec_new)
        check(lib().ppafm_grid_interp_at(_lib.dptr(_lv(self.lvec)), _lib.iptr(_n3(self.shape)), self.dev.ptr, _lib.dptr(lv_new),
                                         _lib.iptr(n_new), array_out.ptr, _lib.dptr(_f64(rot)), _lib.dptr(_f64(rot_center)), None),
              "interp_at")
        return type(self)(array_out, lvec_new, tuple(int(v) for v in n_new), self.ctx)


class HartreePotential(DataGrid):
    """Sample Hartree potential. Units should be in Volts."""


class ElectronDensity(DataGrid):
    """Sample electron density. Units should be in e/A^3."""


class TipDensity(DataGrid):
    """Tip electron density. Units should be in e/A^3."""

    def subCores(self, xyzs, Zs, Rcore=0.7, valElDict=None):
        """A new TipDensity with the core densities subtracted
        (field.py:586-605): each atom of xyzs (N,3) / Zs, and its periodic
        images within 1 A of the cell, removes its valence electrons
        (valElDict, default HighLevel.loadValenceElectronDict) spread over
        radius Rcore (HighLevel.subtractCoreDensities, on the GPU)."""
        from .. import HighLevel

        array = np.array(self.array, dtype=np.float64, order="C")
        Rs, elems = HighLevel._getAtomsWhichTouchPBCcell(np.asarray(xyzs, dtype=np.float64).T, Zs, self.shape, self.lvec, 1.0, False)
        HighLevel.subtractCoreDensities(array, self.lvec, elems=elems, Rs=Rs, valElDict=valElDict, Rcore=Rcore, bSaveDebugGeom=False)
        return TipDensity(array, self.lvec, ctx=self.ctx)

    def interp_at(self, lvec_new, shape_new, array_out=None, local_size=(32,), queue=None):
        """Interpolate onto a new grid with the tip on the origin of both
        (field.py:607-660): the grid is resized about its middle, the
        origins are ignored, points beyond half the input cell are 0."""
        if len(self.shape) == 4 and self.shape[3] > 1:
            raise NotImplementedError("Interpolation for 4D grids is not implemented.")
        n_new = _n3(shape_new)
        size_new = 8 * int(np.prod(n_new))
        if array_out is None:
            array_out = D.DeviceArray(tuple(int(v) for v in n_new), np.float64)
        else:
            assert array_out.nbytes == size_new, f"array size does not match ({array_out.nbytes} != {size_new})"
        check(lib().ppafm_grid_interp_tip_at(_lib.dptr(_lv(self.lvec)), _lib.iptr(_n3(self.shape)), self.dev.ptr, _lib.dptr(_lv(lvec_new)),
                                             _lib.iptr(n_new), array_out.ptr, None), "interp_tip_at")
        return TipDensity(array_out, lvec_new, tuple(int(v) for v in n_new), self.ctx)


def _linspace_wrapped(length, n, center):
    """The tip grid coordinates along one lattice vector (field.py:680-685)."""
    c = np.linspace(0, length * (1 - 1 / n), n) - center
    c[c >= length / 2] -= length
    c[c <= -length / 2] += length
    return _f64(c)


class MultipoleTipDensity(TipDensity):
    """Multipole probe tip charge density on a periodic grid (field.py:663-705),
    built on the device.

    Arguments:
        lvec: (3, 3) grid lattice vectors.
        shape: grid shape (3 ints).
        center: centre of the density in the grid.
        sigma: width of the Gaussian.
        multipole: dict of weights of 's', 'px', 'py', 'pz', 'dz2', 'dy2',
            'dx2', 'dxy', 'dxz', 'dyz' (other keys contribute 0), or None for
            the bare normalised Gaussian.
        tilt: tilt angle of the density in the zx plane (radians).
    """

    def __init__(self, lvec, shape, center=[0, 0, 0], sigma=0.71, multipole={"dz2": -0.1}, tilt=0.0, ctx=None):
        lvec = np.asarray(lvec, dtype=np.float64)
        n = _n3(shape)
        dev = D.DeviceArray(tuple(int(v) for v in n), np.float64)
        self._make_tip_density(lvec, n, center, sigma, multipole, tilt, dev)
        super().__init__(dev, np.concatenate([[[0, 0, 0]], lvec], axis=0), shape=tuple(int(v) for v in n), ctx=ctx)

    @staticmethod
    def _make_tip_density(lvec, n, center, sigma, multipole, tilt, out):
        lvec_len = np.linalg.norm(lvec, axis=1)
        center = np.array(center)
        if (center < 0).any() or (center > lvec.sum(axis=0)).any():
            raise ValueError("Center position is outside the grid.")
        cs = [_linspace_wrapped(lvec_len[i], int(n[i]), center[i]) for i in range(3)]
        # getNormalizedBasisMatrix(lvec).getT() (fieldFFT.py:155-163, 213-217)
        Lmat = np.stack([lvec[i] / np.linalg.norm(lvec[i]) for i in range(3)])
        mat = _f64(Lmat.T)
        step = lvec_len / n
        norm = _f64([abs(np.linalg.det(mat)), step[0], step[1], step[2]])
        if multipole is None:
            kinds, coefs, nk = np.zeros(1, np.int32), np.zeros(1), -1
        else:
            kinds = np.array([_KINDS.get(k, -1) for k in multipole], dtype=np.int32)
            coefs = _f64([float(v) for v in multipole.values()])
            nk = len(kinds)
            if nk == 0:
                kinds, coefs = np.zeros(1, np.int32), np.zeros(1)
        check(lib().ppafm_tip_density_multipole(_lib.iptr(n), _lib.dptr(cs[0]), _lib.dptr(cs[1]), _lib.dptr(cs[2]), _lib.dptr(mat),
                                                float(sigma), float(tilt), int(nk), _lib.iptr(kinds), _lib.dptr(coefs), _lib.dptr(norm),
                                                out.ptr, None), "MultipoleTipDensity")


class FFTCrossCorrelation:
    """Circular cross-correlation of grids with a tip density via FFT
    (field.py:707-843): E[r] = scale * cell_vol * sum_s rho[s] in[s + r].

    Arguments:
        rho: TipDensity. Tip charge density (3D).
    """

    def __init__(self, rho, queue=None):
        self.shape = tuple(rho.shape[:3])
        p = _lib.c_void_p()
        check(lib().ppafm_fft_corr_create(_lib.iptr(_n3(self.shape)), _lib.ctypes.byref(p)), "FFTCrossCorrelation")
        self._h = p.value
        self._lib = lib()
        self._set_rho(rho)

    def _set_rho(self, rho):
        if tuple(rho.shape[:3]) != self.shape:
            raise ValueError(f"rho shape {rho.shape} does not match {self.shape}")
        self.rho = rho
        check(lib().ppafm_fft_corr_set_rho(self._h, rho.dev.ptr, float(rho.cell_vol), None), "FFTCrossCorrelation._set_rho")

    def correlate(self, array, E=None, scale=1):
        """Cross-correlate `array` (DataGrid or float64 DeviceArray of the rho
        shape) with the tip density; E (DataGrid) receives the result, else a
        new DataGrid with rho's lvec."""
        src = array.dev if isinstance(array, DataGrid) else array
        assert src.nbytes == 8 * int(np.prod(self.shape)), f"array size {src.nbytes} does not match rho array shape {self.shape}"
        if E is None:
            E = DataGrid(D.DeviceArray(self.shape, np.float64), lvec=self.rho.lvec, shape=self.shape)
        else:
            assert E.shape == self.shape, "E data grid shape does not match"
            E._array = None
        check(lib().ppafm_fft_corr_correlate(self._h, src.ptr, float(scale), E.dev.ptr, None), "FFTCrossCorrelation.correlate")
        return E

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                self._lib.ppafm_fft_corr_destroy(self._h)
                self._h = None
        except Exception:
            pass


# ------------------------------------------------------- force-field assembly


def _check_rectangular(lvec_ff):
    cell = np.asarray(lvec_ff)[1:]
    if (cell != np.diag(np.diag(cell))).any():
        raise NotImplementedError(
            "Forcefield calculation via FFT for non-rectangular grids is not implemented. "
            "Note that the forcefield grid does not need to match the Hartree potential grid."
        )


class Workspace:
    """Device scratch grids of one force-field shape, allocated on first
    use and reused by every later call (no counterpart in the reference,
    which allocates per call)."""

    def __init__(self, n):
        self.n = tuple(int(v) for v in n[:3])
        self._b = {}

    def buf(self, key, comps=1):
        shape = self.n + ((comps,) if comps > 1 else ())
        b = self._b.get(key)
        if b is None or b.shape != shape:
            b = self._b[key] = D.DeviceArray(shape, np.float64)
        return b

    def grid(self, key, lvec, cls=None):
        return (cls or DataGrid)(self.buf(key), lvec, shape=self.n)


def calc_force_hartree(grid4, lvec_ff, n, pot, fft_corr, rot=np.eye(3), rot_center=np.zeros(3), ws=None):
    """ForceField_LJC.calc_force_hartree (field.py:1446-1504) on a double4
    force-field grid (n, 4) that already holds the Lennard-Jones force and
    energy: the potential interpolated onto the force-field grid (points
    rotated by rot about rot_center), cross-correlated with the tip
    density, and its gradient (-dE/dx, -dE/dy, -dE/dz, E) added.  ws: a
    Workspace of shape n for the intermediate grid."""
    _check_rectangular(lvec_ff)
    ws = ws or Workspace(n)
    pot_i = pot.interp_at(lvec_ff, n, array_out=ws.buf("a"), rot=rot, rot_center=rot_center)
    E = fft_corr.correlate(pot_i, E=pot_i)
    E.grad(scale=[-1.0, -1.0, -1.0, 1.0], array_out=grid4, order="C", accumulate=True)
    return grid4


def _matching_grid(pot, lvec_ff, n):
    # field.py:1268-1273
    cell = np.asarray(lvec_ff)[1:]
    return (
        np.allclose(pot.shape, n[:3])
        and (np.abs(pot.origin - np.asarray(lvec_ff)[0]) < 1e-3).all()
        and (np.abs(np.diag(pot.step) * pot.shape[:3] - np.diag(cell)) < 1e-3).all()
        and (pot.step == np.diag(np.diag(pot.step))).all()
    )


def calc_force_hartree_tip_charges(grid4, lvec_ff, n, pot, Qs, QZs, h=None, ws=None):
    """The Hartree method with point-charge tips: run_gradPotentialGrid
    (field.py:1225-1317) then the charge terms of run_evalLJC_Hartree
    (field.py:1168-1223) added to a double4 grid holding the Lennard-Jones
    force.  Returns the field grid (n, 4)."""
    h = h or DEFAULT_FD_STEP
    n = tuple(int(v) for v in n[:3])
    ws = ws or Workspace(n)
    field = ws.buf("field", 4)
    if _matching_grid(pot, lvec_ff, n):
        pot.grad(scale=[-1.0, -1.0, -1.0, 1.0], array_out=field, order="C")
    else:
        check(lib().ppafm_grid_grad_potential(_lib.dptr(_lv(pot.lvec)), _lib.iptr(_n3(pot.shape)), pot.dev.ptr, _lib.dptr(_lv(lvec_ff)),
                                              _lib.iptr(_n3(n)), float(h), field.ptr, None), "gradPotentialGrid")
    check(lib().ppafm_grid_add_tip_charges(_lib.dptr(_lv(lvec_ff)), _lib.iptr(_n3(n)), field.ptr, _lib.dptr(_f64(Qs)), _lib.dptr(_f64(QZs)),
                                           grid4.ptr, None), "evalLJC_Hartree")
    return field


def calc_force_fdbm(grid4, lvec_ff, n, pot, rho_sample, fft_corr, fft_corr_delta, A=18.0, B=1.0, rot=np.eye(3), rot_center=np.zeros(3),
                    ws=None):
    """The density part of ForceField_LJC.calc_force_fdbm (field.py:1506-1606)
    written to a double4 grid (n, 4): electrostatics of the potential with
    the tip delta density (scale -1), Pauli repulsion A * (sample density^B
    correlated with the tip density), gradient with scale (-1, -1, -1, 1).
    The vdW term is added by the caller (add_vdw / add_d3)."""
    _check_rectangular(lvec_ff)
    lvec_ff = np.asarray(lvec_ff)
    ws = ws or Workspace(n)
    pot_same = np.allclose(lvec_ff, pot.lvec) and np.allclose(pot.shape, n[:3])
    rs_same = np.allclose(lvec_ff, rho_sample.lvec) and np.allclose(rho_sample.shape, n[:3])
    pot_i = pot if pot_same else pot.interp_at(lvec_ff, n, array_out=ws.buf("a"), rot=rot, rot_center=rot_center)
    E = fft_corr_delta.correlate(pot_i, E=ws.grid("b", lvec_ff), scale=-1.0)
    rs = rho_sample if rs_same else rho_sample.interp_at(lvec_ff, n, array_out=ws.buf("a"), rot=rot, rot_center=rot_center)
    if not np.allclose(B, 1.0):
        # power_positive(p=B, in_place=False) into scratch
        scale = rs._get_normalization_factor()
        assert scale > 0, "Normalizing scaling factor should be positive."
        out = ws.grid("c", lvec_ff)
        check(lib().ppafm_grid_power(rs.npts, rs.dev.ptr, out.dev.ptr, float(B), float(scale**B), None), "power_positive")
        rs = out
    E_pauli = fft_corr.correlate(rs, E=ws.grid("c", lvec_ff))  # may alias its input
    E = E.add_mult(E_pauli, scale=A, in_place=True)
    E.grad(scale=[-1.0, -1.0, -1.0, 1.0], array_out=grid4, order="C")
    return grid4


def add_vdw(grid4, n, cell, xyz_rel, coefs, damp_method):
    """addVdW (field.py:1341-1366, FF.cl:548-665) onto a double4 grid: atoms
    xyz_rel (N,3) relative to the grid origin, coefs (N,2) = (c6, c12) for
    damp_method <= 0, else (R, E0)."""
    xyz = D.DeviceArray.from_host(_f64(xyz_rel).reshape(-1, 3) if len(xyz_rel) else np.zeros((1, 3)))
    cf = D.DeviceArray.from_host(_f64(coefs).reshape(-1, 2) if len(coefs) else np.zeros((1, 2)))
    g = D.grid_desc(n, cell, f64=True)
    check(lib().ppafm_grid_add_vdw(_lib.ctypes.byref(g), int(damp_method), len(xyz_rel), xyz.ptr, cf.ptr, grid4.ptr, None), "addvdW")
    D.synchronize()


def add_d3(grid4, n, cell, xyz_rel, coeffs):
    """addDFTD3_BJ (field.py:1396-1444, FF.cl:790-857) onto a double4 grid;
    coeffs (N,4) from computeD3Coeffs."""
    xyz = D.DeviceArray.from_host(_f64(xyz_rel).reshape(-1, 3) if len(xyz_rel) else np.zeros((1, 3)))
    cf = D.DeviceArray.from_host(_f64(coeffs).reshape(-1, 4) if len(coeffs) else np.zeros((1, 4)))
    g = D.grid_desc(n, cell, f64=True)
    check(lib().ppafm_grid_add_d3(_lib.ctypes.byref(g), len(xyz_rel), xyz.ptr, cf.ptr, grid4.ptr, None), "addDFTD3_BJ")
    D.synchronize()


# ------------------------------------------------------------------ DFT-D3

_D3 = {"tables": None, "functionals": {}}


def set_d3_tables(R_COV, REF_CN, REF_C6, R4R2, K=(16.0, 4.0 / 3.0, 4.0), functionals=None):
    """Register the Grimme-D3 tables used by the density-based force field's
    DFT-D3 vdW term.  ppafm reads them from ppafm.defaults.d3 (R_COV (94,),
    REF_CN (94,5), load_ref_c6() (94,94,5,5), R4R2 (94,), K1..K3, and the
    functional table DF_DEFAULT_PARAMS); this package takes them from the
    caller.  functionals: {name: {'s6', 's8', 'a1', 'a2' (bohr)}} for
    d3_params given by name.  K defaults to the published k1 = 16,
    k2 = 4/3, k3 = 4 of Grimme et al. (2010)."""
    _D3["tables"] = dict(R_COV=_f64(R_COV), REF_CN=_f64(REF_CN), REF_C6=_f64(REF_C6), R4R2=_f64(R4R2), K=_f64(K))
    _D3["functionals"] = dict(functionals or {})


def get_df_params(params):
    """Functional parameters (d3.get_df_params, defaults/d3.py:278-286): a
    dict with s6, s8, a1, a2, or the name of a registered functional."""
    if isinstance(params, str):
        if params not in _D3["functionals"]:
            raise ValueError(f"Default parameters not available for functional `{params}`. Register them with set_d3_tables(functionals=...).")
        params = _D3["functionals"][params]
    for key in ["s6", "s8", "a1", "a2"]:
        if key not in params:
            raise ValueError(f"DFT-D3 parameter dictionary is missing entry for `{key}`")
    return params


def d3_coeffs(xyzs, Zs, iZPP, params):
    """Per-atom D3 coefficients (c6 s6, c8 s8, R0^6, R0^8) as
    ForceField_LJC._get_dftd3_params computes them (field.py:1368-1394,
    kernel d3_coeffs FF.cl:667-788): a2 converted from bohr to angstrom."""
    from .. import core

    if _D3["tables"] is None:
        raise RuntimeError("DFT-D3 tables are not set; register them with ppafm_amd.ocl.field.set_d3_tables")
    p = get_df_params(params)
    p = {"s6": p["s6"], "s8": p["s8"], "a1": p["a1"], "a2": p["a2"] * io.bohrRadius2angstroem}
    return core.computeD3Coeffs(_f64(xyzs).reshape(-1, 3), np.asarray(Zs, dtype=np.int32), int(iZPP), p, _D3["tables"])
