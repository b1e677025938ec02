"""Device buffers and thin wrappers over the device entry points of
libppafm_hip.so.  All compute runs in the HIP kernels; this module only
moves arrays and fills the descriptors declared in include/ppafm_hip.h."""

import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib

COULOMB_CONST_CPU = 14.399645  # common.py:21 (ProbeParticle semantics)
COULOMB_CONST_OCL = 14.399644  # FF.cl:2 (OpenCL semantics)


class DeviceArray:
    """A device allocation with numpy-like shape/dtype bookkeeping."""

    def __init__(self, shape, dtype=np.float32):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape)) if not isinstance(shape, tuple) else tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(lib().ppafm_malloc(ctypes.byref(p), self.nbytes), "ppafm_malloc")
        self.ptr = p.value
        self._lib = lib()

    @classmethod
    def from_host(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a, stream=None):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.shape, self.shape)
        check(lib().ppafm_memcpy_h2d(self.ptr, a.ctypes.data, a.nbytes, stream), "upload")

    def download(self, out=None, stream=None):
        if out is None:
            out = np.empty(self.shape, self.dtype)
        assert out.nbytes == self.nbytes and out.flags.c_contiguous
        check(lib().ppafm_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes, stream), "download")
        return out

    def zero(self, stream=None):
        check(lib().ppafm_memset(self.ptr, 0, self.nbytes, stream), "memset")

    def free(self):
        if getattr(self, "ptr", None):
            self._lib.ppafm_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize(stream=None):
    check(lib().ppafm_synchronize(stream), "synchronize")


PEAK_KINDS = {"fp32_fma": (0, 500_000), "fp64_fma": (1, 500_000), "bf16_mfma": (2, 250_000)}


def measure_peak(kind):
    """Measured peak TFLOP/s of the current device for ``kind`` in
    PEAK_KINDS (microbenchmark kernels of the library)."""
    k, iters = PEAK_KINDS[kind]
    v = ctypes.c_double()
    check(lib().ppafm_measure_peak(k, iters, ctypes.byref(v)), "measure_peak")
    return float(v.value)


def device_count():
    return lib().ppafm_device_count()


class Event:
    """HIP event (timing on the stream the kernels run on)."""

    def __init__(self):
        p = ctypes.c_void_p()
        check(lib().ppafm_event_create(ctypes.byref(p)), "event_create")
        self.ptr = p.value
        self._lib = lib()

    def record(self, stream=None):
        check(lib().ppafm_event_record(self.ptr, stream), "event_record")

    def elapsed_ms(self, end):
        ms = ctypes.c_float()
        check(lib().ppafm_event_elapsed_ms(self.ptr, end.ptr, ctypes.byref(ms)), "event_elapsed")
        return float(ms.value)

    def __del__(self):
        try:
            self._lib.ppafm_event_destroy(self.ptr)
        except Exception:
            pass


# ---------------------------------------------------------------- descriptors


def grid_desc(n, cell, f64=False):
    g = _lib.Grid()
    g.f64 = int(bool(f64))
    for k in range(3):
        g.n[k] = int(n[k])
    c = np.asarray(cell, dtype=np.float64).reshape(3, 3)
    for k in range(9):
        g.cell[k] = float(c.flat[k])
    return g


# [A] float4 grids of the ProbeParticle model: pairs closer than this are
# summed in FP64 (the grid then meets 1e-5 relative); the OpenCL model's
# float4 grids are FP32 throughout (0), as the reference's own
FP64_RADIUS = {0: 4.0, 1: 0.0}


def ff_model(mode, with_lj=True, stencil=(), with_energy=False, coulomb_const=None, probe_q=0.0, fmax=10.0, fp64_radius=None):
    """stencil: sequence of (weight, dz) Coulomb evaluation points."""
    m = _lib.FFModel()
    m.mode = int(mode)
    m.with_lj = int(bool(with_lj))
    st = [(float(w), float(z)) for (w, z) in stencil if w != 0]
    if len(st) > 4:
        raise ValueError("at most 4 Coulomb stencil points")
    m.n_stencil = len(st)
    for k, (w, z) in enumerate(st):
        m.stencil_w[k] = w
        m.stencil_dz[k] = z
    m.with_energy = int(bool(with_energy))
    if coulomb_const is None:
        coulomb_const = COULOMB_CONST_CPU if mode == 0 else COULOMB_CONST_OCL
    m.coulomb_const = float(coulomb_const)
    m.probe_q = float(probe_q)
    m.fmax = float(fmax) if fmax is not None else 0.0
    m.fp64_radius = float(FP64_RADIUS[int(mode)] if fp64_radius is None else fp64_radius)
    return m


# CPU probe kinds (ProbeParticle.cpp:193-209): s, pz, dz2 as Coulomb stencils
PROBE_STENCILS = {
    "s": ((1.0, 0.0),),
    "pz": ((-10.0, -0.1), (10.0, 0.1)),
    "dz2": ((-200.0, 0.0), (100.0, -0.1), (100.0, 0.1)),
}


def relax_params(mode=0, alg=1, max_iters=1000, probe_start=1, conv_f2=1e-8, dt=0.1, damping=0.1, finc=1.1, fdec=0.5, falpha=0.99, dtmin_frac=0.0, f2safe=0.0):
    p = _lib.RelaxParams()
    p.mode, p.alg, p.max_iters, p.probe_start = int(mode), int(alg), int(max_iters), int(probe_start)
    p.conv_f2, p.dt, p.damping = float(conv_f2), float(dt), float(damping)
    p.finc, p.fdec, p.falpha = float(finc), float(fdec), float(falpha)
    p.dtmin_frac, p.f2safe = float(dtmin_frac), float(f2safe)
    return p


def tip_desc(d0, kspring, krad, lrad, rot=None, surf_ff=0.0):
    t = _lib.Tip()
    R = np.eye(3) if rot is None else np.asarray(rot, dtype=np.float64).reshape(3, 3)
    for k in range(9):
        t.rot[k] = float(R.flat[k])
    for k in range(3):
        t.d0[k] = float(d0[k])
        t.kspring[k] = float(kspring[k])
    t.krad, t.lrad, t.surf_ff = float(krad), float(lrad), float(surf_ff)
    return t


def scan_desc(nx, ny, nz, p0=(0, 0, 0), da=(0, 0, 0), db=(0, 0, 0), dstep=(0, 0, 0)):
    s = _lib.Scan()
    s.nx, s.ny, s.nz = int(nx), int(ny), int(nz)
    for k in range(3):
        s.p0[k], s.da[k], s.db[k], s.dstep[k] = float(p0[k]), float(da[k]), float(db[k]), float(dstep[k])
    return s


def cpu_tip(lrad, krad, rpp0_xy, kspring):
    """ProbeParticle tip (setTip + makeConsistent) as a descriptor: rPP0.z on the sphere."""
    x, y = float(rpp0_xy[0]), float(rpp0_xy[1])
    z = -np.sqrt(lrad * lrad - x * x - y * y) if abs(krad) > 1e-8 else float(rpp0_xy[2]) if len(rpp0_xy) > 2 else 0.0
    return tip_desc((x, y, z), kspring, krad, lrad)


# ------------------------------------------------------------------ kernels


def pack_atoms(xyzq, cLJs, coulomb_const):
    """Pack (N,4) positions+charges (grid frame) and (N,2) c6/c12 into the
    device atom record layout (host side)."""
    xyzq = np.ascontiguousarray(xyzq, dtype=np.float64).reshape(-1, 4)
    n = len(xyzq)
    out = np.zeros((max(n, 1), 6), dtype=np.float64)
    cl = None if cLJs is None else np.ascontiguousarray(cLJs, dtype=np.float64).reshape(-1, 2)
    check(lib().ppafm_pack_atoms(n, _lib.dptr(xyzq), _lib.dptr(cl), float(coulomb_const), out.ctypes.data), "pack_atoms")
    return out[:n]


def ff_build(n, cell, xyzq, cLJs, model, out=None, stream=None, atoms_dev=None, f64=False):
    """Build the (nx,ny,nz,4) grid on the device: float32, or float64 when f64."""
    n = tuple(int(v) for v in n)
    if out is None:
        out = DeviceArray(n + (4,), np.float64 if f64 else np.float32)
    g = grid_desc(n, cell, f64=out.dtype == np.float64)
    if atoms_dev is None:
        packed = pack_atoms(xyzq, cLJs, model.coulomb_const)
        atoms_dev = DeviceArray.from_host(packed if len(packed) else np.zeros((1, 6)))
        natoms = len(packed)
    else:
        natoms = atoms_dev.shape[0]
    check(lib().ppafm_ff_build_dev(ctypes.byref(g), ctypes.byref(model), natoms, atoms_dev.ptr, out.ptr, stream), "ff_build")
    out._atoms = atoms_dev  # keep alive until the kernel has run
    return out


def relax(n, cell, grid_dev, rp, tip, scan, tips_dev=None, pos=None, force=None, iters=None, counters=None, pos_f32=True,
          stream=None, status=None):
    """Run the relaxation kernel; outputs are DeviceArrays (allocated if None
    and requested with True).  Returns (pos, force, iters), and the uint8
    per-position status array as a fourth item when ``status`` is given
    (0 converged, bit 0 the iteration cap ended the step, bit 1 non-finite)."""
    shp = (scan.nx, scan.ny, scan.nz)
    if pos is True:
        pos = DeviceArray(shp + (3,), np.float32 if pos_f32 else np.float64)
    if force is True:
        force = DeviceArray(shp + (4,), np.float32)
    if iters is True:
        iters = DeviceArray(shp, np.int32)
    if status is True:
        status = DeviceArray(shp, np.uint8)
    o = _lib.RelaxOut()
    o.pos = pos.ptr if pos is not None else None
    o.pos_f32 = int(pos_f32)
    o.force = force.ptr if force is not None else None
    o.iters = iters.ptr if iters is not None else None
    o.counters = counters.ptr if counters is not None else None
    o.status = status.ptr if status is not None else None
    g = grid_desc(n, cell, f64=grid_dev.dtype == np.float64)
    check(
        lib().ppafm_relax(ctypes.byref(g), grid_dev.ptr, ctypes.byref(rp), ctypes.byref(tip), ctypes.byref(scan), tips_dev.ptr if tips_dev is not None else None, ctypes.byref(o), stream),
        "relax",
    )
    if status is not None:
        return pos, force, iters, status
    return pos, force, iters


def stiffness(n, cell, grid_dev, tip, scan, pos, which=0, ddisp=0.05, tips_dev=None, out=None, stream=None):
    """stiffnessMatrix on a device grid at the relaxed positions ``pos`` (a
    DeviceArray (nx,ny,nz,3), float32 or float64, as relax() writes it):
    returns DeviceArrays eigenvalues (nx,ny,nz,3) ascending and a list of
    three eigenvector arrays, the first ``which`` allocated (others None).
    ``out`` = (evals, evecs) from an earlier call is reused instead."""
    shp = (scan.nx, scan.ny, scan.nz)
    if tuple(pos.shape) != shp + (3,) or pos.dtype not in (np.float32, np.float64):
        raise ValueError(f"pos must be a float32/float64 DeviceArray of shape {shp + (3,)}, got {pos.shape} {pos.dtype}")
    if out is None:
        evals = DeviceArray(shp + (3,), np.float64)
        evecs = [DeviceArray(shp + (3,), np.float64) if k < which else None for k in range(3)]
    else:
        evals, evecs = out
        if tuple(evals.shape) != shp + (3,) or any((v is None) != (k >= which) for k, v in enumerate(evecs)):
            raise ValueError("out does not match the scan shape and `which`")
    g = grid_desc(n, cell, f64=grid_dev.dtype == np.float64)
    check(
        lib().ppafm_stiffness(ctypes.byref(g), grid_dev.ptr, ctypes.byref(tip), ctypes.byref(scan),
                              tips_dev.ptr if tips_dev is not None else None, pos.ptr, int(pos.dtype == np.float32),
                              float(ddisp), int(which), evals.ptr, *[v.ptr if v is not None else None for v in evecs], stream),
        "stiffness",
    )
    return evals, evecs


def convolve_z_dev(in_dev, ncol, nz, w, scale=1.0, in_stride=None, in_step=1, in_offset=0, out=None, f64=False, stream=None):
    w = np.ascontiguousarray(w, dtype=np.float64)
    nzout = nz - len(w) + 1
    if in_stride is None:
        in_stride = nz * in_step
    if out is None:
        out = DeviceArray((ncol, nzout), np.float64 if f64 else np.float32)
    check(
        lib().ppafm_convolve_z(int(ncol), int(nz), int(in_stride), int(in_step), int(in_offset), in_dev.ptr, len(w), _lib.dptr(w), float(scale), out.ptr, int(bool(f64)), stream),
        "convolve_z",
    )
    return out


def convolve_z(F, w, scale=1.0):
    """Host float64 array in, host float64 array out; correlation along the
    last axis on the GPU: out[..., j] = scale * sum_k F[..., j+k] w[k]."""
    F = np.ascontiguousarray(F, dtype=np.float64)
    nz = F.shape[-1]
    ncol = F.size // nz if nz else 0
    dF = DeviceArray.from_host(F.reshape(ncol, nz) if ncol else np.zeros((1, nz)))
    out = convolve_z_dev(dF, ncol, nz, w, scale=scale, f64=True)
    return out.download().reshape(F.shape[:-1] + (nz - len(w) + 1,))
