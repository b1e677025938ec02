"""The allocation-free grid build (ppafm_ff_build_noalloc) against
ppafm_ff_build, called through the C ABI: two calls in a row, the second
with fewer atoms (the fixed staging array keeps the first call's records
beyond them), for double4 and float4 grids, both force-field models; the
outputs must be identical.  Too many atoms is a clear error."""

import ctypes

import numpy as np
import pytest

from ppafm_amd import common as C
from ppafm_amd import device as D
from ppafm_amd import molgen as M
from ppafm_amd._lib import HipError, check, dptr, lib

pytestmark = pytest.mark.gpu

N = (48, 40, 56)
CELL = np.diag([4.8, 4.0, 5.6])


def _build(fn, natoms, seed, model, f64):
    xyz, Zs, qs = M.molecule(natoms, seed)
    xyzq = np.ascontiguousarray(np.concatenate([xyz + np.array([2.4, 2.0, 2.0]), qs[:, None]], 1))
    cLJs = np.ascontiguousarray(C.getAtomsLJ(M.PROBE_Z, Zs, C.loadSpecies()))
    out = D.DeviceArray(N + (4,), np.float64 if f64 else np.float32)
    g = D.grid_desc(N, CELL, f64=f64)
    check(fn(ctypes.byref(g), ctypes.byref(model), len(xyzq), dptr(xyzq), dptr(cLJs), out.ptr, None), "grid build")
    D.synchronize()
    return out.download()


@pytest.mark.parametrize("f64", [True, False], ids=["double4", "float4"])
@pytest.mark.parametrize("mode", [0, 1])
def test_noalloc_build_equals_ff_build(f64, mode):
    print(f"case: allocation-free grid build, mode {mode}, f64 {f64}, 41 then 17 atoms", flush=True)
    stencil = D.PROBE_STENCILS["s"] if mode == 0 else ((-10.0, 0.1), (20.0, 0.0), (-10.0, -0.1))
    model = D.ff_model(mode, with_lj=True, stencil=stencil, with_energy=True, probe_q=M.PROBE_Q if mode == 0 else 0.0,
                       fmax=10.0 if mode == 0 else None)
    for natoms, seed in ((41, 3), (17, 4)):
        a = _build(lib().ppafm_ff_build_noalloc, natoms, seed, model, f64)
        b = _build(lib().ppafm_ff_build, natoms, seed, model, f64)
        assert np.isfinite(a).all()
        assert np.array_equal(a, b), natoms


def test_noalloc_build_rejects_too_many_atoms():
    model = D.ff_model(1, with_lj=True, stencil=(), with_energy=True, fmax=None)
    g = D.grid_desc(N, CELL, f64=True)
    out = D.DeviceArray(N + (4,), np.float64)
    n = 65537
    xyzq = np.zeros((n, 4))
    cLJs = np.zeros((n, 2))
    rc = lib().ppafm_ff_build_noalloc(ctypes.byref(g), ctypes.byref(model), n, dptr(xyzq), dptr(cLJs), out.ptr, None)
    assert rc != 0
    assert b"65536" in lib().ppafm_last_error()
    lib().ppafm_clear_error()
    with pytest.raises(HipError):
        check(rc, "grid build")
