"""ctypes binding of libppafm_hip.so (declarations in include/ppafm_hip.h).

The library is the product path: importing this module fails loudly when it
has not been built (``python -m ppafm_amd.build``); nothing falls back to CPU.
"""

import ctypes
import pathlib

import numpy as np

LIB_PATH = pathlib.Path(__file__).resolve().parent / "libppafm_hip.so"

c_int, c_double, c_long, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_long, ctypes.c_void_p
c_ull = ctypes.c_ulonglong
_dp = ctypes.POINTER(c_double)
_ip = ctypes.POINTER(c_int)


class SplineParams(ctypes.Structure):
    _fields_ = [("rff_n", c_int), ("rff_xs", _dp), ("rff_ydys", _dp)]


class Grid(ctypes.Structure):
    _fields_ = [("n", c_int * 3), ("cell", c_double * 9), ("f64", c_int)]


class FFModel(ctypes.Structure):
    _fields_ = [
        ("mode", c_int),
        ("with_lj", c_int),
        ("n_stencil", c_int),
        ("with_energy", c_int),
        ("stencil_w", c_double * 4),
        ("stencil_dz", c_double * 4),
        ("coulomb_const", c_double),
        ("probe_q", c_double),
        ("fmax", c_double),
        ("fp64_radius", c_double),
    ]


class Atom(ctypes.Structure):
    _fields_ = [("x", c_double), ("y", c_double), ("z", c_double), ("q", c_double), ("c6", c_double), ("c12", c_double)]


class RelaxParams(ctypes.Structure):
    _fields_ = [
        ("mode", c_int),
        ("alg", c_int),
        ("max_iters", c_int),
        ("probe_start", c_int),
        ("conv_f2", c_double),
        ("dt", c_double),
        ("damping", c_double),
        ("finc", c_double),
        ("fdec", c_double),
        ("falpha", c_double),
        ("dtmin_frac", c_double),
        ("f2safe", c_double),
    ]


class Tip(ctypes.Structure):
    _fields_ = [("rot", c_double * 9), ("d0", c_double * 3), ("kspring", c_double * 3), ("krad", c_double), ("lrad", c_double), ("surf_ff", c_double)]


class Scan(ctypes.Structure):
    _fields_ = [("nx", c_int), ("ny", c_int), ("nz", c_int), ("p0", c_double * 3), ("da", c_double * 3), ("db", c_double * 3), ("dstep", c_double * 3)]


class RelaxOut(ctypes.Structure):
    _fields_ = [("pos", c_void_p), ("pos_f32", c_int), ("force", c_void_p), ("iters", c_void_p), ("counters", c_void_p),
                ("status", c_void_p)]


class BatchItem(ctypes.Structure):
    _fields_ = [("n", c_int * 3), ("cell", c_double * 9), ("natoms", c_int), ("scan", Scan)]


class BatchTuning(ctypes.Structure):
    _fields_ = [("tile_y", c_int), ("waves_per_block", c_int), ("blocks_per_cu", c_int), ("refill", c_int)]

    def as_dict(self):
        return {"tile_y": self.tile_y, "waves_per_block": self.waves_per_block, "blocks_per_cu": self.blocks_per_cu,
                "refill": self.refill}


class WarmParams(ctypes.Structure):
    _fields_ = [("enable", c_int), ("net_passes", c_int), ("newton_max", c_int), ("newton_max0", c_int),
                ("newton_retry", c_int), ("newton_tol", c_double), ("newton_cap", c_double), ("extrap_max", c_double),
                ("residual_max", c_double)]


# (name, restype, argtypes) of every exported symbol declared in the header
SIGNATURES = [
    ("setGridN", None, [_ip]),
    ("setGridCell", None, [_dp]),
    ("setFF_Fpointer", None, [_dp]),
    ("setFF_Epointer", None, [_dp]),
    ("deleteFF_Fpointer", None, []),
    ("deleteFF_Epointer", None, []),
    ("setTip", None, [c_double, c_double, _dp, _dp]),
    ("setRelax", None, [c_int, c_double, c_double, c_double]),
    ("setFIRE", None, [c_double, c_double, c_double]),
    ("getLennardJonesFF", None, [c_int, _dp, _dp]),
    ("getCoulombFF", None, [c_int, _dp, _dp, c_int]),
    ("relaxTipStroke", c_int, [c_int, c_int, c_int, _dp, _dp, _dp, ctypes.POINTER(SplineParams)]),
    ("relaxTipStrokes_omp", c_int, [c_int, c_int, c_int, c_int, c_int, _dp, _dp, _dp, ctypes.POINTER(SplineParams)]),
    ("stiffnessMatrix", None, [c_double, c_int, c_int, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.POINTER(SplineParams)]),
    ("getVdWFF", None, [c_int, _dp, _dp]),
    ("getVdWFF_RE", None, [c_int, _dp, _dp, c_int, c_double]),
    ("getDFTD3FF", None, [c_int, _dp, _dp]),
    ("getMorseFF", None, [c_int, _dp, _dp, c_double]),
    ("getGaussDensity", None, [c_int, _dp, _dp]),
    ("getSlaterDensity", None, [c_int, _dp, _dp]),
    ("getDensityR4spline", None, [c_int, _dp, _dp]),
    ("getInPoints_LJ", None, [c_int, _dp, _dp, c_int, _dp, _dp]),
    ("evalRadialFF", None, [c_int, _dp, _dp, _dp, _dp, c_int, c_double]),
    ("computeD3Coeffs", None, [c_int, _dp, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp, c_int, _dp]),
    ("subsample_uniform_spline", None, [c_double, c_double, c_int, _dp, c_int, _dp, _dp]),
    ("subsample_nonuniform_spline", None, [c_int, _dp, _dp, c_int, _dp, _dp]),
    ("test_force", None, [c_int, c_int, _dp, _dp, _dp, _dp, ctypes.POINTER(SplineParams)]),
    ("test_eigen3x3", None, [_dp, _dp]),
    ("ppafm_pack_atoms", c_int, [c_int, _dp, _dp, c_double, c_void_p]),
    ("ppafm_ff_build_dev", c_int, [ctypes.POINTER(Grid), ctypes.POINTER(FFModel), c_int, c_void_p, c_void_p, c_void_p]),
    ("ppafm_ff_build", c_int, [ctypes.POINTER(Grid), ctypes.POINTER(FFModel), c_int, _dp, _dp, c_void_p, c_void_p]),
    ("ppafm_ff_build_noalloc", c_int, [ctypes.POINTER(Grid), ctypes.POINTER(FFModel), c_int, _dp, _dp, c_void_p, c_void_p]),
    (
        "ppafm_relax",
        c_int,
        [ctypes.POINTER(Grid), c_void_p, ctypes.POINTER(RelaxParams), ctypes.POINTER(Tip), ctypes.POINTER(Scan), c_void_p, ctypes.POINTER(RelaxOut), c_void_p],
    ),
    (
        "ppafm_stiffness",
        c_int,
        [ctypes.POINTER(Grid), c_void_p, ctypes.POINTER(Tip), ctypes.POINTER(Scan), c_void_p, c_void_p, c_int, c_double, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    ),
    ("ppafm_convolve_z", c_int, [c_long, c_int, c_int, c_int, c_int, c_void_p, c_int, _dp, c_double, c_void_p, c_int, c_void_p]),
    ("ppafm_limit_vec_field", c_int, [c_void_p, c_long, c_double, c_void_p]),
    ("ppafm_copy_kernel", c_int, [c_void_p, c_void_p, c_ull, c_void_p]),
    (
        "ppafm_batch_create",
        c_void_p,
        [c_int, ctypes.POINTER(BatchItem), ctypes.POINTER(FFModel), ctypes.POINTER(RelaxParams), ctypes.POINTER(Tip), c_int, c_int, _dp, c_double, c_int, ctypes.POINTER(BatchTuning), c_void_p],
    ),
    ("ppafm_batch_set_atoms", c_int, [c_void_p, c_void_p]),
    ("ppafm_batch_set_atoms_n", c_int, [c_void_p, c_void_p, c_void_p]),
    ("ppafm_batch_run", c_int, [c_void_p]),
    ("ppafm_batch_run_stages", c_int, [c_void_p, c_int]),
    ("ppafm_batch_tune", c_int, [c_void_p, c_int, c_int, ctypes.POINTER(BatchTuning), _dp]),
    ("ppafm_batch_set_tuning", c_int, [c_void_p, ctypes.POINTER(BatchTuning)]),
    ("ppafm_batch_get_tuning", c_int, [c_void_p, ctypes.POINTER(BatchTuning)]),
    ("ppafm_batch_buffer", c_void_p, [c_void_p, c_int]),
    ("ppafm_batch_offsets", c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    ("ppafm_batch_stream", c_void_p, [c_void_p]),
    ("ppafm_batch_destroy", None, [c_void_p]),
    ("ppafm_device_info", c_int, [ctypes.c_char_p, c_int, _ip]),
    ("ppafm_batch_set_warmstart", c_int, [c_void_p, c_void_p, c_void_p]),
    ("ppafm_measure_peak", c_int, [c_int, c_int, ctypes.POINTER(c_double)]),
    ("ppafm_batch_polish", c_int, [c_void_p, c_double, c_int, c_void_p, c_void_p]),
    ("ppafm_batch_ws_samples", c_int, [c_void_p, c_void_p, ctypes.c_longlong, c_void_p, c_void_p, c_void_p, c_double, ctypes.c_uint]),
    ("ppafm_ws_create", c_void_p, [c_int]),
    ("ppafm_ws_destroy", None, [c_void_p]),
    ("ppafm_ws_nparams", c_int, []),
    ("ppafm_ws_set_params", c_int, [c_void_p, c_void_p, c_int]),
    ("ppafm_ws_get_params", c_int, [c_void_p, c_void_p, c_int]),
    ("ppafm_ws_get_state", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.POINTER(ctypes.c_longlong)]),
    ("ppafm_ws_set_state", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_longlong]),
    ("ppafm_ws_forward", c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    ("ppafm_ws_loss_grad", c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("ppafm_ws_train", c_int, [c_void_p, c_long, c_void_p, c_void_p, c_int, c_long, ctypes.c_float, ctypes.c_float, c_void_p, c_void_p]),
    ("ppafm_malloc", c_int, [ctypes.POINTER(c_void_p), c_ull]),
    ("ppafm_free", c_int, [c_void_p]),
    ("ppafm_memcpy_h2d", c_int, [c_void_p, c_void_p, c_ull, c_void_p]),
    ("ppafm_memcpy_d2h", c_int, [c_void_p, c_void_p, c_ull, c_void_p]),
    ("ppafm_memcpy_d2d", c_int, [c_void_p, c_void_p, c_ull, c_void_p]),
    ("ppafm_memset", c_int, [c_void_p, c_int, c_ull, c_void_p]),
    ("ppafm_synchronize", c_int, [c_void_p]),
    ("ppafm_device_count", c_int, []),
    ("ppafm_event_create", c_int, [ctypes.POINTER(c_void_p)]),
    ("ppafm_event_record", c_int, [c_void_p, c_void_p]),
    ("ppafm_event_elapsed_ms", c_int, [c_void_p, c_void_p, ctypes.POINTER(ctypes.c_float)]),
    ("ppafm_event_destroy", c_int, [c_void_p]),
    ("ppafm_stream_create", c_int, [ctypes.POINTER(c_void_p)]),
    ("ppafm_stream_destroy", c_int, [c_void_p]),
    ("ppafm_xsf_write_data", c_int, [ctypes.c_char_p, c_int, c_void_p, c_int, c_int, c_int, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_longlong, c_int]),
    ("ppafm_xsf_write_data_dev", c_int, [ctypes.c_char_p, c_int, c_void_p, c_int, c_int, c_int, c_int, c_int, c_int]),
    ("ppafm_read_numbers", c_int, [ctypes.c_char_p, c_int, ctypes.c_longlong, c_void_p, c_int]),
    ("ppafm_last_error", ctypes.c_char_p, []),
    ("ppafm_clear_error", None, []),
    ("ppafm_last_itrsum", ctypes.c_longlong, []),
    # (4) Hartree / density-based force-field grids
    ("ppafm_grid_interp_at", c_int, [_dp, _ip, c_void_p, _dp, _ip, c_void_p, _dp, _dp, c_void_p]),
    ("ppafm_grid_interp_tip_at", c_int, [_dp, _ip, c_void_p, _dp, _ip, c_void_p, c_void_p]),
    ("ppafm_grid_grad", c_int, [_ip, _dp, c_void_p, _dp, c_int, c_int, c_void_p, c_void_p]),
    ("ppafm_grid_grad_potential", c_int, [_dp, _ip, c_void_p, _dp, _ip, c_double, c_void_p, c_void_p]),
    ("ppafm_grid_add_tip_charges", c_int, [_dp, _ip, c_void_p, _dp, _dp, c_void_p, c_void_p]),
    ("ppafm_grid_sums", c_int, [ctypes.c_longlong, c_void_p, _dp, c_void_p]),
    ("ppafm_grid_power", c_int, [ctypes.c_longlong, c_void_p, c_void_p, c_double, c_double, c_void_p]),
    ("ppafm_grid_add_mult", c_int, [ctypes.c_longlong, c_void_p, c_void_p, c_double, c_void_p, c_void_p]),
    ("ppafm_grid_clamp", c_int, [ctypes.c_longlong, c_void_p, c_void_p, c_double, c_double, c_int, c_double, c_void_p]),
    ("ppafm_tip_density_multipole", c_int, [_ip, _dp, _dp, _dp, _dp, c_double, c_double, c_int, _ip, _dp, _dp, c_void_p, c_void_p]),
    ("ppafm_fft_corr_create", c_int, [_ip, ctypes.POINTER(c_void_p)]),
    ("ppafm_fft_corr_set_rho", c_int, [c_void_p, c_void_p, c_double, c_void_p]),
    ("ppafm_fft_corr_correlate", c_int, [c_void_p, c_void_p, c_double, c_void_p, c_void_p]),
    ("ppafm_fft_corr_destroy", c_int, [c_void_p]),
    ("ppafm_grid_add_vdw", c_int, [ctypes.POINTER(Grid), c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("ppafm_grid_add_d3", c_int, [ctypes.POINTER(Grid), c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
]

_lib = None


class HipError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise HipError(f"libppafm_hip.so is not built ({LIB_PATH}); run `python -m ppafm_amd.build`")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != 0:
        msg = lib().ppafm_last_error().decode(errors="replace")
        raise HipError(f"{what} failed ({rc}): {msg}")


def dptr(a):
    """ctypes double* of a C-contiguous float64 numpy array (or None)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def iptr(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)
