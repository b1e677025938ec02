"""MI355X (gfx950) implementation of the ppafm probe-particle hot path.

Modules mirror the reference package layout so callers can switch imports:
``ppafm_amd.core`` (ppafm.core), ``ppafm_amd.HighLevel`` (ppafm.HighLevel),
``ppafm_amd.common`` (ppafm.common), ``ppafm_amd.io`` (ppafm.io) and
``ppafm_amd.ocl.AFMulator`` (ppafm.ocl.AFMulator).  The compute runs in
``libppafm_hip.so`` (HIP kernels for gfx950); build it with
``python -m ppafm_amd.build``.
"""

__version__ = "0.1.0"
