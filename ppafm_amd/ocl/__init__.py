"""GPU-facade mirror of ``ppafm.ocl`` (the AFMulator class)."""
