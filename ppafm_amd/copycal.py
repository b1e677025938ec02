"""Copy-kernel calibration for HBM byte counters.

``python -m ppafm_amd.copycal [MiB] [reps]`` copies a buffer of the given size
with the library's plain copy kernel; under ``rocprofv3 --pmc FETCH_SIZE`` (or
WRITE_SIZE) the counter per dispatch divided by the known bytes moved gives
the scale applied to the same counter on other kernels."""

import sys

import numpy as np

from . import _lib
from . import device as D


def main():
    mib = int(sys.argv[1]) if len(sys.argv) > 1 else 256
    reps = int(sys.argv[2]) if len(sys.argv) > 2 else 3
    nbytes = mib << 20
    src = D.DeviceArray((nbytes // 4,), np.float32)
    dst = D.DeviceArray((nbytes // 4,), np.float32)
    src.zero()
    for _ in range(reps):
        _lib.check(_lib.lib().ppafm_copy_kernel(src.ptr, dst.ptr, nbytes, None), "copy")
    D.synchronize()
    print(f"copied {mib} MiB x {reps}: {nbytes} bytes read and {nbytes} bytes written per dispatch", flush=True)


if __name__ == "__main__":
    main()
