"""Build libppafm_hip.so for gfx950 in-tree (``python -m ppafm_amd.build``).

Each .hip file is compiled with hipcc, each host-only .cpp file with g++,
into an object under ppafm_amd/_build/; all are linked into
ppafm_amd/libppafm_hip.so.  Packed FP32 instructions are disabled
(-packed-fp32-ops) for every device compile.
"""

import concurrent.futures
import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libppafm_hip.so"
ARCH = os.environ.get("PPAFM_HIP_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CXX = shutil.which("g++") or "g++"
# hipFFT (the FFT cross-correlations of the Hartree force field)
ROCM_LIB = pathlib.Path(os.environ.get("ROCM_PATH", "/opt/rocm")) / "lib"

COMMON_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def sources():
    return sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))


def _compile(src, verbose):
    obj = BUILD / (src.stem + ".o")
    if src.suffix == ".cpp":
        cmd = [CXX, "-O3", "-std=c++17", "-fPIC", "-pthread", "-Wall", "-c", str(src), "-o", str(obj)]
    else:
        cmd = [HIPCC, *COMMON_FLAGS, "-c", str(src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{cmd[0]} failed on {src.name}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=False, jobs=None):
    BUILD.mkdir(exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), os.cpu_count() or 1, 8)
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", *map(str, objs), "-o", str(tmp),
           f"-L{ROCM_LIB}", "-lhipfft", f"-Wl,-rpath,{ROCM_LIB}"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
