"""The parts of ``ppafm.io`` the hot path needs: the xyz reader
(io.py:19-75), geometry loading for xyz input (io.py:398-495) and the force
clamp ``limit_vec_field`` (io.py:766-778), which runs on the GPU."""

import copy

import numpy as np

from . import _lib, device

# Symbol -> Z for the xyz reader (the elements this package ships parameters
# for, plus the rest of periods 1-3 and common substrates).
ELEMENT_Z = {
    "H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "Ne": 10, "Na": 11, "Mg": 12, "Al": 13,
    "Si": 14, "P": 15, "S": 16, "Cl": 17, "Ar": 18, "K": 19, "Ca": 20, "Cu": 29, "Zn": 30, "Br": 35, "Ag": 47, "I": 53,
    "Pt": 78, "Au": 79,
}


def loadXYZ(fname):
    """Read an xyz file: xyzs (N,3), Zs (N,), qs (N,) from the fifth column
    (zeros if absent), and the comment line."""
    xyzs, Zs, qcol = [], [], []
    with open(fname) as f:
        line = f.readline().strip()
        try:
            N = int(line)
        except ValueError:
            raise ValueError(f"The first line of an xyz file should have the number of atoms, but got `{line}`")
        comment = f.readline().strip()
        for i, line in enumerate(f):
            if i >= N:
                break
            wds = line.split()
            try:
                Z = wds[0]
                Z = ELEMENT_Z[Z] if Z in ELEMENT_Z else int(Z)
                xyzs.append((float(wds[1]), float(wds[2]), float(wds[3])))
                Zs.append(Z)
                qcol.append(float(wds[4]) if len(wds) > 4 else None)
            except (ValueError, IndexError):
                raise ValueError(f"Could not interpret line in xyz file: `{line}`")
    xyzs = np.array(xyzs, dtype=np.float64)
    Zs = np.array(Zs, dtype=np.int32)
    if qcol and qcol[0] is not None:
        qs = np.array([q if q is not None else 0.0 for q in qcol], dtype=np.float64)
    else:
        qs = np.zeros(len(Zs), dtype=np.float64)
    return xyzs, Zs, qs, comment


def parseLvecASE(comment):
    """Lattice from an extended-xyz comment ``Lattice="ax ay az bx ..."`` (or None)."""
    if 'Lattice="' not in comment:
        return None
    vals = comment.split('Lattice="')[1].split('"')[0].split()
    v = np.array([float(x) for x in vals]).reshape(3, 3)
    return [np.zeros(3), v[0], v[1], v[2]]


def loadGeometry(fname=None, format=None, parameters=None):
    """xyz geometry -> (atoms [Zs, xs, ys, zs, qs], nDim, lvec (4,3)), filling
    automatic lattice vectors and grid sizes as io.py:468-495 does."""
    if fname is None:
        raise ValueError("Please provide the name of the file with coordinates")
    if parameters is None:
        raise ValueError("Please provide the parameters dictionary here")
    fmt = (format or "").lower() or ("xyz" if fname.lower().endswith(".xyz") else None)
    if fmt != "xyz":
        raise ValueError(f"Unsupported geometry format {fmt!r}: this package reads xyz")
    xyzs, Zs, qs, comment = loadXYZ(fname)
    lvec = parseLvecASE(comment)
    atoms = [list(Zs), list(xyzs[:, 0]), list(xyzs[:, 1]), list(xyzs[:, 2]), list(qs)]
    nDim = np.array(parameters.gridN)
    if lvec is None:
        lvec = np.zeros((4, 3))
        lvec[0, :] = copy.copy(parameters.gridO)
        lvec[1, :] = copy.copy(parameters.gridA)
        lvec[2, :] = copy.copy(parameters.gridB)
        lvec[3, :] = copy.copy(parameters.gridC)
    lvec = np.array(lvec, dtype=np.float64)
    probe_min = np.array(parameters.scanMin, dtype=np.float64)
    probe_max = np.array(parameters.scanMax, dtype=np.float64)
    probe_min[:2] += parameters.r0Probe[:2]
    probe_max[:2] += parameters.r0Probe[:2]
    probe_min[2] -= parameters.r0Probe[2]
    probe_max[2] -= parameters.r0Probe[2]
    pad, step = 3.0, 0.1
    for i in range(3):
        if np.allclose(lvec[i + 1, :], 0):
            parameters.PBC = False
            lvec[i + 1, i] = probe_max[i] - probe_min[i] + 2 * pad
            lvec[0, i] = probe_min[i] - pad
        if not nDim[i] > 0:
            nDim[i] = round(np.linalg.norm(lvec[i + 1, :]) / step)
    parameters.gridO = lvec[0].tolist()
    parameters.gridA, parameters.gridB, parameters.gridC = (lvec[k].tolist() for k in (1, 2, 3))
    parameters.gridN = [int(v) for v in nDim]
    return atoms, nDim, lvec


def limit_vec_field(FF, Fmax=100.0):
    """Scale vectors longer than Fmax to length Fmax, in place (on the GPU)."""
    if not (isinstance(FF, np.ndarray) and FF.dtype == np.float64 and FF.flags.c_contiguous and FF.shape[-1] == 3):
        raise TypeError("FF must be a C-contiguous float64 array with a last axis of 3")
    npts = FF.size // 3
    if npts == 0:
        return
    d = device.DeviceArray.from_host(FF)
    _lib.check(_lib.lib().ppafm_limit_vec_field(d.ptr, npts, float(Fmax), None), "limit_vec_field")
    d.download(out=FF)
