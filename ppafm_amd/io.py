"""The parts of ``ppafm.io`` the hot path and its file formats need: the xyz
reader (io.py:19-75), geometry loading for xyz input (io.py:398-495), the
force clamp ``limit_vec_field`` (io.py:766-778, on the GPU), and the grid
files -- XSF (io.py:531-640), npz (io.py:716-750), cube input (io.py:643-681)
and the vector/scalar field wrappers over them (io.py:752-880).  XSF data
blocks are formatted and parsed by the library's multithreaded native code;
the files are byte-identical to the reference's and load to the same bits."""

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


# ------------------------------------------------------------------ grid files

bohrRadius2angstroem = 0.5291772109217  # io.py:13
Hartree2eV = 27.211396132  # io.py:14

# The reference's default XSF header (io.py:563-572), byte for byte: files
# written with it are what ppafm's tools expect to read.
XSF_HEAD_DEFAULT = headScan = """
ATOMS
 1   0.0   0.0   0.0

BEGIN_BLOCK_DATAGRID_3D
   some_datagrid
   BEGIN_DATAGRID_3D_whatever
"""


def _as_f64_view(a):
    """float64 array with non-negative strides that are whole elements
    (a copy only when ``a`` is not one already)."""
    a = np.asarray(a)
    if a.dtype != np.float64 or any(st < 0 or st % 8 for st in a.strides):
        a = np.ascontiguousarray(a, dtype=np.float64)
    return a


def _line(values):
    # io.py:550-556 writes each entry with str()
    return " ".join(str(v) for v in values) + "\n"


def _orthoLvec(sh, dd):
    """Orthogonal cell of a (nz,ny,nx) grid with steps dd (io.py:559-560)."""
    return [[0, 0, 0], [sh[2] * dd[0], 0, 0], [0, sh[1] * dd[1], 0], [0, 0, sh[0] * dd[2]]]


def saveXSFData(fname, data, lvec=None, dd=None, head=XSF_HEAD_DEFAULT, data_is_xyz_order=True, nthreads=0):
    """Write a scalar grid as XSF (io.py:575-599): data (nx,ny,nz), or
    (nz,ny,nx) with data_is_xyz_order=False."""
    a = _as_f64_view(data)
    if a.ndim != 3:
        raise ValueError(f"data must be 3-dimensional, got shape {a.shape}")
    if data_is_xyz_order:
        nx, ny, nz = a.shape
        sx, sy, sz = (st // 8 for st in a.strides)
    else:
        nz, ny, nx = a.shape
        sz, sy, sx = (st // 8 for st in a.strides)
    if lvec is None:
        lvec = _orthoLvec((nz, ny, nx), [1.0, 1.0, 1.0] if dd is None else dd)
    write_xsf_header(fname, head, (nx, ny, nz), lvec)
    _lib.check(_lib.lib().ppafm_xsf_write_data(str(fname).encode(), 1, a.ctypes.data, nx, ny, nz, sx, sy, sz, int(nthreads)),
               "saveXSFData")


def write_xsf_header(fname, head, n, lvec):
    """Start an XSF file: the header, the extended grid size and the four
    cell rows (io.py:588-590); the data block is appended by the library."""
    with open(fname, "w") as f:
        f.write("".join(head))
        f.write(_line((n[0] + 1, n[1] + 1, n[2] + 1)))
        for row in lvec:
            f.write(_line(row))


def readNumsUpTo(filename, dimensions, noline, nthreads=0):
    """prod(dimensions) numbers after ``noline`` lines (GridUtils.py:161-164)."""
    n = int(np.prod(np.asarray(dimensions, dtype=np.int64)))
    out = np.zeros(n)
    _lib.check(_lib.lib().ppafm_read_numbers(str(filename).encode(), int(noline), n, out.ctypes.data, int(nthreads)),
               "readNumsUpTo")
    return out


def _read_until(f, keyword):
    lines = []
    while True:
        line = f.readline()
        lines.append(line)
        if not line or keyword in line:
            return len(lines), lines


def loadXSFData(fname, xyz_order=True):
    """Read an XSF scalar grid (io.py:601-620): (FF, lvec, nDim, head) with
    FF (nx,ny,nz) for xyz_order (else (nz,ny,nx)) without the repeated
    periodic planes, lvec (4,3) origin and cell, nDim its shape, head the
    lines up to the DATAGRID_3D one."""
    with open(fname) as f:
        nhead, head = _read_until(f, "BEGIN_DATAGRID_3D")
        dims = [int(v) for v in f.readline().split()]
        lvec = np.array([[float(v) for v in f.readline().split()] for _ in range(4)])
    nDim = np.array(dims[::-1])
    F = readNumsUpTo(fname, nDim.astype(np.int32), nhead + 5)
    FF = F.reshape(nDim)[:-1, :-1, :-1]
    if xyz_order:
        FF = FF.transpose((2, 1, 0))
        nDim = nDim[::-1].copy()
    return FF.copy(), lvec, nDim - 1, head


def getFromHead_PRIMCOORD(head):
    """Atoms (Zs, Rs) listed under PRIMCOORD in an XSF header (io.py:623-637)."""
    for i, line in enumerate(head):
        if "PRIMCOORD" in line:
            n = int(head[i + 1].split()[0])
            rows = [head[i + 2 + j].split() for j in range(n)]
            return np.array([int(r[0]) for r in rows], dtype=np.int32), np.array([[float(v) for v in r[1:4]] for r in rows])
    return None, None


def primcoords2Xsf(iZs, xyzs, lvec):
    """XSF header with the cell and atoms of a geometry (io.py:323-347)."""
    rows = "".join(f"{lvec[k][0]:f} {lvec[k][1]:f} {lvec[k][2]:f}\n" for k in (1, 2, 3))
    atoms = "".join("%i %5.6f %5.6f %5.6f\n" % (iZs[i], xyzs[0][i], xyzs[1][i], xyzs[2][i]) for i in range(len(iZs)))
    return ("CRYSTAL\nPRIMVEC\n" + rows + "CONVVEC\n" + rows + "PRIMCOORD\n" + "%i 1\n" % len(iZs) + atoms +
            "\nBEGIN_BLOCK_DATAGRID_3D\nsome_datagrid\nBEGIN_DATAGRID_3D_whatever\n")


def loadCUBE(fname, xyz_order=False):
    """Gaussian cube grid (io.py:643-681) -> (FF in eV, lvec in A, nDim, head);
    FF (nx,ny,nz) with xyz_order, else transposed to (nz,ny,nx)."""
    with open(fname) as f:
        f.readline()
        f.readline()
        h0 = f.readline().split()
        axes = [f.readline().split() for _ in range(3)]
    nDim = np.array([int(a[0]) for a in axes])
    lvec = np.zeros((4, 3))
    for j in range(3):
        lvec[0, j] = float(h0[j + 1]) * bohrRadius2angstroem
        for k in range(3):
            lvec[k + 1, j] = float(axes[k][j + 1]) * int(axes[k][0]) * bohrRadius2angstroem
    F = readNumsUpTo(fname, nDim.astype(np.int32), 6 + int(h0[0]))
    FF = F.reshape(nDim)
    if not xyz_order:
        FF = FF.transpose((2, 1, 0)).copy()
        nDim = [nDim[2], nDim[1], nDim[0]]
    head = ["BEGIN_BLOCK_DATAGRID_3D \n", "g98_3D_unknown \n", "DATAGRID_3D_g98Cube \n"]
    FF *= Hartree2eV
    return FF, lvec, nDim, head


def loadXSFGeom(fname):
    """Atoms and grid cell of an XSF file (io.py:229-268): ([Zs, xs, ys, zs,
    qs], nDim, lvec) with the atoms listed under PRIMCOORD (charges 0), nDim
    the grid shape without the repeated periodic planes and lvec the four
    rows after it.  A file without PRIMCOORD is an error (ValueError or
    IndexError, as in the reference)."""
    e, x, y, z, q = [], [], [], [], []
    with open(fname) as f:
        for _ in range(10000):
            if "PRIMCOORD" in f.readline():
                break
        n = int(f.readline().split()[0])
        for _ in range(n):
            ws = f.readline().split()
            e.append(int(ws[0]))
            x.append(float(ws[1]))
            y.append(float(ws[2]))
            z.append(float(ws[3]))
            q.append(0)
        for _ in range(10000):
            line = f.readline()
            if "BEGIN_DATAGRID_3D" in line or "DATAGRID_3D_DENSITY" in line:
                break
        ws = f.readline().split()
        nDim = [int(ws[0]), int(ws[1]), int(ws[2])]
        lvec = [[float(v) for v in f.readline().split()[:3]] for _ in range(4)]
    return [e, x, y, z, q], (nDim[0] - 1, nDim[1] - 1, nDim[2] - 1), lvec


def loadAtomsCUBE(fname):
    """Atoms of a Gaussian cube file (io.py:291-320): [Zs, xs, ys, zs, qs] in
    angstrom (bohr * 0.5291772109217), charges 0."""
    b2a = 0.5291772109217
    e, x, y, z, q = [], [], [], [], []
    with open(fname) as f:
        f.readline()
        f.readline()
        natoms = int(f.readline().split()[0])
        for _ in range(3):
            f.readline()
        for _ in range(natoms):
            ws = f.readline().split()
            x.append(float(ws[2]) * b2a)
            y.append(float(ws[3]) * b2a)
            z.append(float(ws[4]) * b2a)
            e.append(int(ws[0]))
            q.append(0.0)
    return [e, x, y, z, q]


def saveNpy(fname, data, lvec, atomic_info):
    """Scalar grid, its cell and the source atoms as ``fname``.npz (io.py:716-726)."""
    np.savez(fname + ".npz", data=data, lvec=lvec, atoms=atomic_info[0], lvec0=atomic_info[1])


def loadNpy(fname):
    """io.py:729-745: (data, lvec, (atoms, lvec0)) from ``fname``.npz."""
    with np.load(fname + ".npz") as z:
        return z["data"].copy(), z["lvec"], (z["atoms"], z["lvec0"])


def packVecGrid(Fx, Fy, Fz, FF=None):
    """Stack three scalar grids into (..., 3) (io.py:752-759)."""
    if FF is None:
        FF = np.zeros(np.shape(Fx) + (3,))
    FF[..., 0], FF[..., 1], FF[..., 2] = Fx, Fy, Fz
    return FF


def unpackVecGrid(FF):
    return FF[..., 0].copy(), FF[..., 1].copy(), FF[..., 2].copy()


def save_vec_field(fname, data, lvec, data_format="xsf", head=XSF_HEAD_DEFAULT, atomic_info=None):
    """Vector grid (nx,ny,nz,3) as three XSF files fname_x/_y/_z.xsf or one
    fname.npz (io.py:781-802).  An unknown format writes nothing, as there
    (which only logs it)."""
    if data_format == "xsf":
        for k, c in enumerate("xyz"):
            saveXSFData(f"{fname}_{c}.xsf", data[..., k], lvec, head=head)
    elif data_format == "npy":
        atomic_info = atomic_info if atomic_info is not None else (np.zeros((4, 1)), lvec)
        np.savez(fname + ".npz", FF=data, lvec=lvec, atoms=atomic_info[0], lvec0=atomic_info[1])


def load_vec_field(fname, data_format="xsf"):
    """io.py:805-839: (data (nx,ny,nz,3), lvec, nDim, head or (atoms, lvec0))."""
    if data_format == "xsf":
        parts = [loadXSFData(f"{fname}_{c}.xsf") for c in "xyz"]
        data = packVecGrid(parts[0][0], parts[1][0], parts[2][0])
        _, lvec, nDim, info = parts[2]
    elif data_format == "npy":
        with np.load(fname + ".npz") as z:
            data, lvec, info = z["FF"].copy(), z["lvec"], (z["atoms"], z["lvec0"])
        nDim = data.shape
    else:
        raise ValueError(f"Cannot load vector field in format: {data_format}")
    return data, lvec, nDim, info


def save_scal_field(fname, data, lvec, data_format="xsf", head=XSF_HEAD_DEFAULT, atomic_info=None):
    """Scalar grid as fname.xsf or fname.npz (io.py:842-861)."""
    if data_format == "xsf":
        saveXSFData(fname + ".xsf", data, lvec, head=head)
    elif data_format == "npy":
        saveNpy(fname, data, lvec, atomic_info if atomic_info is not None else (np.zeros((4, 1)), lvec))


def load_scal_field(fname, data_format="xsf"):
    """io.py:864-880: (data, lvec, nDim, head or (atoms, lvec0)); also cube."""
    if data_format == "xsf":
        data, lvec, nDim, info = loadXSFData(fname + ".xsf")
    elif data_format == "npy":
        data, lvec, info = loadNpy(fname)
        nDim = data.shape
    elif data_format == "cube":
        data, lvec, nDim, info = loadCUBE(fname + ".cube")
    else:
        raise ValueError(f"Cannot load scalar field in format: {data_format}")
    return data.copy(), lvec, nDim, info
