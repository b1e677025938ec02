"""Host-side parameters and small helpers of the probe-particle path.

Mirrors the parts of ``ppafm.common`` that the hot path uses (same names,
arguments and units):

  * constants ``eVA_Nm``, ``CoulombConst``          common.py:20-21
  * Lennard-Jones mixing ``get_C612`` / ``getAtomsLJ`` common.py:851-868
  * ``REA2LJ`` / ``getAtomsREA``                     common.py:871-892
  * periodic images ``PBCAtoms`` / ``PBCAtoms3D_np`` common.py:608-628, 728-776
  * ``prepareScanGrids``                             common.py:922-946
  * ``genFFSampling``                                common.py:967-977
  * df weights ``get_df_weight`` / ``get_simple_df_weight`` and ``Fz2df``
    (the convolution itself runs on the GPU)        common.py:333-430
  * ``loadSpecies`` (reads an atomtypes.ini-format file; the built-in table
    carries the rows this path needs)               common.py:526-543
"""

import math
import os

import numpy as np

eVA_Nm = 16.0217662  # [eV/A^2] / [N/m]
CoulombConst = 14.399645  # [eV A / e^2]

# Lennard-Jones element parameters (R_i [A], eps_i [eV], alpha, Z, symbol):
# the values of the rows H, C, N, O, Ag, Au of ppafm/defaults/atomtypes.ini
# (lines 1, 6, 7, 8, 47, 79).  Any other element: pass a species file.
_BUILTIN_SPECIES = [
    (1.4870, 0.000681, 1.80, 1, "H"),
    (2.0800, 0.003729, 1.80, 5, "B"),
    (1.9080, 0.003729, 1.80, 6, "C"),
    (1.7800, 0.007372, 1.80, 7, "N"),
    (1.6612, 0.009106, 1.80, 8, "O"),
    (1.7500, 0.002645, 1.80, 9, "F"),
    (2.7700, 0.010000, 1.80, 11, "Na"),
    (1.9000, 0.025490, 1.80, 14, "Si"),
    (2.1000, 0.008673, 1.80, 15, "P"),
    (2.0000, 0.010841, 1.80, 16, "S"),
    (1.9480, 0.011491, 1.80, 17, "Cl"),
    (2.2300, 0.010000, 1.80, 29, "Cu"),
    (2.2400, 0.010000, 1.80, 30, "Zn"),
    (2.2200, 0.013876, 1.80, 35, "Br"),
    (2.3700, 0.010000, 1.80, 47, "Ag"),
    (2.3500, 0.017345, 1.80, 53, "I"),
    (2.1815, 0.024344, 1.80, 54, "Xe"),
    (2.3700, 0.010000, 1.80, 78, "Pt"),
    (2.4100, 0.010000, 1.80, 79, "Au"),
]

_SPECIES_DTYPE = [("rmin", np.float64), ("epsilon", np.float64), ("alpha", np.float64), ("atom", int), ("symbol", "|S10")]


class SpeciesTable:
    """Indexable by Z-1 like the reference's FFparams array (common.py:867)."""

    def __init__(self, rows):
        self.rows = {int(r[3]): r for r in rows}
        self.by_symbol = {str(r[4]).strip(): int(r[3]) for r in rows}

    def __getitem__(self, i):
        z = int(i) + 1
        if z not in self.rows:
            raise KeyError(f"no Lennard-Jones parameters for Z={z}; pass a species file to loadSpecies")
        return self.rows[z]

    def __len__(self):
        return max(self.rows) if self.rows else 0


def loadSpecies(fname=None):
    """Species table; reads ``fname`` (atomtypes.ini format: R eps alpha Z symbol)
    when it exists, the built-in rows otherwise."""
    if fname is not None and os.path.exists(fname):
        rows = []
        with open(fname) as f:
            for line in f:
                p = line.split()
                if len(p) >= 5:
                    try:
                        rows.append((float(p[0]), float(p[1]), float(p[2]), int(p[3]), p[4]))
                    except ValueError:
                        continue
        return SpeciesTable(rows)
    return SpeciesTable(_BUILTIN_SPECIES)


def get_C612(i, j, FFparams):
    """c6 = 2 eps R^6, c12 = eps R^12 with R = R_i + R_j, eps = sqrt(eps_i eps_j)."""
    Rij = FFparams[i][0] + FFparams[j][0]
    Eij = np.sqrt(FFparams[i][1] * FFparams[j][1])
    return 2 * Eij * (Rij**6), Eij * (Rij**12)


def getAtomsLJ(iZprobe, iZs, FFparams):
    n = len(iZs)
    cLJs = np.zeros((n, 2))
    for i in range(n):
        cLJs[i, 0], cLJs[i, 1] = get_C612(iZprobe - 1, iZs[i] - 1, FFparams)
    return cLJs


def getAtomsRE(iZprobe, iZs, FFparams):
    """(R_pp + R_i, sqrt(eps_pp eps_i)) per atom, the (R, E0) coefficients of
    the Morse and damped-vdW grids (common.py:899-905)."""
    n = len(iZs)
    REs = np.zeros((n, 2))
    rpp, epp = FFparams[iZprobe - 1][0], FFparams[iZprobe - 1][1]
    for i in range(n):
        REs[i, 0] = rpp + FFparams[iZs[i] - 1][0]
        REs[i, 1] = epp * FFparams[iZs[i] - 1][1]
    REs[:, 1] = np.sqrt(REs[:, 1])
    return REs


def REA2LJ(cREAs, cLJs=None):
    if cLJs is None:
        cLJs = np.zeros((len(cREAs), 2))
    R6 = cREAs[:, 0] ** 6
    cLJs[:, 0] = 2 * cREAs[:, 1] * R6
    cLJs[:, 1] = cREAs[:, 1] * (R6**2)
    return cLJs


def getAtomsREA(iZprobe, iZs, FFparams, alphaFac=-1.0):
    n = len(iZs)
    REAs = np.zeros((n, 4))
    i = iZprobe - 1
    for ii in range(n):
        j = iZs[ii] - 1
        REAs[ii, 0] = FFparams[i][0] + FFparams[j][0]
        REAs[ii, 1] = np.sqrt(FFparams[i][1] * FFparams[j][1])
        REAs[ii, 2] = FFparams[j][2] * alphaFac
    return REAs


def PBCAtoms(Zs, Rs, Qs, avec, bvec, na=1, nb=1):
    """xy periodic images, image-major order, z not replicated."""
    Zs_, Rs_, Qs_ = [], [], []
    for i in range(-na, na + 1):
        for j in range(-nb, nb + 1):
            for ia in range(len(Zs)):
                Zs_.append(Zs[ia])
                Rs_.append((Rs[ia][0] + i * avec[0] + j * bvec[0], Rs[ia][1] + i * avec[1] + j * bvec[1], Rs[ia][2]))
                Qs_.append(Qs[ia])
    return np.array(Zs_), np.array(Rs_), np.array(Qs_)


def PBCAtoms3D_np(Zs, Rs, Qs, cLJs, REAs, lvec, npbc=(1, 1, 1)):
    """3-D images with the (0,0,0) cell first, then ia, ib, ic ascending."""
    natom = len(Zs)
    shifts = [(0, 0, 0)] + [
        (a, b, c)
        for a in range(-npbc[0], npbc[0] + 1)
        for b in range(-npbc[1], npbc[1] + 1)
        for c in range(-npbc[2], npbc[2] + 1)
        if (a, b, c) != (0, 0, 0)
    ]
    lvec = np.asarray(lvec, dtype=np.float64)
    m = len(shifts)
    Zs_ = np.tile(np.asarray(Zs, dtype=np.int32), m)
    xyzs_ = np.empty((m * natom, 3), np.float32)
    for k, (a, b, c) in enumerate(shifts):
        xyzs_[k * natom : (k + 1) * natom] = np.asarray(Rs)[:, :3] + (a * lvec[0] + b * lvec[1] + c * lvec[2])[None, :]
    qs_ = np.tile(np.asarray(Qs, dtype=np.float32), m)
    cLJs_ = None if cLJs is None else np.tile(np.asarray(cLJs, dtype=np.float32), (m, 1))
    REAs_ = None if REAs is None else np.tile(np.asarray(REAs, dtype=np.float32), (m, 1))
    return Zs_, xyzs_, qs_, cLJs_, REAs_


def findPBCAtoms3D_cutoff(Rs, lvec, Rcut=1.0):
    """Periodic images (cells -1, 0, 1 along each lattice vector) of the
    atoms Rs (3, N) that lie in the cell lvec (3, 3) enlarged by Rcut on
    every side (common.py:667-706): returns (indices of the source atoms,
    image positions (3, M)), images in the order ia, ib, ic ascending."""
    lvec = np.asarray(lvec, dtype=np.float64)
    Rs = np.asarray(Rs, dtype=np.float64)
    inv = np.linalg.inv(lvec)
    a, b, c = np.dot(inv, Rs)
    mA, mB, mC = (np.sqrt(np.dot(inv[k], inv[k])) * Rcut for k in range(3))
    inds, out = [], []
    for ia in (-1, 0, 1):
        in_a = (a > (-mA - ia)) & (a < (mA + 1 - ia))
        for ib in (-1, 0, 1):
            in_ab = in_a & (b > (-mB - ib)) & (b < (mB + 1 - ib))
            for ic in (-1, 0, 1):
                sel = np.nonzero(in_ab & (c > (-mC - ic)) & (c < (mC + 1 - ic)))[0]
                if len(sel) == 0:
                    continue
                shift = ia * lvec[0] + ib * lvec[1] + ic * lvec[2]
                inds.append(sel)
                out.append(Rs[:, sel] + shift[:, None])
    if not inds:
        return np.zeros(0, dtype=np.int64), np.zeros((3, 0))
    return np.concatenate(inds), np.hstack(out)


# Valence electrons per element for the core-density subtraction: the s and
# p electrons of the outer shell for main-group elements, plus the d
# electrons for the transition metals (groups 3-12); numbers depend on the
# pseudopotentials of the DFT code, so callers may pass their own table.
def _valence_table():
    table = {}
    period_starts = [1, 3, 11, 19, 37, 55, 87]
    for z in range(1, 87):
        p = max(k for k, z0 in enumerate(period_starts) if z0 <= z)
        pos = z - period_starts[p] + 1  # position in the period
        if p == 0:
            table[z] = float(pos)  # H 1, He 2
        elif p <= 2:
            table[z] = float(pos)  # Li..Ne, Na..Ar: 1..8
        elif p <= 4:
            table[z] = float(pos if pos <= 12 else pos - 10)  # groups 1-12, then 3..8
        else:
            if pos <= 2:
                table[z] = float(pos)
            elif pos <= 17:
                table[z] = 3.0  # La and the lanthanides
            else:
                g = pos - 14  # groups 4..18 after the lanthanides
                table[z] = float(g if g <= 12 else g - 10)
    return table


VALENCE_ELECTRONS = _valence_table()


def prepareScanGrids(parameters):
    """Tip apex axes and the scan lattice vectors (common.py:922-946)."""
    scanMin, scanMax, scanStep, r0Probe = parameters.scanMin, parameters.scanMax, parameters.scanStep, parameters.r0Probe
    xTips = np.arange(scanMin[0], scanMax[0] + 0.5 * scanStep[0], scanStep[0])
    yTips = np.arange(scanMin[1], scanMax[1] + 0.5 * scanStep[1], scanStep[1])
    zTips = np.arange(scanMin[2], scanMax[2] + 0.5 * scanStep[2], scanStep[2])
    lvecScan = np.array(
        [
            [scanMin[0] + r0Probe[0], scanMin[1] + r0Probe[1], scanMin[2] - r0Probe[2]],
            [scanMax[0] - scanMin[0], 0.0, 0.0],
            [0.0, scanMax[1] - scanMin[1], 0.0],
            [0.0, 0.0, scanMax[2] - scanMin[2]],
        ]
    )
    return xTips, yTips, zTips, lvecScan


def genFFSampling(lvec, pixPerAngstrome=10):
    return np.array([int(round(pixPerAngstrome * math.sqrt(np.dot(lvec[k], lvec[k])))) for k in (1, 2, 3)] + [4], np.int32)


# ------------------------------------------------------------------ Fz -> df


def get_df_weight(Amp, dz=0.1):
    """Giessibl (2001) exact weights for the Fz -> df correlation.

    n = ceil(Amp/dz) intervals and n+1 taps.  The nodes t_k = -1 + 2(n-k)dz/Amp
    (k = 0..n) are taken in descending order so the returned kernel is
    already reversed for ``np.convolve``.  Integrating t/sqrt(1-t^2) against the
    piecewise-linear Fz with F(t)=sqrt(1-t^2) and G(t)=(arccos(-t) - t F(t))/2,
    F = 0, G = pi/2 at the first node and F = G = 0 at the last, interval k
    contributes t_k dF_k + dG_k to tap k+1 and -(t_{k+1} dF_k + dG_k) to tap k;
    the sum is divided by pi dz."""
    n = int(np.ceil(Amp / dz))
    nodes = np.linspace(-1 + 2 * n * dz / Amp, -1, n + 1)
    F = np.zeros(n + 1)
    G = np.zeros(n + 1)
    if n > 1:
        inner = nodes[1:-1]
        F[1:-1] = np.sqrt(1.0 - inner * inner)
        G[1:-1] = 0.5 * (np.arccos(-inner) - inner * F[1:-1])
    G[0] = 0.5 * np.pi
    dF = F[:-1] - F[1:]
    dG = G[:-1] - G[1:]
    w = np.zeros(n + 1)
    w[1:] += nodes[:-1] * dF + dG
    w[:-1] -= nodes[1:] * dF + dG
    return w / (np.pi * dz)


def get_simple_df_weight(n=10, dz=0.1):
    """Simplified df kernel of n taps (common.py:395-418)."""
    n = max(int(n), 2)
    mid = np.linspace(-1 + 0.5 / n, 1 - 0.5 / n, n - 1)
    s = np.sqrt(1.0 - mid * mid)
    w = np.zeros(n)
    w[:-1] += s
    w[1:] -= s
    return w / (np.dot(np.linspace(-1, 1, n), w) * n * dz)


def Fz2df(F, dz, k0, f0, amplitude=1.0, units=16.0217656):
    """Valid-mode np.convolve of F along its last axis with get_df_weight,
    times units*f0/k0 -- run on the GPU (common.py:421-430)."""
    from . import device

    W = get_df_weight(amplitude, dz=dz)
    return device.convolve_z(np.asarray(F, dtype=np.float64), W[::-1].copy(), scale=units * f0 / k0)


def Fz2df_tilt(F, d, k0, f0, amplitude=1.0, units=16.0217656):
    """common.py:433-445: project the per-axis convolutions on the tilt direction."""
    from . import device

    dr = math.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2)
    W = get_df_weight(amplitude, dz=dr)[::-1].copy()
    F = np.asarray(F, dtype=np.float64)
    out = 0
    for k in range(3):
        out = out + device.convolve_z(np.ascontiguousarray(F[..., k]), W, scale=1.0) * d[k]
    return out * units * f0 / k0


def maxAlongDir(atoms, hdir):
    """Index and projection of the atom furthest along hdir (common.py:500-503)."""
    proj = np.asarray(atoms)[:, :3] @ np.asarray(hdir, dtype=np.float64)
    i = int(np.argmax(proj))
    return i, proj[i]


def maxAlongDirEntropy(atoms, hdir, beta=1.0):
    """maxAlongDir plus sum_k exp(beta (p_k - p_max)) over the projections p
    (common.py:506-510): how many atoms share the top along hdir."""
    proj = np.asarray(atoms)[:, :3] @ np.asarray(hdir, dtype=np.float64)
    i = int(np.argmax(proj))
    return i, proj[i], np.sum(np.exp(beta * (proj - proj[i])))


# ------------------------------------------------------------------ parameters


class PpafmParameters:
    """Simulation parameters with the names and defaults of
    ppafm.common.PpafmParameters (common.py:26-75); ``from_file`` reads the
    params.ini format of ``load_ini`` (common.py:119-160)."""

    _DEFAULTS = dict(
        PBC=True, nPBC=[1, 1, 1], gridN=[-1, -1, -1], gridO=[0.0, 0.0, 0.0], gridA=[20.0, 0.0, 0.0], gridB=[0.0, 20.0, 0.0],
        gridC=[0.0, 0.0, 20.0], FFgrid0=[-1.0, -1.0, -1.0], FFgridA=[-1.0, -1.0, -1.0], FFgridB=[-1.0, -1.0, -1.0],
        FFgridC=[-1.0, -1.0, -1.0], moleculeShift=[0.0, 0.0, 0.0], probeType="O", charge=0.00, Apauli=18.0, Bpauli=1.0,
        ffModel="LJ", Rcore=0.7, r0Probe=[0.00, 0.00, 4.00], stiffness=[-1.0, -1.0, -1.0], klat=0.5, krad=20.00, tip="s",
        sigma=0.7, scanStep=[0.10, 0.10, 0.10], scanMin=[0.0, 0.0, 5.0], scanMax=[20.0, 20.0, 8.0], scanTilt=[0.0, 0.0, -0.1],
        tiltedScan=False, kCantilever=1800.0, f0Cantilever=30300.0, Amplitude=1.0, plotSliceFrom=16, plotSliceTo=22,
        plotSliceBy=1, imageInterpolation="bicubic", colorscale="gray", colorscale_kpfm="seismic", ddisp=0.05, aMorse=-1.6,
        Rtip=30.0, permit=0.00552634959, vdWDampKind=2, Vbias=0.0,
    )

    def __init__(self, **kw):
        import copy

        for k, v in self._DEFAULTS.items():
            setattr(self, k, copy.deepcopy(v))
        for k, v in kw.items():
            if k not in self._DEFAULTS:
                raise ValueError(f"Parameter {k} is not known")
            setattr(self, k, v)

    @classmethod
    def from_file(cls, file_path="params.ini"):
        obj = cls()
        try:
            with open(file_path) as f:
                if str(file_path).endswith(".ini"):
                    obj.load_ini(f.readlines())
                else:
                    raise ValueError(f"Unsupported file extension: {os.path.splitext(str(file_path))[1]}")
        except FileNotFoundError:
            raise ValueError(f"File {file_path} not found")
        return obj

    def apply_options(self, obj):
        for key, value in obj.items():
            if hasattr(self, key) and value is not None:
                setattr(self, key, value)

    def load_ini(self, lines):
        for line in lines:
            line = line.split("#")[0].strip()
            words = line.split()
            if len(words) < 2:
                continue
            key = words[0]
            if key not in self._DEFAULTS:
                raise ValueError(f"Parameter {key} is not known")
            val = getattr(self, key)
            try:
                if isinstance(val, bool):
                    setattr(self, key, words[1].strip()[0] in "Tt")
                elif isinstance(val, float):
                    setattr(self, key, float(words[1]))
                elif isinstance(val, int):
                    setattr(self, key, int(words[1]))
                elif isinstance(val, str):
                    setattr(self, key, words[1])
                elif isinstance(val, list):
                    if isinstance(val[0], float):
                        setattr(self, key, [float(words[1]), float(words[2]), float(words[3])])
                    elif isinstance(val[0], int):
                        setattr(self, key, [int(words[1]), int(words[2]), int(words[3])])
                    else:
                        setattr(self, key, [str(words[1]), float(words[2])])
            except Exception:
                raise ValueError(f"Error parsing parameters on line: {line}") from None
        if self.gridN[0] <= 0:
            autoGridN(self)
        self.tip = self.tip.replace('"', "").replace("'", "")


def autoGridN(parameters):
    """Grid points at 0.1 A from the cell lengths (common.py:516-522)."""
    parameters.gridN = [
        round(np.linalg.norm(parameters.gridA) * 10),
        round(np.linalg.norm(parameters.gridB) * 10),
        round(np.linalg.norm(parameters.gridC) * 10),
    ]
    return parameters.gridN
