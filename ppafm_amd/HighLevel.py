"""Drop-in for the hot-path parts of ``ppafm.HighLevel`` (HighLevel.py:36-387):
grid setup, the LJ and point-charge grids, and the relaxed scan, running on
the GPU through :mod:`ppafm_amd.core`."""

import numpy as np

from . import common as PPU
from . import core, io

Fmax_DEFAULT = 10.0
Vmax_DEFAULT = 10.0


def trjByDir(n, d, p0):
    trj = np.zeros((n, 3))
    for k in range(3):
        trj[:, k] = p0[k] + (np.arange(n)[::-1]) * d[k]
    return trj


def shift_positions(R, s):
    Rs = np.array(R, dtype=np.float64).copy()
    Rs += np.asarray(s, dtype=np.float64)[None, :]
    return Rs


def relaxedScan3D(xTips, yTips, zTips, trj=None, bF3d=False, tip_spline=None):
    """Relaxed scan along a tip trajectory ``trj`` (ntips,3), default straight
    down zTips (HighLevel.py:63-101): column (ix, iy) is the stroke
    trj + (x, y, 0).  The reference relaxes the columns one relaxTipStroke
    call at a time; here all of them go to the GPU in one
    relaxTipStrokes_omp call with explicit tip positions, which relaxes each
    column as the same stroke.  Returns (fzs, PPpos); fzs is (ntips, ny, nx)
    Fz, or (nx, ny, ntips, 3) forces with bF3d."""
    ntips, nx, ny = len(zTips), len(xTips), len(yTips)
    if trj is None:
        trj = np.zeros((ntips, 3))
        trj[:, 2] = np.asarray(zTips)[::-1]
    trj = np.asarray(trj, dtype=np.float64)
    rTips = np.empty((nx, ny, ntips, 3))
    rTips[..., 0] = trj[None, None, :, 0] + np.asarray(xTips, dtype=np.float64)[:, None, None]
    rTips[..., 1] = trj[None, None, :, 1] + np.asarray(yTips, dtype=np.float64)[None, :, None]
    rTips[..., 2] = trj[None, None, :, 2]
    rs = np.zeros_like(rTips)
    fs = np.zeros_like(rTips)
    core.relaxTipStrokes_omp(rTips, rs, fs, tip_spline=tip_spline)
    PPpos = rs[:, :, ::-1, :].copy()
    if bF3d:
        fzs = fs[:, :, ::-1, :].copy()
    else:
        fzs = np.ascontiguousarray(fs[:, :, ::-1, 2].transpose(2, 1, 0))
    return fzs, PPpos


def relaxedScan3D_omp(xTips, yTips, zTips, trj=None, bF3d=False, tip_spline=None):
    """All columns in one GPU launch (HighLevel.py:104-124); like the
    reference it ignores ``trj``."""
    nx, ny, nz = len(xTips), len(yTips), len(zTips)
    rTips = np.zeros((nx, ny, nz, 3))
    rs = np.zeros((nx, ny, nz, 3))
    fs = np.zeros((nx, ny, nz, 3))
    rTips[:, :, :, 0] = np.asarray(xTips)[:, None, None]
    rTips[:, :, :, 1] = np.asarray(yTips)[None, :, None]
    rTips[:, :, :, 2] = np.asarray(zTips)[::-1][None, None, :]
    core.relaxTipStrokes_omp(rTips, rs, fs, tip_spline=tip_spline)
    rs = rs[:, :, ::-1, :].copy()
    fzs = fs[:, :, ::-1, :].copy() if bF3d else fs[:, :, ::-1, 2].copy()
    return fzs, rs


def setFF(FF=None, computeVpot=False, n=None, lvec=None, parameters=None):
    """HighLevel.py:190-257."""
    if FF is not None:
        gridN = np.array(np.shape(FF)[0:3], dtype=np.int32)
        if parameters is not None:
            parameters.gridN = gridN
        if n is not None:
            n[0:3] = gridN
    else:
        if n is not None:
            gridN = np.array(n, dtype=np.int32)
        elif parameters is not None:
            gridN = np.array(PPU.autoGridN(parameters) if parameters.gridN[0] <= 0 else parameters.gridN, dtype=np.int32)
        else:
            raise ValueError("FF dimensions not set !!")
        FF = np.zeros((gridN[0], gridN[1], gridN[2], 3))
    core.setGridN(gridN)
    scalar = False
    if len(FF.shape) == 4 and FF.shape[-1] == 3:
        core.setFF_Fpointer(FF)
    elif len(FF.shape) == 3 or FF.shape[-1] == 1:
        core.setFF_Epointer(FF)
        computeVpot = False
        scalar = True
    else:
        raise ValueError("setFF: Array dimensions wrong for both vector and array field !!")
    if computeVpot:
        V = np.zeros((gridN[0], gridN[1], gridN[2]))
        core.setFF_Epointer(V)
    else:
        V = None
        if not scalar:
            core.deleteFF_Epointer()  # no stale energy grid from an earlier call
    if (lvec is None) and (parameters is not None):
        lvec = np.array([parameters.gridA, parameters.gridB, parameters.gridC], dtype=np.float64)
    lvec = np.array(lvec, dtype=np.float64)
    if lvec.shape == (4, 3):
        lvec = lvec[1:, :].copy()
    elif lvec.shape != (3, 3):
        raise ValueError("lvec matrix has a wrong format !!")
    core.setGridCell(lvec)
    return FF, V


def _parse_atoms(atoms, elem_dict_unused, parameters, lvec):
    Zs = np.array(atoms[0], dtype=np.int32)
    Rs = np.array([atoms[1], atoms[2], atoms[3]], dtype=np.float64).T.copy()
    Qs = np.array(atoms[4], dtype=np.float64)
    if parameters.PBC:
        Zs, Rs, Qs = PPU.PBCAtoms(Zs, Rs, Qs, avec=lvec[1], bvec=lvec[2], na=parameters.nPBC[0], nb=parameters.nPBC[1])
    return Zs, Rs, Qs


def computeLJ(geomFile, speciesFile=None, geometry_format=None, save_format=None, computeVpot=False, Fmax=Fmax_DEFAULT, Vmax=Vmax_DEFAULT, ffModel="LJ", parameters=None):
    """LJ, Morse or vdW force field of a geometry file (HighLevel.py:260-307):
    ffModel "Morse" uses getMorseFF with parameters.aMorse; "vdW" uses
    getVdWFF for parameters.vdWDampKind 0 and getVdWFF_RE with that kind
    otherwise; anything else is LJ.  With save_format ("xsf" or "npy") the
    grids are written as FF<model> (and E<model>) like the reference."""
    FFparams = PPU.loadSpecies(speciesFile)
    atoms, nDim, lvec = io.loadGeometry(geomFile, format=geometry_format, parameters=parameters)
    iZs, Rs, Qs = _parse_atoms(atoms, None, parameters, lvec)
    iPP = FFparams.by_symbol.get(parameters.probeType) if not str(parameters.probeType).isdigit() else int(parameters.probeType)
    FF, V = setFF(None, computeVpot, lvec=lvec, parameters=parameters)
    Rs0 = shift_positions(Rs, -lvec[0])
    if ffModel == "Morse":
        core.getMorseFF(Rs0, PPU.getAtomsRE(iPP, iZs, FFparams), parameters=parameters)
    elif ffModel == "vdW":
        if parameters.vdWDampKind == 0:
            core.getVdWFF(Rs0, PPU.getAtomsLJ(iPP, iZs, FFparams))
        else:
            core.getVdWFF_RE(Rs0, PPU.getAtomsRE(iPP, iZs, FFparams), kind=parameters.vdWDampKind)
    else:
        core.getLennardJonesFF(Rs0, PPU.getAtomsLJ(iPP, iZs, FFparams))
    if Fmax is not None:
        io.limit_vec_field(FF, Fmax=Fmax)
    if (Vmax is not None) and computeVpot:
        V[V > Vmax] = Vmax
    if save_format is not None:
        _save_fields("FF" + ffModel, "E" + ffModel, FF, V if computeVpot else None, atoms, lvec, save_format)
    return FF, V, nDim, lvec


def _save_fields(fname, ename, FF, V, atoms, lvec, save_format):
    # HighLevel.py:301-305, 381-385: the header carries the cell and atoms
    head = io.primcoords2Xsf(np.asarray(atoms[0], dtype=np.int32), [atoms[1], atoms[2], atoms[3]], lvec)
    info = (atoms[:4], lvec)
    io.save_vec_field(fname, FF, lvec, data_format=save_format, head=head, atomic_info=info)
    if V is not None:
        io.save_scal_field(ename, V, lvec, data_format=save_format, head=head, atomic_info=info)


def computeELFF_pointCharge(geomFile, geometry_format=None, tip="s", save_format=None, computeVpot=False, Fmax=Fmax_DEFAULT, Vmax=Vmax_DEFAULT, parameters=None):
    """Unit-probe-charge Coulomb field of a geometry file (HighLevel.py:353-387);
    with save_format the grids are written as FFel (and Vel)."""
    tipKind = {"s": 0, "pz": 1, "dz2": 2}[tip]
    atoms, nDim, lvec = io.loadGeometry(geomFile, format=geometry_format, parameters=parameters)
    _, Rs, Qs = _parse_atoms(atoms, None, parameters, lvec)
    FF, V = setFF(None, computeVpot, lvec=lvec, parameters=parameters)
    Rs0 = shift_positions(Rs, -lvec[0])
    core.getCoulombFF(Rs0, np.ascontiguousarray(Qs * PPU.CoulombConst), kind=tipKind)
    if Fmax is not None:
        io.limit_vec_field(FF, Fmax=Fmax)
    if (Vmax is not None) and computeVpot:
        V[V > Vmax] = Vmax
    if save_format is not None:
        _save_fields("FFel", "Vel", FF, V if computeVpot else None, atoms, lvec, save_format)
    return FF, V, nDim, lvec


FF = None


def perform_relaxation(lvec, FFLJ, FFel=None, FFpauli=None, FFboltz=None, FFkpfm_t0sV=None, FFkpfm_tVs0=None, tip_spline=None, bPPdisp=False, bFFtotDebug=False, parameters=None):
    """Relaxed scan over the parameters' scan window (HighLevel.py:127-184)."""
    global FF
    xTips, yTips, zTips, lvecScan = PPU.prepareScanGrids(parameters=parameters)
    FF = FFLJ.copy()
    if FFel is not None:
        FF += FFel * parameters.charge
    if FFkpfm_t0sV is not None and FFkpfm_tVs0 is not None:
        FF += (parameters.charge * FFkpfm_t0sV - FFkpfm_tVs0) * parameters.Vbias
    if FFpauli is not None:
        FF += FFpauli * parameters.Apauli
    if FFboltz is not None:
        FF += FFboltz
    FF = np.ascontiguousarray(FF)
    setFF(FF, lvec=lvec, parameters=parameters)
    if (np.array(parameters.stiffness) < 0.0).any():
        parameters.stiffness = np.array([parameters.klat, parameters.klat, parameters.krad])
    core.setTip(kSpring=np.array((parameters.stiffness[0], parameters.stiffness[1], 0.0)) / -PPU.eVA_Nm, kRadial=parameters.stiffness[2] / -PPU.eVA_Nm, parameters=parameters)
    trj = None
    if parameters.tiltedScan:
        trj = trjByDir(len(zTips), d=parameters.scanTilt, p0=[0.0, 0.0, parameters.scanMin[2] - lvec[0, 2]])
    fzs, PPpos = relaxedScan3D_omp(xTips - lvec[0, 0], yTips - lvec[0, 1], zTips - lvec[0, 2], trj=trj, bF3d=parameters.tiltedScan, tip_spline=tip_spline)
    PPpos += np.asarray(lvec[0], dtype=np.float64)[None, None, None, :]
    if bPPdisp:
        init_pos = np.array(np.meshgrid(xTips, yTips, zTips)).transpose(3, 1, 2, 0) + np.array([parameters.r0Probe[0], parameters.r0Probe[1], -parameters.r0Probe[2]])
        PPdisp = PPpos - init_pos
    else:
        PPdisp = None
    core.deleteFF_Fpointer()
    return fzs, PPpos, PPdisp, lvecScan


# --------------------------------------------------------- core densities


def loadValenceElectronDict():
    """Valence electrons per atomic number (HighLevel.py:414-432).  The
    reference executes a local valelec_dict.py when one exists; this package
    runs no file and returns its built-in table (common.VALENCE_ELECTRONS):
    pass valElDict to override it."""
    return dict(PPU.VALENCE_ELECTRONS)


def _getAtomsWhichTouchPBCcell(Rs, elems, nDim, lvec, Rcut, bSaveDebug=False, fname=None):
    """Periodic images of the atoms Rs (3, N) within Rcut of the cell
    (HighLevel.py:435-441): (positions (M, 3), their elements)."""
    inds, Rs_ = PPU.findPBCAtoms3D_cutoff(Rs, np.array(lvec[1:]), Rcut=Rcut)
    elems = [elems[i] for i in inds]
    return Rs_.transpose().copy(), elems


def subtractCoreDensities(rho, lvec, elems=None, Rs=None, fname=None, valElDict=None, Rcore=0.7, bSaveDebugDens=False, bSaveDebugGeom=True,
                          head=io.XSF_HEAD_DEFAULT, parameters=None):
    """Subtract the valence charge of each atom, spread as the normalised
    (1 - r^2/Rcore^2)^2 density, from the density grid rho (float64
    (nx,ny,nz), in place) (HighLevel.py:452-478).  The projection runs on
    the GPU (getDensityR4spline)."""
    if fname is not None:
        # getAtomsWhichTouchPBCcell (HighLevel.py:444-449) on an xyz geometry
        atoms, nDim_g, lvec_g = io.loadGeometry(fname, parameters=parameters)
        Rs, elems = _getAtomsWhichTouchPBCcell(np.array(atoms[1:4]), np.array(atoms[0]), nDim_g, lvec_g, Rcore)
    if valElDict is None:
        valElDict = loadValenceElectronDict()
    cRAs = np.array([(-valElDict[elem], Rcore) for elem in elems], dtype=np.float64).reshape(-1, 2)
    setFF(rho, lvec=lvec, parameters=parameters)
    core.getDensityR4spline(shift_positions(Rs, -np.asarray(lvec)[0]), cRAs.copy())
    if bSaveDebugDens:
        io.saveXSFData("rho_subCoreChg.xsf", rho, lvec, head=head)
    return rho

