"""Seeded parity cases shared by the tests: a molecule from the benchmark
generator, its FP64 oracle grids and the matching scan geometry."""

import functools

import numpy as np

import oracle as O
from ppafm_amd import common as C
from ppafm_amd import molgen as M

RELAX_DEFAULT = O.DEFAULT_RELAX


def small_config(n=(64, 64, 80), scan_n=(24, 24), nz=20, natoms=20, seed=7):
    """A scaled-down config-0-like case (0.1 A grid step)."""
    nx, ny, nzg = n
    cfg = dict(natoms=natoms, origin=(-nx * 0.05, -ny * 0.05, -3.0), cell=(nx * 0.1, ny * 0.1, nzg * 0.1), n=n,
               scan_n=scan_n, scan_x=(-1.2, 1.2), scan_y=(-1.2, 1.2), z_top=7.0, nz=nz, dz=0.1)
    return cfg, seed


@functools.lru_cache(maxsize=8)
def case(config_id, small=False):
    if small:
        cfg, seed = small_config()
        xyz, Zs, qs = M.molecule(cfg["natoms"], seed)
    else:
        cfg = M.CONFIGS[config_id]
        xyz, Zs, qs = M.config_molecule(config_id, 0, cfg["natoms"])
    lvec = M.config_lvec(cfg)
    cLJs = C.getAtomsLJ(M.PROBE_Z, Zs, C.loadSpecies())
    Rs0 = xyz - lvec[0]
    n = cfg["n"]
    FL = O.lj_ff(n, lvec[1:], Rs0, cLJs)
    FE = O.coulomb_ff(n, lvec[1:], Rs0, qs * C.CoulombConst)
    FLc = O.limit_vec_field(FL.copy(), 10.0)
    FEc = O.limit_vec_field(FE.copy(), 10.0)
    FF = FLc + M.PROBE_Q * FEc
    return dict(cfg=cfg, xyz=xyz, Zs=Zs, qs=qs, lvec=lvec, cLJs=cLJs, Rs0=Rs0, FL=FL, FE=FE, FF=FF)


def tip_params():
    return O.tip_params(M.TIP_R0[2], M.K_RAD / -C.eVA_Nm, (0.0, 0.0, 0.0), np.array([M.K_LAT, M.K_LAT, 0.0]) / -C.eVA_Nm)


def grid_rel_err(F, Fref, floor=1e-3):
    """|dF| / max(|Fref|, floor) on the vector norm, per point."""
    d = np.linalg.norm(F[..., :3] - Fref[..., :3], axis=-1)
    return d / np.maximum(np.linalg.norm(Fref[..., :3], axis=-1), floor)
