"""GPU side of the grid files (ppafm/io.py:575-880): a device grid written
as XSF straight from the GPU copy, read back with io.load_vec_field, fed to
the GPU as the grid of a new scan, and relaxed -- against the oracle
relaxing the same loaded field; and HighLevel.computeLJ /
computeELFF_pointCharge writing their grids like the reference.

Bars: the written file equals the reference's formatting of the device
values ("%10.5e", Python's rounding); the loaded field equals those values;
positions from the loaded field within 1e-3 A of the oracle's (the field is
FP64 on both sides, rounded to 6 digits by the file)."""

import numpy as np
import pytest

import oracle as O
from cases import small_config, tip_params
from ppafm_amd import HighLevel as PPH
from ppafm_amd import common as C
from ppafm_amd import io
from ppafm_amd import molgen as M
from ppafm_amd.scan import PlainScan

pytestmark = pytest.mark.gpu


def test_device_grid_xsf_round_trip_and_relax(tmp_path):
    cfg, seed = small_config(scan_n=(12, 12), nz=10)
    print("case: small molecule (seed 7), 64x64x80 grid -> xsf -> device grid, 12x12x10 scan", flush=True)
    xyz, Zs, qs = M.molecule(cfg["natoms"], seed)
    lvec = M.config_lvec(cfg)
    sc = PlainScan(lvec, cfg["n"], cfg, tipR0=M.TIP_R0, k_lat=M.K_LAT, k_rad=M.K_RAD, probe_q=M.PROBE_Q)
    sc.set_molecule(xyz, Zs, qs)
    sc.build_grid()
    base = str(tmp_path / "FF")
    sc.save_grid(base)
    G = sc.grid_host()
    for k, c in enumerate("xyz"):
        text = open(f"{base}_{c}.xsf").read()
        data_part = text[text.index("BEGIN_DATAGRID_3D_whatever") :].split("\n", 6)[6]
        assert data_part == O.xsf_data_text(G[..., k]), c
    FF, lv, nDim, head = io.load_vec_field(base)
    rounded = np.vectorize(lambda v: float("%10.5e" % v))(G[..., :3])
    assert np.array_equal(FF, rounded) and np.array_equal(lv, lvec)
    # the loaded field as the grid of a fresh scan
    sc2 = PlainScan(lvec, cfg["n"], cfg, tipR0=M.TIP_R0, k_lat=M.K_LAT, k_rad=M.K_RAD, probe_q=M.PROBE_Q)
    sc2.set_grid(FF)
    sc2.relax()
    pos = sc2.positions()
    tips = np.ascontiguousarray(M.scan_tips(cfg) - lvec[0])
    rr, fr, it, tot = O.relax_strokes(cfg["n"], lvec[1:], np.ascontiguousarray(FF), tip_params(), tips)
    d = np.abs(pos - (rr + lvec[0])).max()
    print(f"relaxed on the loaded field: max |dr| vs oracle {d:.3g} A")
    assert d < 1e-3


def test_highlevel_saves_like_the_reference(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    xyz, Zs, qs = M.molecule(10, 3)
    xyz = xyz + np.array([2.0, 2.0, 2.0])
    with open("m.xyz", "w") as f:
        f.write(f"{len(Zs)}\n")
        f.write('Lattice="4.0 0.0 0.0 0.0 4.0 0.0 0.0 0.0 4.0"\n')
        sym = {1: "H", 6: "C", 7: "N", 8: "O"}
        for z, p, q in zip(Zs, xyz, qs):
            f.write(f"{sym[int(z)]} {p[0]:.6f} {p[1]:.6f} {p[2]:.6f} {q:.6f}\n")
    p = C.PpafmParameters(gridN=[20, 20, 20], PBC=False, probeType="O")
    FF, V, nDim, lvec = PPH.computeLJ("m.xyz", computeVpot=True, save_format="xsf", parameters=p)
    F2, lv, nd, head = io.load_vec_field("FFLJ")
    assert np.array_equal(F2, np.vectorize(lambda v: float("%10.5e" % v))(FF))
    Zh, Rh = io.getFromHead_PRIMCOORD(head)
    assert list(Zh) == list(Zs)
    E2 = io.load_scal_field("ELJ")[0]
    assert np.array_equal(E2, np.vectorize(lambda v: float("%10.5e" % v))(V))
    FE, VE, _, _ = PPH.computeELFF_pointCharge("m.xyz", save_format="npy", computeVpot=True, parameters=p)
    F3, lv3, nd3, info = io.load_vec_field("FFel", data_format="npy")
    assert np.array_equal(F3, FE) and np.array_equal(lv3, lvec)
    assert np.array_equal(io.load_scal_field("Vel", data_format="npy")[0], VE)
