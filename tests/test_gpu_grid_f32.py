"""The FP32 grid kernel (ff_grid_f32.hip: every pair in FP32, FP64 partial
sums, FP64 only where an FP32 pair overflows) against the FP64 oracle.

* OpenCL model (evalLJC_QZs_noPos, FF.cl:366-423; the reference computes it
  in FP32): the AFMulator's dz2-like tip charges (3 stencil points) on an
  orthogonal grid, and a monopole probe on a sheared cell;
* ProbeParticle model (Forces.h:120-132, 237-244, with the separate clamps
  of io.limit_vec_field) with fp64_radius 0;
* the overflow guard: an atom 0.0045 A from a grid node, where FP32
  intermediates of the repulsion overflow but the force itself is finite;
  an atom exactly on a node in the clamped ProbeParticle model;
* batches with uncharged atoms equal single-grid launches bit for bit.

The FP32 error is reported as max and p99 of |dF| / max(|F|, 1e-3 eV/A);
the bars: p99 below 3e-5 and max below 1e-3 (FP32 pairs lose ~1e-7 of each
term, more in the r^-14 repulsion, and points where large terms cancel
reach ~1e-4 at the floor); the guarded points match the oracle to 1e-6."""

import numpy as np
import pytest

import oracle as O
from cases import grid_rel_err
from ppafm_amd import batch as B
from ppafm_amd import common as C
from ppafm_amd import device as D
from ppafm_amd import molgen as M

pytestmark = pytest.mark.gpu

QS_DZ2 = np.array([-10.0, 20.0, -10.0, 0.0])
QZS_DZ2 = np.array([0.1, 0.0, -0.1, 0.0])


def _ocl_atoms(natoms, seed, shift):
    xyz, Zs, qs = M.molecule(natoms, seed)
    xyz = xyz + shift
    cLJs = C.REA2LJ(C.getAtomsREA(8, Zs, C.loadSpecies(), alphaFac=-1.0))
    return xyz, qs, cLJs


def _ocl_grid(n, lvec, xyz, qs, cLJs, Qs, QZs):
    model = D.ff_model(1, with_lj=True, stencil=list(zip(Qs, QZs)), with_energy=True, fmax=None)
    assert model.fp64_radius == 0.0  # the OpenCL model's float4 grid is FP32
    xyzq = np.concatenate([xyz - lvec[0], qs[:, None]], 1)
    G = D.ff_build(n, lvec[1:], xyzq, cLJs, model, f64=False).download()
    FE = O.ljc_qzs_ff(n, lvec, np.concatenate([xyz, qs[:, None]], 1), cLJs, Qs, QZs)
    return G, FE


def _report(name, G, FE):
    err = grid_rel_err(G, FE)
    de = np.abs(G[..., 3] - FE[..., 3]) / np.maximum(np.abs(FE[..., 3]), 1e-3)
    print(f"{name}: |dF|/max(|F|,1e-3) max {err.max():.3g} p99 {np.percentile(err, 99):.3g} p50 {np.median(err):.3g}; "
          f"E max {de.max():.3g} p99 {np.percentile(de, 99):.3g}", flush=True)
    return err, de


def test_ocl_model_dz2_tip_orthogonal():
    print("case: OpenCL model, dz2 tip charges, 40-atom molecule, 96x96x80 grid", flush=True)
    n = (96, 96, 80)
    lvec = np.array([[-4.8, -4.8, -3.0], [9.6, 0, 0], [0, 9.6, 0], [0, 0, 8.0]])
    xyz, qs, cLJs = _ocl_atoms(40, 5, np.zeros(3))
    G, FE = _ocl_grid(n, lvec, xyz, qs, cLJs, QS_DZ2, QZS_DZ2)
    assert np.isfinite(G).all()
    err, de = _report("dz2 tip", G, FE)
    assert np.percentile(err, 99) < 3e-5 and err.max() < 1e-3
    assert np.percentile(de, 99) < 3e-5


def test_ocl_model_monopole_sheared_cell():
    print("case: OpenCL model, monopole probe, sheared cell, 30 atoms, 72x64x80 grid", flush=True)
    n = (72, 64, 80)
    lvec = np.array([[-4.0, -3.5, -3.0], [7.2, 0, 0], [1.2, 6.4, 0], [0, 0.4, 8.0]])
    xyz, qs, cLJs = _ocl_atoms(30, 9, np.array([0.3, 0.1, 0.0]))
    G, FE = _ocl_grid(n, lvec, xyz, qs, cLJs, np.array([-0.1, 0, 0, 0]), np.zeros(4))
    assert np.isfinite(G).all()
    err, de = _report("monopole, sheared", G, FE)
    assert np.percentile(err, 99) < 3e-5 and err.max() < 1e-3


def _pp_grid(n, lvec, xyz, Zs, qs, fp64_radius=0.0):
    cLJs = C.getAtomsLJ(M.PROBE_Z, Zs, C.loadSpecies())
    model = D.ff_model(0, with_lj=True, stencil=D.PROBE_STENCILS["s"], with_energy=True,
                       coulomb_const=C.CoulombConst, probe_q=M.PROBE_Q, fmax=10.0, fp64_radius=fp64_radius)
    xyzq = np.concatenate([xyz - lvec[0], qs[:, None]], 1)
    G = D.ff_build(n, lvec[1:], xyzq, cLJs, model, f64=False).download()
    Rs0 = xyz - lvec[0]
    FL = O.limit_vec_field(O.lj_ff(n, lvec[1:], Rs0, cLJs), 10.0)
    FEl = O.limit_vec_field(O.coulomb_ff(n, lvec[1:], Rs0, qs * C.CoulombConst), 10.0)
    return G, FL + M.PROBE_Q * FEl


def test_pp_model_fp32():
    print("case: ProbeParticle model, FP32 pairs (fp64_radius 0), 36 atoms, 80x80x96 grid", flush=True)
    n = (80, 80, 96)
    lvec = np.array([[-4.0, -4.0, -3.0], [8.0, 0, 0], [0, 8.0, 0], [0, 0, 9.6]])
    xyz, Zs, qs = M.molecule(36, 4)
    G, FF = _pp_grid(n, lvec, xyz, Zs, qs)
    assert np.isfinite(G).all()
    err = grid_rel_err(G, FF)
    print(f"ProbeParticle FP32: max {err.max():.3g} p99 {np.percentile(err, 99):.3g} p50 {np.median(err):.3g}")
    assert np.percentile(err, 99) < 3e-5 and err.max() < 1e-3


def test_overflow_guard_ocl():
    print("case: OpenCL model, an atom 0.0045 A from a grid node", flush=True)
    n = (32, 32, 32)
    lvec = np.array([[0.0, 0.0, 0.0], [3.2, 0, 0], [0, 3.2, 0], [0, 0, 3.2]])
    node = np.array([1.6, 1.6, 1.6])
    off = 0.0045 / np.sqrt(3) * np.ones(3)
    xyz = np.array([node + off, [0.7, 2.2, 1.05]])
    Zs = np.array([79, 6])
    qs = np.array([0.1, -0.1])
    cLJs = C.REA2LJ(C.getAtomsREA(8, Zs, C.loadSpecies(), alphaFac=-1.0))
    G, FE = _ocl_grid(n, lvec, xyz, qs, cLJs, np.array([-0.1, 0, 0, 0]), np.zeros(4))
    i = (16, 16, 16)
    print(f"at the node: GPU {G[i]} oracle {FE[i]}")
    assert np.isfinite(FE[i]).all() and np.abs(FE[i][:3]).max() < 3e38
    assert np.isfinite(G[i]).all()
    assert np.abs(G[i] - FE[i]).max() <= 1e-6 * np.abs(FE[i]).max()
    m = np.ones(n, bool)
    m[i] = False
    fin = np.isfinite(FE[..., :3]).all(-1) & (np.abs(FE[..., :3]).max(-1) < 3e38)
    err = grid_rel_err(G, FE)[m & fin]
    assert err.max() < 1e-3


def test_overflow_guard_pp_atom_near_node():
    """An atom 0.001 A from a grid node: the FP32 repulsion term 12 c12 r^-14
    overflows there (r^-2 ~ 1e6), the guard sums the point again in FP64,
    and the clamped fields agree with the oracle to FP64 accuracy (FP32
    pairs would leave the direction of a 0.001 A offset wrong at 1e-4).
    (An atom exactly on a node is a singular point: the pair's force there
    is 0 * inf, and its direction depends on 1e-16 offsets of the node's
    coordinates, so no result pins it.)"""
    print("case: ProbeParticle model, an atom 0.001 A from a grid node (clamped fields)", flush=True)
    n = (32, 32, 32)
    lvec = np.array([[0.0, 0.0, 0.0], [3.2, 0, 0], [0, 3.2, 0], [0, 0, 3.2]])
    node = np.array([1.6, 1.6, 1.6])
    xyz = np.array([node + 0.001 / np.sqrt(3) * np.array([1.0, -1.0, 1.0]), [0.75, 2.3, 1.1]])
    G, FF = _pp_grid(n, lvec, xyz, np.array([6, 8]), np.array([0.2, -0.2]))
    # the clamped forces stay finite; the energy at the node (~1e52 eV)
    # exceeds the float range of the storage
    assert np.isfinite(G[..., :3]).all()
    err = grid_rel_err(G, FF)
    print(f"atom near a node: max {err.max():.3g}; at the node GPU {G[16, 16, 16]} oracle {FF[16, 16, 16]}")
    assert err[16, 16, 16] < 1e-6
    assert err.max() < 1e-3


def test_batch_with_uncharged_atoms_equals_single_grids():
    print("case: OpenCL-model FP32 batch of 3 molecules with uncharged atoms vs single grids", flush=True)
    lvec = np.array([[-3.2, -3.2, -3.0], [6.4, 0, 0], [0, 6.4, 0], [0, 0, 6.4]])
    n = (64, 64, 64)
    mols = []
    for k in range(3):
        xyz, Zs, qs = M.molecule(12 + 9 * k, 40 + k)
        qs = qs.copy()
        qs[::3] = 0.0
        mols.append((xyz, Zs, qs))
    b = B.BatchScan.tilted([lvec] * 3, [n] * 3, pos0s=[(0.0, 0.0, 6.0)] * 3, start=(-1, -1), end=(1, 1), scan_n=(8, 8),
                           nz=10, molecules=mols, grid_f64=False)
    b.run_stages(1)
    b.synchronize()
    for k, (xyz, Zs, qs) in enumerate(mols):
        cLJs = C.REA2LJ(C.getAtomsREA(8, Zs, C.loadSpecies(), alphaFac=-1.0))
        xyzq = np.concatenate([xyz - lvec[0], qs[:, None]], 1)
        G = D.ff_build(n, lvec[1:], xyzq, cLJs, b.model, f64=False).download()
        assert np.array_equal(b.grid(k), G), k
    b.close()
