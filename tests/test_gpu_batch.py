"""GPU tests of the batched path (one captured graph for many molecules).

The batch runs the same per-brick and per-column device code as the
single-molecule path, so every output of every item must equal the
single-molecule PlainScan / AFMulator result bit for bit; one item is also
checked against the FP64 oracle at the plain path's tolerances (positions
1e-3 A, df 1e-3 Hz).  Items differ in atom count, grid size, scan size and
number of z steps (two df groups)."""

import numpy as np
import pytest

import oracle as O
from cases import small_config, tip_params
from ppafm_amd import batch as B
from ppafm_amd import common as C
from ppafm_amd import molgen as M
from ppafm_amd._lib import HipError
from ppafm_amd.scan import PlainScan

pytestmark = pytest.mark.gpu

# (grid n, scan (nx, ny), nz, natoms, seed)
SPECS = [
    ((64, 64, 80), (24, 24), 20, 20, 7),
    ((48, 56, 72), (17, 9), 20, 8, 8),
    ((72, 64, 64), (30, 21), 16, 33, 9),
    ((64, 48, 80), (8, 40), 16, 11, 10),
    ((56, 56, 88), (13, 13), 20, 47, 11),
]


def _cases():
    out = []
    for n, scan_n, nz, natoms, seed in SPECS:
        cfg, _ = small_config(n=n, scan_n=scan_n, nz=nz, natoms=natoms, seed=seed)
        out.append((cfg, M.config_lvec(cfg), M.molecule(natoms, seed)))
    return out


def _plain(cfg, lvec, mol, **kw):
    s = PlainScan(lvec, cfg["n"], cfg, tipR0=M.TIP_R0, k_lat=M.K_LAT, k_rad=M.K_RAD, iZPP=M.PROBE_Z, probe_q=M.PROBE_Q,
                  amplitude=M.DF_AMPLITUDE, **kw)
    s.set_molecule(*mol)
    s.run()
    return s


def _batch(cases, **kw):
    return B.BatchScan.plain([c[1] for c in cases], [c[0]["n"] for c in cases], [c[0] for c in cases], tipR0=M.TIP_R0,
                             k_lat=M.K_LAT, k_rad=M.K_RAD, iZPP=M.PROBE_Z, probe_q=M.PROBE_Q,
                             amplitude=M.DF_AMPLITUDE, molecules=[c[2] for c in cases], keep_iters=True, **kw)


@pytest.fixture(scope="module")
def run():
    print("case: plain batch of 5 molecules (8..47 atoms, 4 grid sizes, nz 20/16)", flush=True)
    cases = _cases()
    b = _batch(cases)
    b.run()
    b.synchronize()
    singles = [_plain(*c) for c in cases]
    return cases, b, singles


def test_batch_equals_single_scans(run):
    cases, b, singles = run
    for k, s in enumerate(singles):
        assert np.array_equal(b.grid(k), s.grid_host()), k
        assert np.array_equal(b.positions_raw(k), s.pos.download()), k
        assert np.array_equal(b.forces(k), s.forces()), k
        assert np.array_equal(b.iterations(k), s.iterations()), k
        assert np.array_equal(b.status(k), s.status()), k
        # mode 0: a capped step reports max_iters, a converged one fewer
        assert np.array_equal(b.status(k), (b.iterations(k) >= b.rp.max_iters).astype(np.uint8)), k
        assert np.array_equal(b.df_host(k), s.df_host()), k
    ne, npos = b.counts()
    assert npos == b.total_positions()
    assert ne == sum(s.counts()[0] for s in singles)
    # evaluations over lane slots of the waves' loop passes
    assert 0 < b.lane_use() <= 1


def test_batch_item_against_oracle(run):
    cases, b, _ = run
    k = 4
    cfg, lvec, (xyz, Zs, qs) = cases[k]
    cLJs = C.getAtomsLJ(M.PROBE_Z, Zs, C.loadSpecies())
    Rs0 = xyz - lvec[0]
    FL = O.limit_vec_field(O.lj_ff(cfg["n"], lvec[1:], Rs0, cLJs), 10.0)
    FE = O.limit_vec_field(O.coulomb_ff(cfg["n"], lvec[1:], Rs0, qs * C.CoulombConst), 10.0)
    FF = FL + M.PROBE_Q * FE
    tips = M.scan_tips(cfg) - lvec[0]
    rs, fs, it, _ = O.relax_strokes(cfg["n"], lvec[1:], FF, tip_params(), tips)
    dr = np.abs(b.positions(k) - lvec[0] - rs).max()
    df_ref = O.fz2df(fs[:, :, ::-1, 2], 0.1, M.K0, M.F0, amplitude=M.DF_AMPLITUDE)
    ddf = np.abs(b.df_host(k) - df_ref).max()
    print(f"batch item {k}: max |dr| {dr:.3g} A, max |ddf| {ddf:.3g} Hz")
    assert dr < 1e-3 and ddf < 1e-3


def test_graph_replays_and_tunings_agree(run):
    cases, b, _ = run
    ref = [(b.positions_raw(k), b.forces(k), b.df_raw(k)) for k in range(b.nitems)]
    c0 = b.counts()
    for tuning in ({"tile_y": 4, "waves_per_block": 2, "blocks_per_cu": 3, "refill": 0},
                   {"tile_y": 16, "waves_per_block": 4, "blocks_per_cu": 1, "refill": 0},
                   {"tile_y": 1, "waves_per_block": 1, "blocks_per_cu": 2, "refill": 0},
                   {"tile_y": 8, "waves_per_block": 1, "blocks_per_cu": 16, "refill": 1}):
        b.set_tuning(tuning)
        assert b.tuning == tuning
        for rep in range(2):
            b.run()
            b.synchronize()
            # eager work between replays
            b.run_stages(4)
            b.synchronize()
            assert b.counts() == c0, (tuning, rep)
            for k in range(b.nitems):
                p, f, d = ref[k]
                assert np.array_equal(b.positions_raw(k), p), (tuning, rep, k)
                assert np.array_equal(b.forces(k), f), (tuning, rep, k)
                assert np.array_equal(b.df_raw(k), d), (tuning, rep, k)


def test_autotuner_and_cache(run, tmp_path):
    cases, b, _ = run
    ref = [b.positions_raw(k) for k in range(b.nitems)]
    best, ms = b.tune(nsample=3, reps=1)
    assert best["tile_y"] in (2, 4, 8, 16) and 1 <= best["waves_per_block"] <= 4 and best["blocks_per_cu"] >= 1
    assert best["refill"] == 0  # the items differ in geometry: no lane refill
    assert ms > 0
    assert b.tuning == best
    b.run()
    b.synchronize()
    for k in range(b.nitems):
        assert np.array_equal(b.positions_raw(k), ref[k])
    path = tmp_path / "tuning.json"
    t1 = B.tuned(b, nsample=2, reps=1, path=path)
    assert path.exists()
    t2 = B.tuned(b, nsample=2, reps=1, path=path)  # cache hit
    assert t1 == t2


def test_batch_errors(run):
    cases, b, _ = run
    with pytest.raises(ValueError):
        b.set_atoms([np.zeros((3, 6))] * (b.nitems + 1))
    with pytest.raises(HipError):
        b.set_tuning({"tile_y": 3, "waves_per_block": 1, "blocks_per_cu": 1})
    assert b.tuning["tile_y"] != 3


def test_new_molecules_with_other_atom_counts():
    """A plan takes molecules of other sizes (fewer, then more atoms than it
    was built for, which grows its atom buffer and captures the graph again)
    and gives the single-molecule results for them."""
    print("case: plain batch of 5, then two sets of other molecules (3..90 atoms)", flush=True)
    cases = _cases()
    b = _batch(cases)
    for seed0, sizes in ((300, (3, 30, 12, 5, 9)), (400, (90, 60, 75, 40, 88))):
        mols = [M.molecule(na, seed0 + k) for k, na in enumerate(sizes)]
        b.set_molecules(mols)
        b.run()
        b.synchronize()
        for k, (c, m) in enumerate(zip(cases, mols)):
            s = _plain(c[0], c[1], m)
            assert np.array_equal(b.grid(k), s.grid_host()), (seed0, k)
            assert np.array_equal(b.positions_raw(k), s.pos.download()), (seed0, k)
            assert np.array_equal(b.df_host(k), s.df_host()), (seed0, k)


def test_afmulator_batch_equals_calls():
    from ppafm_amd.ocl.AFMulator import AFMulator

    print("case: AFMulator.batch of 4 molecules vs __call__ per molecule", flush=True)
    afm = AFMulator(scan_dim=(24, 20, 18), scan_window=((-2.0, -2.0, 7.0), (2.0, 2.0, 8.8)), df_steps=8)
    mols = [M.molecule(n, 100 + n) for n in (8, 20, 35, 12)]
    Xs = afm.batch(mols)
    for m, X in zip(mols, Xs):
        Xr = afm(*m)
        assert X.shape == Xr.shape
        assert np.array_equal(X, Xr)


def test_batch_f32_grid_equals_single_scans():
    print("case: plain batch, float4 grids, 3 molecules", flush=True)
    cases = _cases()[:3]
    b = _batch(cases, grid_f64=False)
    b.run()
    b.synchronize()
    for k, c in enumerate(cases):
        s = _plain(*c, grid_f64=False)
        assert np.array_equal(b.grid(k), s.grid_host()), k
        assert np.array_equal(b.positions_raw(k), s.pos.download()), k
        assert np.array_equal(b.df_host(k), s.df_host()), k


@pytest.mark.parametrize("f64", [True, False])
def test_batch_with_uncharged_atoms_equals_single_scans(f64):
    """A batch holding uncharged atoms runs the grid kernel variant that skips
    their Coulomb terms (exact zeros); the grids and every output still equal
    the single-molecule path's (which evaluates them) bit for bit, on double4
    and float4 grids."""
    print(f"case: 2 molecules with every third charge zero, grid_f64={f64}", flush=True)
    cases = []
    for n, scan_n, nz, natoms, seed in SPECS[:2]:
        cfg, _ = small_config(n=n, scan_n=scan_n, nz=nz, natoms=natoms, seed=seed)
        xyz, Zs, qs = M.molecule(natoms, seed)
        qs = qs.copy()
        qs[::3] = 0.0
        cases.append((cfg, M.config_lvec(cfg), (xyz, Zs, qs)))
    b = _batch(cases, grid_f64=f64)
    b.run()
    b.synchronize()
    for k, c in enumerate(cases):
        s = _plain(*c, grid_f64=f64)
        assert np.array_equal(b.grid(k), s.grid_host()), k
        assert np.array_equal(b.positions_raw(k), s.pos.download()), k
        assert np.array_equal(b.df_host(k), s.df_host()), k
    b.close()


@pytest.mark.parametrize("mode,warm", [("plain", False), ("plain", True), ("tilted", True)])
def test_partial_record_chunks(mode, warm):
    """Warm-started batches stage a lane's output records in LDS and write
    the records of 4 consecutive steps of a column together; a column whose
    step count is not a multiple of 4 ends on a partial chunk.  A 7-step
    scan reproduces the first 7 steps of the same 8-step scan bit for bit
    (positions, forces, iterations, status)."""
    print(f"case: 3 molecules, {mode} batch, warm {warm}, 7 against 8 steps", flush=True)
    mols = [M.molecule(12 + 5 * k, 80 + k) for k in range(3)]
    outs = []
    for nz in (7, 8):
        cfg, _ = small_config(n=(48, 48, 80), scan_n=(20, 13), nz=nz, natoms=20, seed=5)
        cfg = dict(cfg, scan_x=(-1.0, 1.0), scan_y=(-0.9, 0.9))
        lvec = M.config_lvec(cfg)
        if mode == "plain":
            b = B.BatchScan.plain([lvec] * 3, [cfg["n"]] * 3, [cfg] * 3, tipR0=M.TIP_R0, k_lat=M.K_LAT, k_rad=M.K_RAD,
                                  iZPP=M.PROBE_Z, probe_q=M.PROBE_Q, amplitude=M.DF_AMPLITUDE, molecules=mols,
                                  keep_iters=True)
        else:
            t = np.deg2rad(10.0)
            rot = np.array([[np.cos(t), 0, np.sin(t)], [0, 1, 0], [-np.sin(t), 0, np.cos(t)]])
            b = B.BatchScan.tilted([lvec] * 3, [cfg["n"]] * 3, rot=rot, pos0s=[(0.0, 0.0, 6.5)] * 3, start=(-1.0, -0.9),
                                   end=(1.0, 0.9), scan_n=(20, 13), nz=nz, molecules=mols, keep_iters=True)
        if warm:
            b.set_warmstart(None)
        b.run()
        b.synchronize()
        outs.append([(b.positions_raw(k), b.forces(k), b.iterations(k), b.status(k)) for k in range(3)])
        b.close()
    for k in range(3):
        for a7, a8 in zip(outs[0][k], outs[1][k]):
            assert a7.shape[2] == 7 and a8.shape[2] == 8
            assert np.array_equal(a7, a8[:, :, :7]), k


@pytest.mark.parametrize("mode", ["plain", "tilted"])
def test_lane_refill_equals_tile_queue(mode):
    """Items of one geometry (here 24 molecules, a 66x45 scan that leaves
    partial tiles; 71,280 pixel columns, more than the 16,384 lanes of one
    wave per CU, so lanes do take further columns): relaxing with lane
    refill (each lane takes the next column when its own is done, from the
    XCD's queue or from a pool of one or two tiles its wave took) gives
    every output bit for bit as the tile queue does, plain and
    warm-started, with the same counters.  Plain FIRE (columns of ~10^2
    evaluations) also gets a higher lane use from refill; warm-started
    columns end after a few Newton iterates per step, so both schedules leave
    lanes idle about equally there (0.50-0.59 each) and only the outputs are
    compared."""
    NI = 24
    print(f"case: {NI} molecules of one geometry, {mode} batch, lane refill vs tile queue", flush=True)
    cfg, _ = small_config(n=(48, 48, 80), scan_n=(66, 45), nz=8, natoms=20, seed=5)
    cfg = dict(cfg, scan_x=(-1.0, 1.0), scan_y=(-0.9, 0.9))
    lvec = M.config_lvec(cfg)
    mols = [M.molecule(8 + 3 * k, 60 + k) for k in range(NI)]
    if mode == "plain":
        b = B.BatchScan.plain([lvec] * NI, [cfg["n"]] * NI, [cfg] * NI, tipR0=M.TIP_R0, k_lat=M.K_LAT, k_rad=M.K_RAD,
                              iZPP=M.PROBE_Z, probe_q=M.PROBE_Q, amplitude=M.DF_AMPLITUDE, molecules=mols, keep_iters=True)
    else:
        t = np.deg2rad(10.0)
        rot = np.array([[np.cos(t), 0, np.sin(t)], [0, 1, 0], [-np.sin(t), 0, np.cos(t)]])
        b = B.BatchScan.tilted([lvec] * NI, [cfg["n"]] * NI, rot=rot, pos0s=[(0.0, 0.0, 6.5)] * NI, start=(-1.0, -0.9),
                               end=(1.0, 0.9), scan_n=(66, 45), nz=8, molecules=mols, keep_iters=True)
    for warm in (False, True):
        if warm:
            b.set_warmstart(None)
        out = {}
        for refill in (0, 1, 2, 3):
            b.set_tuning({"tile_y": 4, "waves_per_block": 1, "blocks_per_cu": 1, "refill": refill})
            b.run()
            b.synchronize()
            out[refill] = ([(b.positions_raw(k), b.forces(k), b.iterations(k), b.status(k), b.df_raw(k)) for k in range(NI)],
                           b.counters(), b.lane_use())
        for rf in (1, 2, 3):
            for k in range(NI):
                for a, c in zip(out[0][0][k], out[rf][0][k]):
                    assert np.array_equal(a, c), (warm, rf, k)
            assert out[0][1] == out[rf][1], (warm, rf, out[0][1], out[rf][1])
        print(f"warm {warm}: lane use tile queue {out[0][2]:.3f}, lane refill {out[1][2]:.3f}, "
              f"wave pools of one tile {out[2][2]:.3f} and two {out[3][2]:.3f}")
        if not warm:
            assert out[1][2] > out[0][2] and out[2][2] > out[0][2]
    b.close()
