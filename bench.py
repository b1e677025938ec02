"""Benchmark of the probe-particle hot path on one MI355X.

A step is one pass of the path over the 120-atom production workload
(256x256x320 grid at 0.1 A, 256x256 pixels x 60 z-steps, plain FIRE to
|F| < 1e-4 eV/A, 9-tap Giessibl df): force-field grid build, relaxation of
all 3,932,160 tip positions and the Fz -> df convolution, with the atoms
already resident on the GPU.  ``value`` is relaxed tip positions per second
of the whole step (grid build included).

Timing: HIP events recorded on the stream the kernels are launched on,
bracketed by device synchronisation; the relaxation kernel is also timed on
its own for the roofline line (algorithmic bytes from device-side counters
of force evaluations).  The CPU baseline is the reference C++ (compiled by
``make -C oracle ref``; the FP64 oracle port if that is absent) timed with
OpenMP on this host's cores on a bounded sample of the same workload.

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
"""

import argparse
import contextlib
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "relaxed tip positions/s & FF grid pts/s, learned+batched vs plain FIRE"
HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-fp32", dest="fp32", action="store_false", help="skip the FP32 grid legs (c1, batch)")
    p.add_argument("--batch-items", type=int, default=512, help="molecules of the batched workload (0 skips it)")
    p.add_argument("--learned-samples", type=int, default=50_000_000,
                   help="training pairs of the warm-start network on the batch (0 skips the learned comparison)")
    p.add_argument("--learned-epochs", type=float, default=2.0)
    p.add_argument("--c4-instances", type=int, default=16,
                   help="instances of the large tilted case (Au slab + molecule, 400x400x300 grid; 0 skips it)")
    p.add_argument("--generator-molecules", type=int, default=64,
                   help="molecules of the generator comparison, per call vs batched (0 skips it)")
    p.add_argument("--c4-grid", choices=["f64", "f32"], default="f32", help="grid storage of the large tilted case")
    p.add_argument("--no-hartree", action="store_true", help="skip the LJ+Hartree AFMulator leg")
    return p.parse_args()


@contextlib.contextmanager
def quiet_stdout():
    """Silence C-level printf (the reference prints progress) on fd 1."""
    libc = ctypes.CDLL(None)
    sys.stdout.flush()
    libc.fflush(None)
    saved = os.dup(1)
    devnull = os.open(os.devnull, os.O_WRONLY)
    os.dup2(devnull, 1)
    try:
        yield
    finally:
        libc.fflush(None)
        os.dup2(saved, 1)
        os.close(saved)
        os.close(devnull)


def cpu_cores():
    """Every core in this process's affinity set (the OpenMP threads of the
    CPU baselines)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, n)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_gpu(args, cfg, xyz, Zs, qs, lvec, grid_f64=True, fp64_radius=None):
    from ppafm_amd import device as D
    from ppafm_amd import molgen as M
    from ppafm_amd.scan import PlainScan

    s = PlainScan(lvec, cfg["n"], cfg, tipR0=M.TIP_R0, k_lat=M.K_LAT, k_rad=M.K_RAD, iZPP=M.PROBE_Z, probe_q=M.PROBE_Q,
                  amplitude=M.DF_AMPLITUDE, grid_f64=grid_f64, fp64_radius=fp64_radius)
    s.set_molecule(xyz, Zs, qs)
    for _ in range(args.warmup):
        s.run()
    D.synchronize()
    ev = [[D.Event() for _ in range(4)] for _ in range(args.steps)]
    start, stop = D.Event(), D.Event()
    s.reset_counters()
    D.synchronize()
    t0 = time.perf_counter()
    start.record()
    for e in ev:
        e[0].record()
        s.build_grid()
        e[1].record()
        s.relax()
        e[2].record()
        s.to_df()
        e[3].record()
    stop.record()
    D.synchronize()
    wall = time.perf_counter() - t0
    t_total = start.elapsed_ms(stop) / args.steps
    t_grid = float(np.mean([e[0].elapsed_ms(e[1]) for e in ev]))
    t_relax = float(np.mean([e[1].elapsed_ms(e[2]) for e in ev]))
    t_df = float(np.mean([e[2].elapsed_ms(e[3]) for e in ev]))
    nev, npos = s.counts()
    nev //= args.steps
    npos //= args.steps
    iters = s.iterations()
    return dict(scan=s, t_step=t_total, t_grid=t_grid, t_relax=t_relax, t_df=t_df, wall_step_ms=1e3 * wall / args.steps, n_eval=nev, n_pos=npos, mean_iters=float(iters.mean()))


def run_stiffness(s, reps=3):
    """stiffnessMatrix (which = 3, ddisp 0.05) at every relaxed position of
    the scan, on its grid; HIP events around the launches."""
    from ppafm_amd import device as D

    outs = D.stiffness(s.n, s.lvec[1:], s.grid, s.tip, s.scan_d, s.pos, which=3, ddisp=0.05)
    D.synchronize()
    e0, e1 = D.Event(), D.Event()
    e0.record()
    for _ in range(reps):
        D.stiffness(s.n, s.lvec[1:], s.grid, s.tip, s.scan_d, s.pos, which=3, ddisp=0.05, out=outs)
    e1.record()
    D.synchronize()
    ms = e0.elapsed_ms(e1) / reps
    evals = outs[0].download().reshape(-1, 3)
    return dict(ms=ms, npos=len(evals), n_unstable=int((evals[:, 0] <= 0).sum()), n_nan=int(np.isnan(evals).any(axis=1).sum()),
                positions=s.pos.download().astype(np.float64).reshape(-1, 3), evals=evals)


def run_cpu_stiffness(cfg, lvec, grid_host, st, ncols=None):
    """The reference's stiffnessMatrix (serial loop, ProbeParticle.cpp:703-750)
    at the GPU's relaxed positions and on its FP64 grid: the whole scan
    (about a second) or its first ncols pixel columns."""
    O, threads = _cpu_env()
    from ppafm_amd import common as C
    from ppafm_amd import molgen as M

    nz = cfg["nz"]
    ntot = cfg["scan_n"][0] * cfg["scan_n"][1]
    ncols = ntot if ncols is None else min(ncols, ntot)
    tips = np.ascontiguousarray((M.scan_tips(cfg) - lvec[0]).reshape(-1, 3)[: ncols * nz])
    rpps = np.ascontiguousarray(st["positions"][: ncols * nz])
    FF = np.ascontiguousarray(grid_host[..., :3], dtype=np.float64)
    kS = np.array([M.K_LAT, M.K_LAT, 0.0]) / -C.eVA_Nm
    kind = "reference" if O.have_reference() else "port"
    with quiet_stdout():
        if kind == "reference":
            ref = O.Reference()
            ref.set_grid(np.array(cfg["n"], dtype=np.int32), lvec[1:], FF)
            ref.set_tip(M.TIP_R0[2], M.K_RAD / -C.eVA_Nm, np.zeros(3), kS)
            t0 = time.perf_counter()
            ev, _ = ref.stiffness(tips, rpps, which=3, ddisp=0.05)
            t = time.perf_counter() - t0
        else:
            tip = O.tip_params(M.TIP_R0[2], M.K_RAD / -C.eVA_Nm, (0.0, 0.0, 0.0), kS)
            t0 = time.perf_counter()
            ev, _, _ = O.stiffness(cfg["n"], lvec[1:], FF, tip, tips, rpps, which=3, ddisp=0.05)
            t = time.perf_counter() - t0
    m = len(tips)
    dev = float(np.abs(ev - st["evals"][:m]).max() / np.abs(ev).max())
    cores = 1 if kind == "reference" else threads
    return dict(value=m / t, unit="positions/s", cores=cores, kind=kind,
                sample=f"{ncols} of {ntot} pixel columns ({m} positions, {t:.2f} s); "
                       f"the reference's loop is serial; CPU {cpu_model()}",
                max_rel_eigenvalue_diff_vs_gpu=dev)


def run_batch(args):
    """Batched workload: seeded molecules of U[8,120] atoms, each with its own
    128x128x160 grid and 128x128x30 scan, run as one captured graph (grid
    build, persistent FIRE relaxation, df) with the autotuned launch shape."""
    from ppafm_amd import batch as B
    from ppafm_amd import device as D
    from ppafm_amd import molgen as M

    cfg = M.CONFIGS[2]
    nb = args.batch_items
    mols = [M.config_molecule(2, k) for k in range(nb)]
    lvec = M.config_lvec(cfg)
    b = B.BatchScan.plain([lvec] * nb, [cfg["n"]] * nb, [cfg] * nb, tipR0=M.TIP_R0, k_lat=M.K_LAT, k_rad=M.K_RAD,
                          iZPP=M.PROBE_Z, probe_q=M.PROBE_Q, amplitude=M.DF_AMPLITUDE, molecules=mols,
                          grid_f64=True, tuning="auto")
    st = b.stream
    for _ in range(max(1, args.warmup)):
        b.run()
    b.synchronize()
    steps = max(1, min(args.steps, 3))
    e0, e1 = D.Event(), D.Event()
    e0.record(st)
    for _ in range(steps):
        b.run()
    e1.record(st)
    b.synchronize()
    t_step = e0.elapsed_ms(e1) / steps
    nev, npos = b.counts()
    lane_use = b.lane_use()
    # stages one by one (same kernels, launched without the graph)
    stage = {}
    for name, mask in (("grid", 1), ("relax", 2), ("df", 4)):
        e0.record(st)
        b.run_stages(mask)
        e1.record(st)
        b.synchronize()
        stage[name] = e0.elapsed_ms(e1)
    natoms = np.array([len(m[1]) for m in mols])
    npts = int(np.prod(cfg["n"]))
    res = dict(t_step=t_step, stage=stage, n_eval=nev, n_pos=npos, lane_use=lane_use, natoms_total=int(natoms.sum()), npts=npts * nb,
               tuning=b.tuning, nitems=nb)
    if args.learned_samples > 0:
        res["learned"] = run_learned(args, b)
    if args.fp32:
        res["fp32"] = run_fp32_batch(args, b, mols, cfg, lvec)
    if not args.no_cpu:
        res["cpu"] = run_cpu_batch(args, b, mols, cfg)
    b.close()
    return res


def _err_stats(d):
    d = np.asarray(d, dtype=np.float64).ravel()
    return {"max": float(d.max()), "p99": float(np.percentile(d, 99)), "p50": float(np.median(d))}


ITER_BINS = [0, 1, 2, 5, 10, 20, 50, 100, 200, 500, 1000, 1001]


def iteration_histogram(it):
    """Counts of FIRE iterations per position in the bins [0,1), [1,2), [2,5),
    ..., [500,1000), [1000] (the cap)."""
    h, _ = np.histogram(np.asarray(it).ravel(), bins=ITER_BINS)
    return {"bins": ITER_BINS, "counts": [int(v) for v in h]}


def run_fp32_batch(args, b64, mols, cfg, lvec):
    """The batch with FP32 grids (every pair FP32, FP64 partial sums): the
    grid stage timed, and its grids, relaxed positions and df against the
    FP64 batch (itself equal to the FP64 oracle within 1e-12 on the grid and
    1e-6 A on positions): grid |dF|/max(|F|, 1e-3), positions |dr| (max,
    p99, fraction beyond 1e-3 A), df |d df|."""
    from ppafm_amd import batch as B
    from ppafm_amd import device as D
    from ppafm_amd import molgen as M

    nb = b64.nitems
    b64.set_plain()
    b64.set_tuning(b64.tuning)
    b64.run()
    b64.synchronize()
    b = B.BatchScan.plain([lvec] * nb, [cfg["n"]] * nb, [cfg] * nb, tipR0=M.TIP_R0, k_lat=M.K_LAT, k_rad=M.K_RAD,
                          iZPP=M.PROBE_Z, probe_q=M.PROBE_Q, amplitude=M.DF_AMPLITUDE, molecules=mols, grid_f64=False,
                          fp64_radius=0.0, tuning=b64.tuning)
    b.run()
    b.synchronize()
    e0, e1 = D.Event(), D.Event()
    e0.record(b.stream)
    b.run_stages(1)
    e1.record(b.stream)
    b.synchronize()
    t_grid = e0.elapsed_ms(e1)
    e0.record(b.stream)
    b.run()
    e1.record(b.stream)
    b.synchronize()
    t_step = e0.elapsed_ms(e1)
    pairs = int(np.prod(cfg["n"])) * sum(len(m[1]) for m in mols)
    gerr, dpos, ddf = [], [], []
    for k in range(nb):
        if k % 64 == 0:  # grids: every 64th item (a grid pair is 84 + 42 MB on the host)
            g64 = b64.grid(k)
            g32 = b.grid(k).astype(np.float64)
            d = np.linalg.norm(g32[..., :3] - g64[..., :3], axis=-1) / np.maximum(np.linalg.norm(g64[..., :3], axis=-1), 1e-3)
            gerr.append(d.ravel())
        dpos.append(np.linalg.norm(b.positions_raw(k).astype(np.float64) - b64.positions_raw(k), axis=-1).ravel())
        ddf.append(np.abs(b.df_host(k) - b64.df_host(k)).ravel())
    b.close()
    dpos = np.concatenate(dpos)
    return {"grid_ms": t_grid, "ms_per_step": t_step, "grid_pts_per_s": int(np.prod(cfg["n"])) * nb / (t_grid * 1e-3),
            "grid_build_tflops": 32 * pairs / (t_grid * 1e-3) / 1e12,
            "grid_rel_err": _err_stats(np.concatenate(gerr)), "grid_err_sample": f"items 0, 64, ..., {nb - 1 - (nb - 1) % 64}",
            "positions_err_A": dict(_err_stats(dpos), frac_above_1e_3=float(np.mean(dpos > 1e-3))),
            "df_err_Hz": _err_stats(np.concatenate(ddf))}


def run_fp32_c1(args, cfg, xyz, Zs, qs, lvec, g64):
    """c1 with the FP32 grid kernel (fp64_radius 0): stage times and the
    FP32 grid, positions and df against the FP64 run (equal to the FP64
    oracle within 5e-13 on the grid and 2e-6 A on positions)."""
    g = run_gpu(args, cfg, xyz, Zs, qs, lvec, grid_f64=False, fp64_radius=0.0)
    s32, s64 = g["scan"], g64["scan"]
    G64 = s64.grid_host()
    G32 = s32.grid_host().astype(np.float64)
    fref = np.maximum(np.linalg.norm(G64[..., :3], axis=-1), 1e-3)
    gerr = np.linalg.norm(G32[..., :3] - G64[..., :3], axis=-1) / fref
    gcomp = {c: _err_stats(np.abs(G32[..., k] - G64[..., k]) / fref) for k, c in enumerate("xyz")}
    dpos = np.linalg.norm(s32.positions().astype(np.float64) - s64.positions(), axis=-1)
    ddf = np.abs(s32.df_host() - s64.df_host())
    it32, it64 = s32.iterations(), s64.iterations()
    pairs = int(np.prod(cfg["n"])) * len(xyz)
    return {"ms_per_step": g["t_step"], "stage_ms": {"grid": g["t_grid"], "relax": g["t_relax"], "df": g["t_df"]},
            "positions_per_s": g["n_pos"] / (g["t_step"] * 1e-3), "grid_pts_per_s": int(np.prod(cfg["n"])) / (g["t_grid"] * 1e-3),
            "grid_build_tflops": 32 * pairs / (g["t_grid"] * 1e-3) / 1e12, "grid_rel_err": _err_stats(gerr),
            "grid_rel_err_per_component": gcomp,
            "positions_err_A": dict(_err_stats(dpos), frac_above_1e_3=float(np.mean(dpos > 1e-3))),
            "df_err_Hz": _err_stats(ddf),
            "iterations": {"fp32_grid": iteration_histogram(it32), "fp64_grid": iteration_histogram(it64),
                           "positions_with_other_count": int((it32 != it64).sum())}}


def _cpu_env():
    """The oracle module and the OpenMP threads it runs with (the
    environment is set at the start of main(), before any OpenMP runtime
    starts: once a bound parallel region has run, the main thread's
    affinity no longer shows the cores)."""
    if os.path.join(ROOT, "oracle") not in sys.path:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    return O, O.num_threads()


def run_cpu_batch(args, b, mols, cfg, nsub=16):
    """Reference C++ on the first ``nsub`` molecules of the batch (grid build,
    then the strokes on the GPU-built FP64 grid), one run each; the batch is
    a seeded sample of one distribution, so the full batch scales by
    positions."""
    O, cores = _cpu_env()
    from ppafm_amd import common as C
    from ppafm_amd import molgen as M

    kind = "reference" if O.have_reference() else "port"
    lvec = M.config_lvec(cfg)
    n = np.array(cfg["n"], dtype=np.int32)
    tips = np.ascontiguousarray(M.scan_tips(cfg) - lvec[0])
    tip = O.tip_params(M.TIP_R0[2], M.K_RAD / -C.eVA_Nm, (0.0, 0.0, 0.0), np.array([M.K_LAT, M.K_LAT, 0.0]) / -C.eVA_Nm)
    t_grid = t_relax = 0.0
    nsub = min(nsub, b.nitems)
    with quiet_stdout():
        ref = O.Reference() if kind == "reference" else None
        for k in range(nsub):
            xyz, Zs, qs = mols[k]
            cLJs = C.getAtomsLJ(M.PROBE_Z, Zs, C.loadSpecies())
            Rs0 = np.ascontiguousarray(xyz - lvec[0])
            kq = np.ascontiguousarray(qs * C.CoulombConst)
            F = np.zeros(tuple(n) + (3,))
            FF = np.ascontiguousarray(b.grid(k)[..., :3], dtype=np.float64)
            if ref is not None:
                ref.set_grid(n, lvec[1:], F)
                t0 = time.perf_counter()
                ref.lj(Rs0, cLJs)
                ref.coulomb(Rs0, kq, 0)
                t_grid += time.perf_counter() - t0
                ref.set_grid(n, lvec[1:], FF)
                ref.set_tip(M.TIP_R0[2], M.K_RAD / -C.eVA_Nm, np.zeros(3), np.array([M.K_LAT, M.K_LAT, 0.0]) / -C.eVA_Nm)
                ref.set_relax(1000, 1e-8, 0.1, 0.1)
                t0 = time.perf_counter()
                ref.relax_strokes(tips)
                t_relax += time.perf_counter() - t0
            else:
                t0 = time.perf_counter()
                O.lj_ff(n, lvec[1:], Rs0, cLJs, F=F)
                O.coulomb_ff(n, lvec[1:], Rs0, kq, F=F)
                t_grid += time.perf_counter() - t0
                t0 = time.perf_counter()
                O.relax_strokes(n, lvec[1:], FF, tip, tips, want_iters=False)
                t_relax += time.perf_counter() - t0
    npos = nsub * int(np.prod(cfg["scan_n"])) * cfg["nz"]
    scale = b.nitems / nsub
    return dict(value=npos / (t_grid + t_relax), unit="relaxed tip positions/s", cores=cores, kind=kind,
                sample=f"first {nsub} of {b.nitems} molecules, one run each: grid {t_grid:.2f} s, relaxation {t_relax:.2f} s; "
                       f"CPU {cpu_model()}, {cores} OpenMP threads",
                full_batch_s_extrapolated=(t_grid + t_relax) * scale)


def run_learned(args, b):
    """Learned warm start on the same batch: the network trained on step 0
    of 95% of the molecules (the last 5% held out for the accuracy
    figures), then plain FIRE, warm starts without the network (Newton from
    the plain start at step 0) and with it, each as the whole captured
    graph (grid build, relaxation, df) timed over 3 runs."""
    from ppafm_amd import batch as B
    from ppafm_amd import device as D
    from ppafm_amd import learn as L
    from ppafm_amd.warmstart import NPARAM, WarmStartNet

    nb = b.nitems
    held = list(range(nb - max(1, nb // 20), nb))
    t0 = time.time()
    b.set_plain()
    b.run()
    b.synchronize()
    P, T, truth = L.training_set(b, args.learned_samples, seed=1, items=range(held[0]))
    t_data = time.time() - t0
    print(f"[bench] learned: {P.shape[0]} training pairs in {t_data:.1f} s", file=sys.stderr, flush=True)
    net = WarmStartNet(max_batch=8192, seed=3)
    t0 = time.time()
    losses = L.train(net, P, T, epochs=args.learned_epochs, batch=8192, lr=3e-3)
    t_train = time.time() - t0
    nsteps = len(losses)
    print(f"[bench] learned: {nsteps} training steps in {t_train:.1f} s, loss {losses[:5].mean():.3f} -> {losses[-5:].mean():.4f}",
          file=sys.stderr, flush=True)
    r = L.compare(b, net, truth=truth, items=held, reps=3, tune=True)
    # stage times of each setting with its tuned launch shape
    st = b.stream
    e0, e1 = D.Event(), D.Event()

    def stages(names):
        out_ms = {}
        for name, mask in names:
            b.run_stages(mask)
            e0.record(st)
            b.run_stages(mask)
            e1.record(st)
            b.synchronize()
            out_ms[name] = e0.elapsed_ms(e1)
        return out_ms

    def steps_ms():
        # W untimed warm-up runs, then exactly K timed runs of the graph
        for _ in range(args.warmup):
            b.run()
        b.synchronize()
        e0.record(st)
        for _ in range(args.steps):
            b.run()
        e1.record(st)
        b.synchronize()
        return e0.elapsed_ms(e1) / args.steps

    b.set_warmstart(net)
    b.set_tuning(r["learned"]["tuning"])
    stage = stages((("grid", 1), ("estimates", 8), ("relax", 16), ("df", 4)))
    r["learned"]["ms"] = steps_ms()
    b.set_warmstart(None)
    b.set_tuning(r["newton"]["tuning"])
    stage_newton = stages((("relax", 16),))
    r["newton"]["ms"] = steps_ms()
    b.set_plain()
    b.set_tuning(r["plain"]["tuning"])
    stage_plain = stages((("relax", 16),))
    r["plain"]["ms"] = steps_ms()
    npos = r["plain"]["positions"]
    ncols = sum(int(np.prod(b._scan_shape(k)[:2])) for k in range(nb))
    flop_col = 2 * (9 * 81 * 32 + 288 * 32 + 32 * 3)  # one network pass per column
    out = {"training": {"pairs": int(P.shape[0]), "held_out_molecules": len(held), "data_s": t_data, "train_s": t_train,
                        "steps": nsteps, "batch": 8192, "epochs": args.learned_epochs,
                        "samples_per_s": nsteps * 8192 / t_train,
                        # forward + backward ~ 3 forward passes of flop_col per sample (bf16 MFMA layers 1-2)
                        "tflops": 3 * flop_col * nsteps * 8192 / t_train / 1e12, "loss_first": float(losses[:5].mean()),
                        "loss_last": float(losses[-5:].mean()), "parameters": NPARAM}}
    for k in ("plain", "newton", "learned"):
        q = r[k]
        out[k] = {"ms_per_step": q["ms"], "positions_per_s": npos / (q["ms"] * 1e-3), "tuning": q["tuning"],
                  "force_evals_per_position": q["evals"] / npos, "fallback_steps": q["fallbacks"],
                  "lane_use": q["lane_use"], "held_out_vs_min_A": q["vs_min"]}
        if "vs_plain" in q:
            out[k]["held_out_vs_plain_A"] = q["vs_plain"]
        if "estimate_vs_min" in q:
            out[k]["held_out_step0_estimate_vs_min_A"] = q["estimate_vs_min"]
    out["learned"]["stage_ms"] = stage
    # step 0 is plain FIRE where the total force at the estimate exceeds this (counted in fallback_steps)
    out["learned"]["residual_max_eV_per_A"] = B.RESIDUAL_MAX
    out["newton"]["stage_ms"] = stage_newton
    out["plain"]["stage_ms"] = stage_plain
    out["learned"]["estimates_mfma"] = {"flop_per_column": flop_col, "columns": ncols,
                                        "achieved_tflops": flop_col * ncols / (stage["estimates"] * 1e-3) / 1e12,
                                        "peak_tflops": 2500.0,
                                        "note": "bf16 MFMA; the stage is bound by the per-node patch arithmetic (VALU), not by MFMA"}
    out["speedup_learned_vs_plain"] = r["plain"]["ms"] / r["learned"]["ms"]
    out["speedup_newton_only_vs_plain"] = r["plain"]["ms"] / r["newton"]["ms"]
    out["speedup_learned_vs_newton_only"] = r["newton"]["ms"] / r["learned"]["ms"]
    # the relaxation alone (the grid build and df are the same work in every setting)
    out["speedup_learned_vs_plain_relaxation"] = stage_plain["relax"] / (stage["estimates"] + stage["relax"])
    out["speedup_newton_only_vs_plain_relaxation"] = stage_plain["relax"] / stage_newton["relax"]
    out["speedup_learned_vs_newton_only_relaxation"] = stage_newton["relax"] / (stage["estimates"] + stage["relax"])
    # agreement with plain FIRE (the reference's results) over the held-out items
    lp = r["plain"]["positions_array"].astype(np.float64)
    lw = r["learned"]["positions_array"].astype(np.float64)
    d = np.linalg.norm(lw - lp, axis=1)
    out["held_out_learned_vs_plain"] = {"p50_A": float(np.median(d)), "p99_A": float(np.percentile(d, 99)), "max_A": float(d.max()),
                                        "frac_above_1e_3_A": float(np.mean(d > 1e-3)), "count_above_0_1_A": int((d > 0.1).sum()),
                                        "positions": int(len(d))}
    return out


def run_generator(args, batch=32):
    """AFM images as the ML generators make them: AFMulator with the
    reference's defaults (128x128x30 scan over a 16x16 A window, CO-like tip
    charges) for molecules of the batched configuration (U[8,120] atoms),
    evaluated one call per molecule (the reference generators' loop,
    ml/Generator.py:96-113) and as batches of requests
    (AFMulator.request/run_requests, which ppafm_amd.ml.Generator uses)."""
    from ppafm_amd import molgen as M
    from ppafm_amd.ocl.AFMulator import AFMulator

    n = args.generator_molecules
    mols = []
    for k in range(n + batch):
        xyz, Zs, qs = M.config_molecule(2, k)
        mols.append((xyz + np.array([10.0, 10.0, 0.0]), Zs, qs))
    warm, mols = mols[:batch], mols[batch:]
    afm = AFMulator()
    afm(*warm[0])
    afm.run_requests([afm.request(*m) for m in warm])  # builds and captures the plan
    t0 = time.perf_counter()
    per_call = [afm(*m) for m in mols]
    t_call = time.perf_counter() - t0
    t0 = time.perf_counter()
    batched = []
    for i in range(0, n, batch):
        batched += afm.run_requests([afm.request(*m) for m in mols[i:i + batch]])
    t_batch = time.perf_counter() - t0
    same = all(np.array_equal(a, b) for a, b in zip(per_call, batched))
    return {
        "workload": f"{n} molecules of the batched configuration (U[8,120] atoms), AFMulator defaults: grid {tuple(int(v) for v in afm.nDim[:3])}, "
                    f"scan {tuple(afm.scan_dim)}, df_steps {afm.df_steps}; batches of {batch}",
        "images_per_s_per_call": n / t_call,
        "images_per_s_batched": n / t_batch,
        "speedup": t_call / t_batch,
        "identical_images": bool(same),
    }


def _plane_wave_potential(lv, n, seed=3):
    """A smooth periodic synthetic Hartree potential (V): a few plane waves
    of the sample cell."""
    rng = np.random.default_rng(seed)
    f = [np.arange(m) / m for m in n]
    V = np.zeros(n)
    for _ in range(6):
        k = rng.integers(-3, 4, 3)
        a, ph = rng.uniform(0.2, 1.0), rng.uniform(0, 2 * np.pi)
        V += a * np.cos(2 * np.pi * (k[0] * f[0][:, None, None] + k[1] * f[1][None, :, None] + k[2] * f[2][None, None, :]) + ph)
    return V


def run_hartree(args, reps=3):
    """AFMulator with the LJ+Hartree method at the reference's defaults
    (force-field grid 256x256x192, 128x128x30 scan, dz2 tip density): one
    molecule of the batched configuration, a synthetic Hartree potential on
    a 24x24x20 A cell (200x200x166 points), rotated by 0.3 rad.  Reports
    images/s of whole AFMulator calls and the stage times of the force
    field: the Lennard-Jones grid, the potential interpolated onto the
    force-field grid, the FFT cross-correlation with the tip density and
    the gradient added to the grid."""
    from ppafm_amd import device as D
    from ppafm_amd import molgen as M
    from ppafm_amd.ocl import field as FFcl
    from ppafm_amd.ocl.AFMulator import AFMulator

    xyz, Zs, qs = M.config_molecule(2, 7)
    xyz = xyz + np.array([10.0, 10.0, 0.0])
    lv = np.array([[-2.0, -2.0, -6.0], [24.0, 0, 0], [0, 24.0, 0], [0, 0, 20.0]])
    pot = FFcl.HartreePotential(_plane_wave_potential(lv, (200, 200, 166)), lv)
    a = 0.3
    rot = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1.0]])
    afm = AFMulator(rho={"dz2": -0.1})
    afm(xyz, Zs, pot, rot=rot)  # tip density, FFT plans, buffers
    D.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        X = afm(xyz, Zs, pot, rot=rot)
    t_eval = (time.perf_counter() - t0) / reps
    # the stages of prepareFF, each bracketed by synchronisation
    n = tuple(int(v) for v in afm.nDim[:3])
    lvec_ff = afm._lvec_ff()
    xyzq, cLJs, _, _, _, rc = afm._prepare_atoms(xyz, Zs, None, rot=rot, full=True)
    stage = {}

    def timed(name, fn):
        # HIP events on the stream the stage's kernels run on (the default one)
        D.synchronize()
        e0, e1 = D.Event(), D.Event()
        e0.record()
        for _ in range(reps):
            r = fn()
        e1.record()
        D.synchronize()
        stage[name] = e0.elapsed_ms(e1) / reps
        return r

    ws = afm._ws
    timed("lj_grid", lambda: D.ff_build(n, lvec_ff[1:], xyzq, cLJs, afm._ff_model(False), out=afm._grid))
    pot_i = timed("interp", lambda: pot.interp_at(lvec_ff, n, array_out=ws.buf("a"), rot=np.linalg.inv(rot), rot_center=rc))
    fft = afm.forcefield.fft_corr
    E = timed("correlate", lambda: fft.correlate(pot_i, E=ws.grid("b", lvec_ff)))
    timed("grad_add", lambda: E.grad(scale=[-1.0, -1.0, -1.0, 1.0], array_out=afm._grid, order="C", accumulate=True))
    npts = int(np.prod(n))
    out = {
        "workload": f"AFMulator LJ+Hartree, dz2 tip density: grid {n}, scan {tuple(afm.scan_dim)}, df_steps {afm.df_steps}; "
                    f"{len(Zs)}-atom molecule, synthetic potential 200x200x166, rotated 0.3 rad",
        "images_per_s": 1.0 / t_eval,
        "ms_per_image": t_eval * 1e3,
        "stage_ms": stage,
        "grid_points": npts,
        "interp_gbs": 16 * npts / (stage["interp"] * 1e-3) / 1e9,
        "grad_add_gbs": 72 * npts / (stage["grad_add"] * 1e-3) / 1e9,
        "bytes_model": "interp: 8 B value in + 8 B out per point; grad_add: 8 B in + 32 B read + 32 B written per point (double4 grid)",
        "image_finite": bool(np.isfinite(X).all()),
    }
    if not args.no_cpu:
        _cpu_env()
        import hartree_oracle as H

        Eh = E.array
        rho = afm.forcefield.rho.array
        t = time.perf_counter()
        Ec = H.correlate_fft(rho, pot_i.array, afm.forcefield.rho.cell_vol)
        G = H.grad(Ec, np.diag(H.steps(lvec_ff, n)), (-1, -1, -1, 1))
        t_cpu = time.perf_counter() - t
        out["cpu_baseline"] = {"value": 1.0 / t_cpu, "unit": "correlate+grad per s", "cores": 1, "kind": "port",
                               "sample": f"numpy FFT cross-correlation and gradient of the same {n} grids (the oracle)"}
        out["correlate_grad_speedup_vs_cpu"] = t_cpu / ((stage["correlate"] + stage["grad_add"]) * 1e-3)
        out["correlate_max_rel_diff_vs_cpu"] = float(np.abs(Eh - Ec).max() / np.abs(Ec).max())
        del G
    return out


def run_c4(args):
    """Large tilted case: per instance an Au(111) slab with a seeded
    60-atom molecule and 3x3 xy images (17,928 atoms), the OpenCL LJ +
    point-charge grid on 400x400x300 points, a 400x400x60 scan with the tip
    tilted 10 degrees (RelaxedScanner.run_relaxStrokesTilted semantics), df.
    All instances in one batch plan (one captured graph); one timed run
    after one warm-up run."""
    from ppafm_amd import batch as B
    from ppafm_amd import device as D
    from ppafm_amd import molgen as M

    cfg = M.CONFIGS[4]
    ni = args.c4_instances
    systems = [M.c4_system(k) for k in range(ni)]
    lvec = M.config_lvec(cfg)
    t = np.deg2rad(cfg["tilt_deg"])
    rot = np.array([[np.cos(t), 0, np.sin(t)], [0, 1, 0], [-np.sin(t), 0, np.cos(t)]])
    b = B.BatchScan.tilted([lvec] * ni, [cfg["n"]] * ni, rot=rot, pos0s=[(0.0, 0.0, cfg["z_top"])] * ni,
                           start=(cfg["scan_x"][0], cfg["scan_y"][0]), end=(cfg["scan_x"][1], cfg["scan_y"][1]),
                           scan_n=cfg["scan_n"], nz=cfg["nz"], zstep=cfg["dz"],
                           molecules=[(x, z, q) for x, z, q, _ in systems], tuning="auto", tune_sample=2,
                           grid_f64=args.c4_grid == "f64")
    st = b.stream
    b.run()
    b.synchronize()
    e0, e1 = D.Event(), D.Event()
    e0.record(st)
    b.run()
    e1.record(st)
    b.synchronize()
    t_step = e0.elapsed_ms(e1)
    nev, npos = b.counts()
    stage = {}
    for name, mask in (("grid", 1), ("relax", 2), ("df", 4)):
        e0.record(st)
        b.run_stages(mask)
        e1.record(st)
        b.synchronize()
        stage[name] = e0.elapsed_ms(e1)
    natoms = len(systems[0][0])
    npts = int(np.prod(cfg["n"]))
    pairs = npts * natoms * ni
    tuning = b.tuning
    learned = run_c4_learned(args, b, stage, nev, npos) if args.learned_samples > 0 else None
    cpu = None if args.no_cpu else run_cpu_c4(b, systems[0], cfg, lvec)
    fp64_check = run_c4_fp64_check(b, systems[0], cfg, lvec, rot) if args.c4_grid == "f32" else None
    b.close()
    return {
        "workload": f"{ni} instances: Au(111) slab (3 layers, 1932 atoms) + 60-atom molecule, 3x3 images ({natoms} atoms), "
                    "OpenCL LJ+charge grid 400x400x300 at 0.1 A, 400x400x60 scan tilted 10 deg, df",
        "instances": ni, "atoms_per_instance": natoms, "grid_points": npts * ni, "positions": npos,
        "grid_storage": args.c4_grid,
        "ms_per_step": t_step, "positions_per_s": npos / (t_step * 1e-3), "stage_ms": stage,
        "grid_pairs_per_s": pairs / (stage["grid"] * 1e-3),
        "grid_build_tflops": 32 * pairs / (stage["grid"] * 1e-3) / 1e12,
        "force_evals_per_position": nev / npos, "tuning": tuning,
        "learned": learned,
        "fp32_grid_vs_fp64_instance0": fp64_check,
        "cpu_baseline": cpu,
        "speedup_vs_cpu": (npos / (t_step * 1e-3)) / cpu["value"] if cpu else None,
    }


def run_c4_learned(args, b, stage_plain, nev_plain, npos):
    """Learned warm start on the tilted instances, beside plain FIRE: the
    network trained on step 0 of every instance but the last two (held out),
    then warm starts (network at step 0, Newton from the continuation) on all
    of them; relaxation-stage times, force evaluations and the held-out
    instances' agreement with plain FIRE."""
    from ppafm_amd import device as D
    from ppafm_amd import learn as L
    from ppafm_amd.batch import tuned
    from ppafm_amd.warmstart import WarmStartNet

    ni = b.nitems
    held = list(range(max(1, ni - 2), ni))
    b.set_plain()
    b.run()
    b.synchronize()
    plain = L.positions_of(b, held).astype(np.float64)
    n_train = min(args.learned_samples, 8_000_000)
    P, T, truth = L.training_set(b, n_train, seed=5, items=range(held[0]))
    net = WarmStartNet(max_batch=8192, seed=3)
    losses = L.train(net, P, T, epochs=2.0, batch=8192, lr=3e-3)
    del P, T
    st = b.stream
    e0, e1 = D.Event(), D.Event()
    out = {"training_pairs": int(n_train), "loss_first": float(losses[:5].mean()), "loss_last": float(losses[-5:].mean())}
    for key, model in (("newton", None), ("learned", net)):
        b.set_warmstart(model)
        b.set_tuning(tuned(b, nsample=2))
        b.run()
        b.synchronize()
        stages = {}
        for name, mask in (("estimates", 8), ("relax", 16)):
            if name == "estimates" and model is None:
                continue
            b.run_stages(mask)
            e0.record(st)
            b.run_stages(mask)
            e1.record(st)
            b.synchronize()
            stages[name] = e0.elapsed_ms(e1)
        b.run()
        b.synchronize()
        ev, _, nfb = b.counters()
        w = L.positions_of(b, held).astype(np.float64)
        d = np.linalg.norm(w - plain, axis=1)
        relax_ms = stages["relax"] + stages.get("estimates", 0.0)
        out[key] = {"stage_ms": stages, "relax_positions_per_s": npos / (relax_ms * 1e-3),
                    "speedup_vs_plain_relaxation": stage_plain["relax"] / relax_ms,
                    "force_evals_per_position": ev / npos, "fallback_steps": nfb, "tuning": b.tuning,
                    "held_out_vs_plain_A": {"p50": float(np.median(d)), "p99": float(np.percentile(d, 99)),
                                            "max": float(d.max()), "frac_above_1e_3": float(np.mean(d > 1e-3)),
                                            "count_above_0_1": int((d > 0.1).sum())}}
    b.set_plain()
    out["plain"] = {"relax_ms": stage_plain["relax"], "force_evals_per_position": nev_plain / npos}
    return out


def run_c4_fp64_check(b, system, cfg, lvec, rot):
    """Instance 0 of the tilted batch again with an FP64 grid (every pair in
    FP64): the FP32 grid's relaxed positions and df against it."""
    from ppafm_amd import batch as B

    xyz, Zs, qs, _ = system
    b64 = B.BatchScan.tilted([lvec], [cfg["n"]], rot=rot, pos0s=[(0.0, 0.0, cfg["z_top"])],
                             start=(cfg["scan_x"][0], cfg["scan_y"][0]), end=(cfg["scan_x"][1], cfg["scan_y"][1]),
                             scan_n=cfg["scan_n"], nz=cfg["nz"], zstep=cfg["dz"], molecules=[(xyz, Zs, qs)], grid_f64=True,
                             tuning=b.tuning)
    b64.run()
    b64.synchronize()
    b.set_plain()
    b.run()
    b.synchronize()
    dp = np.linalg.norm(b.positions_raw(0).astype(np.float64) - b64.positions_raw(0), axis=-1)
    ddf = np.abs(b.df_raw(0) - b64.df_raw(0))
    b64.close()
    return {"positions_err_A": dict(_err_stats(dp), frac_above_1e_3=float(np.mean(dp > 1e-3))), "df_err_Hz": _err_stats(ddf),
            "note": "instance 0 relaxed on its FP32 grid against the same scan on an FP64 grid"}


def run_cpu_c4(b, system, cfg, lvec, planes=(60, 120, 180, 240), stride=10):
    """FP64 port (the reference C++ has neither the OpenCL grid nor tilted
    strokes) on a sample of c4 instance 0: the grid on 4 z-planes of
    400x400 points and the tilted strokes of every stride-th pixel column
    along x and y (1 %) on the GPU-built grid, scaled linearly to the whole
    instance and to all instances."""
    O, cores = _cpu_env()
    from ppafm_amd import common as C

    xyz, Zs, qs, _ = system
    cLJs = C.REA2LJ(C.getAtomsREA(8, Zs, C.loadSpecies(), alphaFac=-1.0))
    atoms = np.concatenate([xyz, qs[:, None]], 1)
    n = np.array(cfg["n"])
    h = np.array(cfg["cell"]) / n
    t0 = time.perf_counter()
    for iz in planes:
        lv = np.array([np.asarray(cfg["origin"]) + np.array([0, 0, iz * h[2]]), [cfg["cell"][0], 0, 0],
                       [0, cfg["cell"][1], 0], [0, 0, h[2]]])
        O.ljc_qzs_ff((n[0], n[1], 1), lv, atoms, cLJs, np.array([-0.1, 0, 0, 0.0]), np.zeros(4))
    t_grid = (time.perf_counter() - t0) * n[2] / len(planes)
    it = b.items[0]
    s = it["scan"]
    ii, jj = np.meshgrid(np.arange(0, s.nx, stride), np.arange(0, s.ny, stride), indexing="ij")
    poss = np.zeros(ii.shape + (4,))
    for c in range(3):
        poss[..., c] = s.p0[c] + ii * s.da[c] + jj * s.db[c]
    cell = np.asarray(it["cell"], dtype=np.float64)
    dinv = np.linalg.inv(cell / n[:, None]).T
    tip = b.tip
    tipRot = np.zeros((3, 4))
    tipRot[:, :3] = np.array(tip.rot).reshape(3, 3)
    tipRot[2, 3] = -cfg["dz"]
    FE = b.grid(0)
    t0 = time.perf_counter()
    O.relax_tilted(FE, dinv, poss.reshape(-1, 4), s.nz, tipRot, np.array([tip.kspring[0], tip.kspring[1], tip.kspring[2], tip.krad]),
                   np.array([tip.d0[0], tip.d0[1], tip.d0[2], tip.lrad]), np.array([b.rp.dt, b.rp.damping]), max_iters=128,
                   conv_f2=float(np.float32(1e-8)), f2safe=float(np.float32(1e-8)))
    t_rel = (time.perf_counter() - t0) * (s.nx * s.ny) / ii.size
    npos = s.nx * s.ny * s.nz
    return dict(value=npos / (t_grid + t_rel), unit="relaxed tip positions/s", cores=cores, kind="port",
                sample=f"instance 0: grid on {len(planes)} of {n[2]} z-planes, tilted strokes of {ii.size} of {s.nx * s.ny} columns, "
                       f"scaled linearly: grid {t_grid:.1f} s, relaxation {t_rel:.1f} s per instance; CPU {cpu_model()}, {cores} OpenMP threads",
                instance_s_extrapolated=t_grid + t_rel)


def run_cpu(args, cfg, xyz, Zs, qs, lvec, grid_host):
    """Reference C++ (or FP64 port) on a bounded sample of the workload."""
    O, cores = _cpu_env()

    from ppafm_amd import common as C
    from ppafm_amd import molgen as M

    cLJs = C.getAtomsLJ(M.PROBE_Z, Zs, C.loadSpecies())
    Rs0 = np.ascontiguousarray(xyz - lvec[0])
    kq = np.ascontiguousarray(qs * C.CoulombConst)
    nx, ny, nz = cfg["n"]
    kind = "reference" if O.have_reference() else "port"
    # the whole workload, median of 3 runs per stage
    reps = 3
    n_full = np.array([nx, ny, nz], dtype=np.int32)
    F = np.zeros((nx, ny, nz, 3))
    FF = np.ascontiguousarray(grid_host[..., :3], dtype=np.float64)  # relax on the GPU's FP64 grid
    tips = np.ascontiguousarray(M.scan_tips(cfg) - lvec[0])
    tip = O.tip_params(M.TIP_R0[2], M.K_RAD / -C.eVA_Nm, (0.0, 0.0, 0.0), np.array([M.K_LAT, M.K_LAT, 0.0]) / -C.eVA_Nm)
    t_grid_r, t_rel_r = [], []
    with quiet_stdout():
        if kind == "reference":
            ref = O.Reference()
        for _ in range(reps):
            F[:] = 0
            if kind == "reference":
                ref.set_grid(n_full, lvec[1:], F)
                t0 = time.perf_counter()
                ref.lj(Rs0, cLJs)
                ref.coulomb(Rs0, kq, 0)
                t_grid_r.append(time.perf_counter() - t0)
                ref.set_grid(n_full, lvec[1:], FF)
                ref.set_tip(M.TIP_R0[2], M.K_RAD / -C.eVA_Nm, np.zeros(3), np.array([M.K_LAT, M.K_LAT, 0.0]) / -C.eVA_Nm)
                ref.set_relax(1000, 1e-8, 0.1, 0.1)
                t0 = time.perf_counter()
                ref.relax_strokes(tips)
                t_rel_r.append(time.perf_counter() - t0)
            else:
                t0 = time.perf_counter()
                O.lj_ff(n_full, lvec[1:], Rs0, cLJs, F=F)
                O.coulomb_ff(n_full, lvec[1:], Rs0, kq, F=F)
                t_grid_r.append(time.perf_counter() - t0)
                t0 = time.perf_counter()
                O.relax_strokes(cfg["n"], lvec[1:], FF, tip, tips, want_iters=False)
                t_rel_r.append(time.perf_counter() - t0)
    t_grid = float(np.median(t_grid_r))
    t_relax = float(np.median(t_rel_r))
    sx, sy = cfg["scan_n"]
    n_pos = sx * sy * cfg["nz"]
    value = n_pos / (t_grid + t_relax)
    sample = (f"whole workload (LJ + Coulomb grid of {nx * ny * nz} points, {n_pos} relaxed positions on the GPU-built FP64 grid), "
              f"median of {reps} runs per stage: grid {t_grid:.2f} s, relaxation {t_relax:.2f} s; CPU {cpu_model()}, {cores} OpenMP threads")
    return dict(value=value, unit="relaxed tip positions/s", cores=cores, kind=kind, sample=sample, grid_pts_per_s=nx * ny * nz / t_grid, t_grid_s=t_grid, t_relax_s=t_relax)


def profile_figures():
    """Figures read from the committed rocprofv3 summaries (profiles/*.json):
    HBM bytes per launch (FETCH_SIZE + WRITE_SIZE scaled by the copy-kernel
    calibration) and counter fractions of the kernels the line quotes."""
    out = {}
    for name, pre in (("relax_c1_pmc.json", "relax_c1_"), ("grid_batch_pmc.json", "grid_batch_"),
                      ("warm_relax_pmc.json", "warm_relax_")):
        try:
            d = json.load(open(os.path.join(ROOT, "profiles", name)))
        except (OSError, ValueError):
            continue
        for k, v in d.items():
            if isinstance(v, (int, float)) and not isinstance(v, bool):
                out[pre + k] = float(v)
    return out


def main():
    args = parse()
    # OpenMP threads of the CPU legs: every core this process may use (set
    # before any OpenMP runtime starts)
    cores = cpu_cores()
    os.environ["OMP_NUM_THREADS"] = str(cores)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)
    from ppafm_amd import device as D
    from ppafm_amd import molgen as M

    peaks = {k: D.measure_peak(k) for k in D.PEAK_KINDS}
    prof = profile_figures()

    # ---- the metric's configuration: learned warm start + batching against
    # plain FIRE on the 512-molecule batch (the headline line)
    bt = run_batch(args) if args.batch_items > 0 else None
    Lr = bt.get("learned") if bt else None
    cfg2 = M.CONFIGS[2]
    if Lr:
        npos = bt["n_pos"]
        ms = Lr["learned"]["ms_per_step"]
        st = Lr["learned"]["stage_ms"]
        pairs = bt["npts"] // bt["nitems"] * bt["natoms_total"]
        grid_tf = 32 * pairs / (st["grid"] * 1e-3) / 1e12
        relax_ms = st["estimates"] + st["relax"]
        if st["grid"] >= relax_ms:
            roof = {"kernel": "ff_build_f64_batch (FP64 pairs, double4 grids; the learned step's longest stage)",
                    "bound": "valu", "achieved": grid_tf, "peak": 157.3, "unit": "TFLOP/s", "frac": grid_tf / 157.3,
                    "traffic": prof.get("grid_batch_hbm_bytes_per_launch"),
                    "algorithmic_unit": "32 FLOP per atom-point pair (LJ + Coulomb, force + energy)",
                    "pairs_per_launch": pairs, "kernel_ms": st["grid"],
                    "fp64_vendor_peak_tflops": 78.6, "frac_of_fp64_vendor_peak": grid_tf / 78.6,
                    "fp64_measured_peak_tflops": peaks["fp64_fma"]}
        else:
            ev = Lr["learned"]["force_evals_per_position"] * npos
            ach = (96 * ev + 24 * npos) / (st["relax"] * 1e-3) / 1e9
            kname = "relax_refill_kernel" if Lr["learned"]["tuning"].get("refill") else "relax_batch_kernel"
            roof = {"kernel": f"{kname} (warm-started FIRE, the learned step's longest stage)", "bound": "latency",
                    "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                    "traffic": prof.get("warm_relax_hbm_bytes_per_launch"),
                    "algorithmic_unit": "96 B per force evaluation + 24 B per relaxed position", "kernel_ms": st["relax"]}
        out = {
            "metric": METRIC,
            "value": npos / (ms * 1e-3),
            "unit": "relaxed tip positions/s",
            "n_gpus": 1,
            "steps": args.steps,
            "warmup": args.warmup,
            "ms_per_step": ms,
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "f64",
            "data": "synthetic",
            "config": {
                "workload": f"learned warm-start FIRE on the batch of {bt['nitems']} seeded molecules (U[8,120] atoms), each a "
                            "128x128x160 LJ+Coulomb grid at 0.1 A (FP64) and a 128x128x30 scan, one captured graph "
                            "(grid, step-0 network estimates, warm-started FIRE to |F| < 1e-4 eV/A, 9-tap df); "
                            "against plain FIRE on the same batch",
                "molecules": bt["nitems"], "tip_positions": npos, "grid_points": bt["npts"],
                "training_pairs": Lr["training"]["pairs"], "held_out_molecules": Lr["training"]["held_out_molecules"],
            },
            "speedup_vs_plain_relaxation": Lr["speedup_learned_vs_plain_relaxation"],
            "speedup_vs_plain_step": Lr["speedup_learned_vs_plain"],
            "speedup_vs_newton_only_relaxation": Lr["speedup_learned_vs_newton_only_relaxation"],
            "speedup_vs_newton_only_step": Lr["speedup_learned_vs_newton_only"],
            "relax_positions_per_s": npos / (relax_ms * 1e-3),
            "plain_relax_positions_per_s": npos / (Lr["plain"]["stage_ms"]["relax"] * 1e-3),
            "grid_pts_per_s": bt["npts"] / (st["grid"] * 1e-3),
            "stage_ms": st,
            "roofline": roof,
            "cpu_baseline": None,
        }
        if "cpu" in bt:
            c = bt["cpu"]
            out["cpu_baseline"] = {k: c[k] for k in ("value", "unit", "cores", "kind", "sample")}
            out["cpu_baseline"]["affinity_cores"] = cores
            out["speedup_vs_cpu"] = out["value"] / c["value"]
        out["learned"] = Lr
    else:
        out = {"metric": METRIC, "value": None, "unit": "relaxed tip positions/s", "n_gpus": 1, "steps": args.steps,
               "warmup": args.warmup, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "f64", "data": "synthetic",
               "config": {"workload": "learned comparison skipped (--batch-items 0 or --learned-samples 0)"},
               "roofline": None, "cpu_baseline": None}
    if bt:
        out["c2_plain_batch"] = {
            "workload": f"{bt['nitems']} seeded molecules (U[8,120] atoms), each a 128x128x160 FP64 grid at 0.1 A and a "
                        "128x128x30 scan, plain FIRE, one captured graph",
            "positions": bt["n_pos"], "grid_points": bt["npts"], "ms_per_step": bt["t_step"],
            "positions_per_s": bt["n_pos"] / (bt["t_step"] * 1e-3), "grid_pts_per_s": bt["npts"] / (bt["stage"]["grid"] * 1e-3),
            "relax_positions_per_s": bt["n_pos"] / (bt["stage"]["relax"] * 1e-3), "stage_ms": bt["stage"],
            "force_evals": bt["n_eval"], "relax_lane_use": bt["lane_use"],
            "grid_build_tflops": 32 * (bt["npts"] // bt["nitems"] * bt["natoms_total"]) / (bt["stage"]["grid"] * 1e-3) / 1e12,
            "grid_peak_tflops": 157.3,
            "relax_algorithmic_gbs": (96 * bt["n_eval"] + 24 * bt["n_pos"]) / (bt["stage"]["relax"] * 1e-3) / 1e9,
            "tuning": bt["tuning"]}
        if "cpu" in bt:
            out["c2_plain_batch"]["cpu_baseline"] = dict(bt["cpu"], affinity_cores=cores)
            out["c2_plain_batch"]["speedup_vs_cpu"] = out["c2_plain_batch"]["positions_per_s"] / bt["cpu"]["value"]
        if "fp32" in bt:
            out["c2_fp32_grid"] = bt["fp32"]

    # ---- c1: plain FIRE, one 120-atom molecule, the production-size scan
    cfg = M.CONFIGS[1]
    xyz, Zs, qs = M.config_molecule(1, 0, cfg["natoms"])
    lvec = M.config_lvec(cfg)
    g = run_gpu(args, cfg, xyz, Zs, qs, lvec)
    npts = int(np.prod(cfg["n"]))
    n_pos = g["n_pos"]
    alg_bytes = 96 * g["n_eval"] + 24 * n_pos
    achieved = alg_bytes / (g["t_relax"] * 1e-3) / 1e9
    traffic = prof.get("relax_c1_hbm_bytes_per_launch")
    from ppafm_amd.batch import device_info

    nsimd = 4 * int(device_info()[1])
    c1 = {
        "workload": "plain FIRE, one 120-atom molecule: LJ+Coulomb grid 256x256x320 at 0.1 A (FP64), 256x256 pixels x 60 "
                    "z-steps, 9-tap df",
        "positions_per_s": n_pos / (g["t_step"] * 1e-3), "ms_per_step": g["t_step"],
        "grid_pts_per_s": npts / (g["t_grid"] * 1e-3), "relax_positions_per_s": n_pos / (g["t_relax"] * 1e-3),
        "stage_ms": {"grid": g["t_grid"], "relax": g["t_relax"], "df": g["t_df"]},
        "host_wall_ms_per_step": g["wall_step_ms"], "force_evals_per_step": g["n_eval"], "mean_fire_iterations": g["mean_iters"],
        "iterations_histogram": iteration_histogram(g["scan"].iterations()),
        "roofline": {
            "kernel": "relax_dev_kernel (FIRE, one lane per pixel column)",
            "bound": "latency",
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": alg_bytes,
            "bytes_model": "96 B per force evaluation (8 corners x 3 FP32) + 24 B per relaxed position",
            "waves_per_simd": (cfg["scan_n"][0] // 8) * (cfg["scan_n"][1] // 8) / nsimd,
            "physical_hbm_gbs": traffic / (g["t_relax"] * 1e-3) / 1e9 if traffic else None,
            "physical_hbm_frac": traffic / (g["t_relax"] * 1e-3) / 1e9 / HBM_PEAK_GBS if traffic else None,
            "valu_busy_frac": prof.get("relax_c1_valu_busy_frac"),
            "note": "each pixel column is a serial chain of dependent FP64 force evaluations and the scan has one wave per "
                    "SIMD, so instruction latency, not bandwidth, bounds the kernel; `traffic` and physical_hbm_* are the "
                    "rocprof FETCH_SIZE + WRITE_SIZE bytes of one launch",
        },
        "grid_build": {
            "bound": "valu", "pairs": npts * len(xyz),
            "achieved_tflops": 32 * npts * len(xyz) / (g["t_grid"] * 1e-3) / 1e12, "peak_tflops": 157.3,
            "frac": 32 * npts * len(xyz) / (g["t_grid"] * 1e-3) / 1e12 / 157.3,
            "fp64_vendor_peak_tflops": 78.6, "fp64_measured_peak_tflops": peaks["fp64_fma"],
            "note": "32 algorithmic FLOP per atom-point pair (LJ + Coulomb, force + energy), FP64 pairs on the double4 grid, "
                    "against the FP32 vector peak the grid roofline is fixed to",
        },
    }
    out["c1"] = c1
    out["peaks_measured_tflops"] = peaks
    if args.fp32:
        print("[bench] c1 with the FP32 grid kernel", file=sys.stderr, flush=True)
        out["c1_fp32_grid"] = run_fp32_c1(args, cfg, xyz, Zs, qs, lvec, g)
    if args.generator_molecules > 0:
        print(f"[bench] generator: {args.generator_molecules} molecules", file=sys.stderr, flush=True)
        out["generator"] = run_generator(args)
    if args.c4_instances > 0:
        print(f"[bench] c4: {args.c4_instances} instances", file=sys.stderr, flush=True)
        out["c4"] = run_c4(args)
    if not args.no_hartree:
        print("[bench] hartree", file=sys.stderr, flush=True)
        out["hartree"] = run_hartree(args)
    st_ = run_stiffness(g["scan"])
    stiff_bytes = 6 * 8 * 4 * 8 + 12 + 96
    out["stiffness"] = {
        "workload": "stiffnessMatrix (which 3, ddisp 0.05 A) at every relaxed position of the c1 scan, on its grid",
        "positions": st_["npos"],
        "ms": st_["ms"],
        "positions_per_s": st_["npos"] / (st_["ms"] * 1e-3),
        "algorithmic_gbs": stiff_bytes * st_["npos"] / (st_["ms"] * 1e-3) / 1e9,
        "bytes_model": "6 force evaluations x 8 corners x 32 B + 12 B position in + 96 B eigenvalues and eigenvectors out (FP64)",
        "positions_with_non_positive_smallest_eigenvalue": st_["n_unstable"],
        "positions_with_nan": st_["n_nan"],
    }
    if not args.no_cpu:
        grid_host = g["scan"].grid_host()
        cs = run_cpu_stiffness(cfg, lvec, grid_host, st_)
        out["stiffness"]["cpu_baseline"] = cs
        out["stiffness"]["speedup_vs_cpu"] = out["stiffness"]["positions_per_s"] / cs["value"]
        c = run_cpu(args, cfg, xyz, Zs, qs, lvec, grid_host)
        c1["cpu_baseline"] = {k: c[k] for k in ("value", "unit", "cores", "kind", "sample", "grid_pts_per_s")}
        c1["cpu_baseline"]["affinity_cores"] = cores
        c1["speedup_vs_cpu"] = c1["positions_per_s"] / c["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
