"""Training and evaluation of the learned warm start on batched scans.

The network (ppafm_amd.warmstart.WarmStartNet) learns, from FIRE runs of the
method itself, where a probe relaxes at the first step of its stroke, where
FIRE has no continuation to start from: a batch is relaxed with plain FIRE,
its positions are polished to a tight threshold on the GPU (the minima), and
training pairs are cut at step 0 of random pixel columns: the total-force
patch around the plain start (moved by a small pseudo-random offset), and
the polished minimum relative to the patch centre.  Training runs on the GPU
kernels of the network.  With warm starts on (BatchScan.set_warmstart), step
0 starts Newton iterations from the network's estimate and every later step
from the continuation start; FIRE then finishes each step to the reference's
threshold, and falls back to the plain start where Newton does not converge
to a stable minimum.
"""

import numpy as np

from . import device as D
from ._lib import check, lib
from .warmstart import WarmStartNet

POLISH_CONV_F2 = 1e-18
POLISH_MAX_ITERS = 4000

MAX_TARGET = 8.0  # 0.8 A in units of 0.1 A: beyond it the pair is a polish outlier
JITTER = 0.1  # A: training patches are centred up to this far from the plain start


def first_steps(batch, items=None):
    """Global position indices of step 0 of every pixel column of ``items``
    (default: all)."""
    ks = range(batch.nitems) if items is None else items
    out = []
    for k in ks:
        nx, ny, nz = batch._scan_shape(k)
        if nz > 0:
            out.append(batch.pos_off[k] + np.arange(nx * ny, dtype=np.int64) * nz)
    return np.concatenate(out) if out else np.zeros(0, np.int64)


def training_set(batch, nsamples, seed=0, items=None, conv_f2=POLISH_CONV_F2, max_iters=POLISH_MAX_ITERS,
                 max_target=MAX_TARGET, jitter=JITTER):
    """Patches and targets (device arrays, shuffled) at step 0 of the pixel
    columns of ``items`` (default: all), from the last run of a plain-FIRE
    batch, and the polished minima of every position (device).

    Each column gives one pair per round, the patch centre moved by up to
    ``jitter`` A by a hash of (position, round seed); rounds repeat until
    ``nsamples`` pairs are collected, each round in a fresh random order of
    the columns.  Pairs whose target exceeds ``max_target`` (a polish that
    ended far away) are dropped.  The pairs stay on the device: a round's
    targets are made once to find the outliers, then the kept pairs are made
    again (the same hash, so the same pairs) straight into the output."""
    truth, _ = batch.polish(conv_f2=conv_f2, max_iters=max_iters)
    batch.synchronize()
    rng = np.random.default_rng(seed)
    pool = first_steps(batch, items)
    nsamples = int(nsamples)
    P = D.DeviceArray((max(nsamples, 1), 232), np.uint16)
    T = D.DeviceArray((max(nsamples, 1), 3), np.float32)
    n, r = 0, 0
    while n < nsamples and len(pool):
        sel = rng.permutation(pool)[: nsamples - n]
        s = seed * 7919 + r
        _, Tr = batch.ws_samples(truth, sel, jitter=jitter, seed=s)
        keep = np.abs(Tr.download()[: len(sel)]).max(axis=1) <= max_target
        del Tr, _
        k = int(keep.sum())
        if k:
            batch.ws_samples(truth, sel[keep], jitter=jitter, seed=s, out=(P, T), offset=n)
        n += k
        r += 1
        if r > 1000:
            break
    if n < nsamples:
        P2 = D.DeviceArray((max(n, 1), 232), np.uint16)
        T2 = D.DeviceArray((max(n, 1), 3), np.float32)
        if n:
            check(lib().ppafm_memcpy_d2d(P2.ptr, P.ptr, 464 * n, None), "copy")
            check(lib().ppafm_memcpy_d2d(T2.ptr, T.ptr, 12 * n, None), "copy")
        P, T = P2, T2
    return P, T, truth


def train(net, P, T, epochs=1.0, batch=8192, lr=3e-3, lr_end_frac=0.02):
    n = int(P.shape[0]) // batch * batch
    steps = max(1, int(round(epochs * n / batch)))
    return net.train(P, T, n=n, batch=batch, steps=steps, lr=lr, lr_end_frac=lr_end_frac)


def positions_all(batch):
    """All relaxed positions of the last run, (total, 3) float32."""
    out = np.empty((batch.total_positions(), 3), np.float32)
    check(lib().ppafm_memcpy_d2h(out.ctypes.data, lib().ppafm_batch_buffer(batch._h, 1), out.nbytes, batch.stream), "download")
    return out


def _rows(batch, ptr, items):
    """Rows (positions, 3) float32 of the items' positions from a device
    array at ``ptr`` laid out like the batch's positions."""
    if items is None:
        items = range(batch.nitems)
    parts = []
    for k in items:
        lo, hi = int(batch.pos_off[k]), int(batch.pos_off[k + 1])
        a = np.empty((hi - lo, 3), np.float32)
        if hi > lo:
            check(lib().ppafm_memcpy_d2h(a.ctypes.data, ptr + 12 * lo, a.nbytes, batch.stream), "download")
        parts.append(a)
    return np.concatenate(parts) if parts else np.zeros((0, 3), np.float32)


def positions_of(batch, items=None):
    """Relaxed positions of the last run for ``items`` (default all)."""
    return _rows(batch, lib().ppafm_batch_buffer(batch._h, 1), items)


def _stats(d):
    return dict(mean=float(d.mean()), p50=float(np.median(d)), p90=float(np.percentile(d, 90)),
                p99=float(np.percentile(d, 99)), max=float(d.max()))


def timed_run(batch, reps=3):
    """Mean wall time (ms) of ``reps`` graph runs after one warm-up run."""
    batch.run()
    e0, e1 = D.Event(), D.Event()
    e0.record(batch.stream)
    for _ in range(reps):
        batch.run()
    e1.record(batch.stream)
    batch.synchronize()
    return e0.elapsed_ms(e1) / reps


def evaluate(batch, truth, net=None, enable=True, items=None, reps=0, tune=False, **warm):
    """Run the batch with the given warm-start setting (``enable`` False:
    plain FIRE) and return its counters, the distances of its results to the
    polished minima ``truth`` (a device array of all positions; statistics
    over the positions of ``items``, default all), the distances of the
    step-0 estimates to the minima (with a network), the positions of
    ``items`` and the mean graph time over ``reps`` runs (0: untimed).
    ``tune``: launch shape from the autotuner cache for this setting (tuned
    on a miss)."""
    batch.set_warmstart(net, enable=enable, **warm)
    if tune:
        from .batch import tuned

        batch.set_tuning(tuned(batch))
    ms = timed_run(batch, reps) if reps else None
    batch.run()
    batch.synchronize()
    ev, npos, nfb = batch.counters()
    lane_use = batch.lane_use()
    r = positions_of(batch, items)
    rt = _rows(batch, truth.ptr, items)
    out = dict(evals=ev, positions=npos, fallbacks=nfb, lane_use=lane_use, vs_min=_stats(np.linalg.norm(r - rt, axis=1)), ms=ms,
               positions_array=r, tuning=batch.tuning)
    if net is not None and enable:
        ks = range(batch.nitems) if items is None else items
        d = []
        for k in ks:
            nx, ny, nz = batch._scan_shape(k)
            if nz == 0:
                continue
            e = batch.predictions(k).reshape(-1, 3)
            t0 = _rows(batch, truth.ptr, [k])[::nz]
            d.append(np.linalg.norm(e - t0, axis=1))
        out["estimate_vs_min"] = _stats(np.concatenate(d))
    return out


def compare(batch, net, truth=None, items=None, reps=0, tune=False, **warm):
    """Plain FIRE, warm starts without the network (Newton from the plain
    start at step 0) and warm starts with it, on the same batch (``tune``:
    each with its own autotuned launch shape)."""
    batch.set_plain()
    batch.run()
    batch.synchronize()
    if truth is None:
        truth, _ = batch.polish()
        batch.synchronize()
    res = {"plain": evaluate(batch, truth, None, enable=False, items=items, reps=reps, tune=tune),
           "newton": evaluate(batch, truth, None, items=items, reps=reps, tune=tune,
                              **{k: v for k, v in warm.items() if k != "newton_max0"}),
           "learned": evaluate(batch, truth, net, items=items, reps=reps, tune=tune, **warm)}
    rp = res["plain"]["positions_array"]
    for k in ("newton", "learned"):
        res[k]["vs_plain"] = _stats(np.linalg.norm(res[k]["positions_array"] - rp, axis=1))
    batch.set_plain()
    return res
