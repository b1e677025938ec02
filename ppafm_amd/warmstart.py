"""Learned warm start for the FIRE relaxation (no counterpart in the
reference): the network object of libppafm_hip.so (``ppafm_ws_*`` in
include/ppafm_hip.h).

The network maps the total-force patch around a probe's start position (5x5x3
grid nodes x 3 force components, each component scaled by its RMS over the
patch, bf16) to the relaxed probe
position as a displacement from the patch's centre node in units of
``SCALE`` = 0.1 A.  Forward, backward and Adam run on the GPU in hand-written
kernels (bf16 MFMA, FP32 accumulation, FP32 master weights).
"""

import ctypes

import numpy as np

from . import device as D
from ._lib import HipError, check, lib

PATCH_STRIDE = 232
PATCH_VALUES = 225
SCALE = 0.1
NPARAM = 11971


def init_params(seed=0):
    """He-style random initial parameters (W1 b1 W2 b2 W3 b3)."""
    rng = np.random.default_rng(seed)
    parts = [rng.normal(0, np.sqrt(2.0 / 81), 32 * 81), np.zeros(32), rng.normal(0, np.sqrt(2.0 / 288), 32 * 288),
             np.zeros(32), rng.normal(0, np.sqrt(1.0 / 32), 3 * 32), np.zeros(3)]
    return np.concatenate(parts).astype(np.float32)


def to_bf16(x):
    """float32 array -> uint16 bf16 bit patterns (round to nearest even)."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    return (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)


def from_bf16(b):
    return (np.asarray(b, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def pack_patches(P):
    """(n, 5, 5, 3, 3) float patches -> (n, 232) bf16 rows."""
    P = np.asarray(P, dtype=np.float32).reshape(len(P), PATCH_VALUES)
    out = np.zeros((len(P), PATCH_STRIDE), dtype=np.uint16)
    out[:, :PATCH_VALUES] = to_bf16(P)
    return out


CHECKPOINT_FORMAT = 1


def read_checkpoint(path):
    """A training checkpoint (WarmStartNet.save) as a dict: theta, m1, m2
    (float32, NPARAM each), step, seed, meta (dict); the arrays are checked
    for shape and type."""
    with np.load(path, allow_pickle=False) as z:
        d = {k: z[k] for k in z.files}
    if int(d.get("format", -1)) != CHECKPOINT_FORMAT:
        raise ValueError(f"{path}: not a warm-start checkpoint of format {CHECKPOINT_FORMAT}")
    for k in ("theta", "m1", "m2"):
        if d[k].dtype != np.float32 or d[k].shape != (NPARAM,):
            raise ValueError(f"{path}: {k} must be {NPARAM} float32 values")
    import json

    return dict(theta=d["theta"], m1=d["m1"], m2=d["m2"], step=int(d["step"]), seed=int(d["seed"]),
                meta=json.loads(str(d["meta"])))


class WarmStartNet:
    """The network and its training state on the GPU."""

    def __init__(self, max_batch=8192, params=None, seed=0):
        self._lib = lib()
        h = lib().ppafm_ws_create(int(max_batch))
        if not h:
            raise HipError("ppafm_ws_create failed: " + lib().ppafm_last_error().decode(errors="replace"))
        self._h = h
        self.max_batch = int(max_batch)
        self.seed = int(seed)
        assert lib().ppafm_ws_nparams() == NPARAM
        self.set_params(init_params(seed) if params is None else params)

    # ------------------------------------------------------------ checkpoints
    def state(self):
        """(theta, Adam m1, Adam m2, step) from the GPU."""
        th, m1, m2 = (np.empty(NPARAM, np.float32) for _ in range(3))
        step = ctypes.c_longlong()
        check(lib().ppafm_ws_get_state(self._h, th.ctypes.data, m1.ctypes.data, m2.ctypes.data, ctypes.byref(step)),
              "ws_get_state")
        return th, m1, m2, int(step.value)

    def save(self, path, meta=None):
        """Write a training checkpoint (npz): the parameters, both Adam moments,
        the Adam step, the initialisation seed and ``meta`` (a JSON-serialisable
        dict, e.g. the training-set seed and sample count) for a resume."""
        import json

        th, m1, m2, step = self.state()
        np.savez(path, format=np.int64(CHECKPOINT_FORMAT), theta=th, m1=m1, m2=m2, step=np.int64(step),
                 seed=np.int64(self.seed), meta=np.array(json.dumps(meta or {})))

    @classmethod
    def load(cls, path, max_batch=8192):
        """A network with the checkpoint's parameters and Adam state: training
        resumes where the saved one stopped."""
        d = read_checkpoint(path)
        net = cls(max_batch=max_batch, params=d["theta"], seed=d["seed"])
        check(lib().ppafm_ws_set_state(net._h, d["theta"].ctypes.data, d["m1"].ctypes.data, d["m2"].ctypes.data,
                                       d["step"]), "ws_set_state")
        net.meta = d["meta"]
        return net

    def set_params(self, theta):
        theta = np.ascontiguousarray(theta, dtype=np.float32)
        if theta.size != NPARAM:
            raise ValueError(f"expected {NPARAM} parameters")
        check(lib().ppafm_ws_set_params(self._h, theta.ctypes.data, NPARAM), "ws_set_params")

    def params(self):
        out = np.empty(NPARAM, np.float32)
        check(lib().ppafm_ws_get_params(self._h, out.ctypes.data, NPARAM), "ws_get_params")
        return out

    def forward_dev(self, n, patches, out, stream=None):
        check(lib().ppafm_ws_forward(self._h, int(n), patches.ptr, out.ptr, stream), "ws_forward")

    def forward(self, patches_bf16):
        """Host convenience: (n, 232) bf16 rows -> (n, 3) outputs."""
        p = D.DeviceArray.from_host(np.ascontiguousarray(patches_bf16, dtype=np.uint16))
        o = D.DeviceArray((len(patches_bf16), 3), np.float32)
        self.forward_dev(len(patches_bf16), p, o)
        D.synchronize()
        return o.download()

    def loss_grad(self, patches, targets):
        """patches: DeviceArray (n, 232) uint16, targets: DeviceArray (n, 3)."""
        g = np.empty(NPARAM, np.float32)
        L = np.empty(1, np.float32)
        check(lib().ppafm_ws_loss_grad(self._h, int(patches.shape[0]), patches.ptr, targets.ptr, g.ctypes.data, L.ctypes.data), "ws_loss_grad")
        return float(L[0]), g

    def train(self, patches, targets, n=None, batch=4096, steps=0, lr=3e-3, lr_end_frac=0.02, stream=None):
        """Adam over consecutive minibatches of device arrays (shuffle them
        first).  Returns the per-step losses."""
        n = int(patches.shape[0] if n is None else n)
        nsteps = int(steps) if steps > 0 else n // batch
        losses = np.zeros(nsteps, np.float32)
        check(lib().ppafm_ws_train(self._h, n, patches.ptr, targets.ptr, int(batch), nsteps, float(lr), float(lr_end_frac),
                                   losses.ctypes.data, stream), "ws_train")
        return losses

    def close(self):
        if getattr(self, "_h", None):
            self._lib.ppafm_ws_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
