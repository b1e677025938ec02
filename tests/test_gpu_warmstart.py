"""GPU tests of the warm-start network kernels against the FP64 CPU
reference (oracle/warmstart.py, itself pinned by finite differences).

The CPU reference rounds to bf16 exactly where the kernels do (the inputs
of every matrix product), so the two differ by FP32 accumulation order, the
device exp, and the occasional bf16 rounding flip that those cause: outputs
within 2e-3 relative, gradients within 1e-3 relative (L2 per parameter
block) with cosine similarity above 0.9999 (measured: 3e-4 and <2e-5)."""

import numpy as np
import pytest

import warmstart as W
from ppafm_amd import device as D
from ppafm_amd import warmstart as WS

pytestmark = pytest.mark.gpu


def _data(n, seed):
    rng = np.random.default_rng(seed)
    P = rng.normal(size=(n, 5, 5, 3, 3)).astype(np.float32)
    Pb = WS.pack_patches(P)
    Pr = WS.from_bf16(Pb[:, :225]).reshape(n, 5, 5, 3, 3).astype(np.float64)  # what the kernel sees
    T = rng.normal(0, 0.5, (n, 3)).astype(np.float32)
    return Pb, Pr, T


def _theta(seed):
    rng = np.random.default_rng(seed)
    th = WS.init_params(seed).astype(np.float64)
    th[32 * 81:32 * 81 + 32] = rng.normal(0, 0.1, 32)  # nonzero biases
    th[32 * 81 + 32 + 32 * 288:32 * 81 + 64 + 32 * 288] = rng.normal(0, 0.1, 32)
    th[-3:] = rng.normal(0, 0.1, 3)
    return th.astype(np.float32)


def test_forward_matches_cpu_reference():
    print("case: 200 random patches, random weights", flush=True)
    Pb, Pr, T = _data(200, 1)
    th = _theta(2)
    net = WS.WarmStartNet(max_batch=256, params=th)
    out = net.forward(Pb)
    ref, _ = W.forward(th.astype(np.float64), Pr, bf=True)
    err = np.abs(out - ref).max() / np.abs(ref).max()
    print(f"forward max rel err {err:.2e}")
    assert err < 2e-3


@pytest.mark.parametrize("n", [64, 200, 1024])
def test_gradient_matches_cpu_reference(n):
    print(f"case: gradient over {n} samples", flush=True)
    Pb, Pr, T = _data(n, 3 + n)
    th = _theta(4)
    net = WS.WarmStartNet(max_batch=1024, params=th)
    L, g = net.loss_grad(D.DeviceArray.from_host(Pb), D.DeviceArray.from_host(T))
    Lr, gr = W.loss_grad(th.astype(np.float64), Pr, T.astype(np.float64), bf=True)
    assert abs(L - Lr) <= 1e-3 * Lr
    blocks = dict(W1=(0, 2592), b1=(2592, 2624), W2=(2624, 11840), b2=(11840, 11872), W3=(11872, 11968), b3=(11968, 11971))
    for name, (a, b) in blocks.items():
        d = np.linalg.norm(g[a:b] - gr[a:b]) / np.linalg.norm(gr[a:b])
        cos = np.dot(g[a:b], gr[a:b]) / (np.linalg.norm(g[a:b]) * np.linalg.norm(gr[a:b]))
        print(f"n={n} {name}: rel L2 {d:.2e} cos {cos:.6f}")
        assert d < 1e-3 and cos > 0.9999, name


def test_training_reduces_loss():
    # targets a smooth function of the patch: the net must fit it
    print("case: 16384 synthetic samples, 200 Adam steps", flush=True)
    rng = np.random.default_rng(7)
    n = 16384
    P = rng.normal(size=(n, 5, 5, 3, 3)).astype(np.float32)
    T = np.stack([P[:, 2, 2, 1, 0] - 0.5 * P[:, 3, 2, 1, 0], np.tanh(P[:, 2, 1, 1, 1]), 0.3 * P[:, 1, 1, 0, 2] * P[:, 2, 2, 2, 2]], 1).astype(np.float32)
    net = WS.WarmStartNet(max_batch=1024, seed=5)
    Pd = D.DeviceArray.from_host(WS.pack_patches(P))
    Td = D.DeviceArray.from_host(T)
    losses = net.train(Pd, Td, batch=1024, steps=400, lr=3e-3)
    print(f"loss first {losses[:8].mean():.4f} last {losses[-8:].mean():.4f}")
    assert np.isfinite(losses).all()
    assert losses[-8:].mean() < 0.1 * losses[:8].mean()
    # the trained parameters reproduce the device loss on the CPU
    th = net.params()
    Pr = WS.from_bf16(WS.pack_patches(P[:1024])[:, :225]).reshape(1024, 5, 5, 3, 3).astype(np.float64)
    Lr, _ = W.loss_grad(th.astype(np.float64), Pr, T[:1024].astype(np.float64), bf=True)
    L, _ = net.loss_grad(D.DeviceArray.from_host(WS.pack_patches(P[:1024])), D.DeviceArray.from_host(T[:1024]))
    assert abs(L - Lr) <= 1e-2 * Lr


def test_checkpoint_resumes_training(tmp_path):
    """Train, save (parameters, Adam moments, step, seed), train on; a network
    loaded from the checkpoint and trained on the same minibatches ends at the
    same parameters and step."""
    print("case: warm-start network checkpoint and resume", flush=True)
    rng = np.random.default_rng(11)
    n = 4096
    P = D.DeviceArray.from_host(WS.pack_patches(rng.normal(size=(n, 5, 5, 3, 3)).astype(np.float32)))
    T = D.DeviceArray.from_host(rng.normal(size=(n, 3)).astype(np.float32))
    net = WS.WarmStartNet(max_batch=1024, seed=5)
    net.train(P, T, batch=1024, steps=6, lr=1e-3)
    path = tmp_path / "ck.npz"
    net.save(path, meta={"pairs": n})
    th0, m10, m20, step0 = net.state()
    net.train(P, T, batch=1024, steps=4, lr=1e-3)
    a = net.state()
    net2 = WS.WarmStartNet.load(path, max_batch=1024)
    th1, m11, m21, step1 = net2.state()
    assert np.array_equal(th0, th1) and np.array_equal(m10, m11) and np.array_equal(m20, m21) and step0 == step1 == 6
    assert net2.seed == 5 and net2.meta == {"pairs": n}
    net2.train(P, T, batch=1024, steps=4, lr=1e-3)
    b = net2.state()
    assert a[3] == b[3] == 10
    for x, y in zip(a[:3], b[:3]):
        assert np.allclose(x, y, rtol=1e-5, atol=1e-7)
