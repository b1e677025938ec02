"""FP64 CPU reference of the warm-start network -- TEST INFRASTRUCTURE.

Only tests/ and bench.py's CPU leg may import this module, and only as the
checker.  The network itself has no counterpart in the reference (it is the
learned addition on top of the FIRE relaxation); this module is the CPU
restatement of what the HIP kernels in ppafm_amd/csrc/warmstart.hip compute.
Its backward pass is pinned by central finite differences of its own loss
(tests/test_oracle_cpu.py), which depend only on the forward pass.

Network (valid 3D convolutions, SiLU after layers 1 and 2):
  patch (5, 5, 3) x 3 channels  -- total force at grid nodes, RMS-normalised
  conv1 3 -> 32, kernel (3, 3, 3)  -> (3, 3, 1) x 32
  conv2 32 -> 32, kernel (3, 3, 1) -> (1, 1, 1) x 32
  conv3 32 -> 3, kernel 1          -> predicted displacement (3)
Parameter layout (one flat float vector, 11,971 values):
  W1 [32][81]  k = ((dx*3 + dy)*3 + dz)*3 + c
  b1 [32]
  W2 [32][288] k = p*32 + c, p = ox*3 + oy
  b2 [32]
  W3 [3][32]
  b3 [3]
Loss: mean over samples and components of (out - target)^2.

``bf16=True`` rounds what the GPU kernels round to bfloat16 (the inputs of
every matrix product), so GPU and CPU then differ only by FP32 accumulation.
"""

import numpy as np

N1, K1, N2, K2, N3 = 32, 81, 32, 288, 3
SIZES = [("W1", (N1, K1)), ("b1", (N1,)), ("W2", (N2, K2)), ("b2", (N2,)), ("W3", (N3, N2)), ("b3", (N3,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(theta):
    out, o = {}, 0
    for name, shp in SIZES:
        n = int(np.prod(shp))
        out[name] = np.asarray(theta[o:o + n], dtype=np.float64).reshape(shp)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).reshape(-1) for name, _ in SIZES])


def bf16(x):
    """Round to bfloat16 (nearest, ties to even) and back to float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def init(seed=0):
    rng = np.random.default_rng(seed)
    d = {"W1": rng.normal(0, np.sqrt(2.0 / K1), (N1, K1)), "b1": np.zeros(N1),
         "W2": rng.normal(0, np.sqrt(2.0 / K2), (N2, K2)), "b2": np.zeros(N2),
         "W3": rng.normal(0, np.sqrt(1.0 / N2), (N3, N2)), "b3": np.zeros(N3)}
    return pack(d)


def im2col(P):
    """P (S, 5, 5, 3, 3) -> A1 (S, 9, 81)."""
    S = P.shape[0]
    A = np.zeros((S, 9, K1))
    for ox in range(3):
        for oy in range(3):
            blk = P[:, ox:ox + 3, oy:oy + 3, :, :]  # (S, dx, dy, dz, c)
            A[:, ox * 3 + oy, :] = blk.reshape(S, K1)
    return A


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(theta, P, bf=False):
    w = unpack(theta)
    r = bf16 if bf else (lambda x: np.asarray(x, dtype=np.float64))
    A1 = r(im2col(P))
    Z1 = A1 @ r(w["W1"]).T + w["b1"]  # (S, 9, 32)
    H1 = silu(Z1)
    A2 = r(H1.reshape(len(P), K2))
    Z2 = A2 @ r(w["W2"]).T + w["b2"]
    H2 = silu(Z2)
    out = r(H2) @ r(w["W3"]).T + w["b3"]
    return out, dict(A1=A1, Z1=Z1, A2=A2, Z2=Z2, H2=H2)


def loss_grad(theta, P, T, bf=False):
    """Loss and its gradient (flat, same layout as theta)."""
    w = unpack(theta)
    r = bf16 if bf else (lambda x: np.asarray(x, dtype=np.float64))
    out, c = forward(theta, P, bf)
    S = len(P)
    L = np.mean((out - T) ** 2)
    dY = 2.0 * (out - T) / (S * N3)
    g = {}
    g["W3"] = dY.T @ r(c["H2"])
    g["b3"] = dY.sum(0)
    dH2 = dY @ r(w["W3"])
    dZ2 = dH2 * dsilu(c["Z2"])
    g["W2"] = r(dZ2).T @ c["A2"]
    g["b2"] = dZ2.sum(0)
    dA2 = r(dZ2) @ r(w["W2"])  # (S, 288)
    dZ1 = dA2.reshape(S, 9, N1) * dsilu(c["Z1"])
    g["W1"] = np.einsum("spj,spk->jk", r(dZ1), c["A1"])
    g["b1"] = dZ1.sum((0, 1))
    return L, pack(g)


def normalise_patch(F):
    """Total-force patch (S, 5, 5, 3, 3) -> each component scaled by its RMS
    over the patch, and the scales (S, 3)."""
    s = np.sqrt(np.mean(F.reshape(len(F), -1, 3) ** 2, axis=1)) + 1e-30
    return F / s[:, None, None, None, :], s
