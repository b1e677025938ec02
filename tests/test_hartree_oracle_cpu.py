"""Pins of the Hartree / density-based grid oracle (oracle/hartree_oracle.py) against
what the mathematics fixes, so that the GPU parity tests compare with a
checked restatement of ppafm/ocl (field.py, FF.cl):

  * trilinear interpolation: node values, exact linear fields, periodicity;
    a 90-degree rotation that maps the grid onto itself permutes values;
  * interp_tip_at: identity on the same grid, zero padding on a larger one;
  * grad: the centred difference of a sampled sine is a cosine times
    sin(2 pi / n) / h;
  * gradPotentialGrid with h = grid step: the same closed form, sign -grad;
  * tip charges in a uniform field: the charges times the field;
  * power_positive: p = 1 preserves the sum;
  * multipole tip density: monopole, dipole and quadrupole moments of the
    normalised Gaussian; tilt rotates the pz dipole into x;
  * FFT cross-correlation: brute force on a small grid, a delta density
    shifts the input, and correlation (not convolution) orientation;
  * the OpenCL vdW variants: force = -dE/dr (numeric) and the r^-6 tail.
"""

import numpy as np
import pytest

import oracle as O
import hartree_oracle as H


def _lvec(o, cell):
    return np.array([o, cell[0], cell[1], cell[2]], dtype=np.float64)


def test_trilinear_nodes_linear_and_periodic():
    rng = np.random.default_rng(0)
    n = (6, 5, 7)
    lv = _lvec([0.3, -0.2, 1.0], [[3.0, 0.2, 0.0], [0.1, 2.5, 0.0], [0.0, 0.3, 3.5]])
    buf = rng.standard_normal(n)
    P = H.points(lv, n)
    assert np.allclose(H.trilinear(buf, lv, P), buf, atol=1e-12)
    # a field linear in the index coordinates is reproduced inside the cell
    s = H.steps(lv, n)
    alpha = np.array([0.7, -1.3, 0.4])
    i, j, k = np.meshgrid(*[np.arange(m) for m in n], indexing="ij")
    lin = alpha[0] * i + alpha[1] * j + alpha[2] * k
    f = rng.uniform(0, 1, (50, 3)) * (np.array(n) - 1)
    pts = lv[0] + f @ s
    assert np.allclose(H.trilinear(lin, lv, pts), f @ alpha, atol=1e-12)
    # periodic in every lattice vector
    q = lv[0] + rng.uniform(-1, 2, (40, 3)) @ lv[1:]
    for a in range(3):
        assert np.allclose(H.trilinear(buf, lv, q + lv[a + 1]), H.trilinear(buf, lv, q), atol=1e-12)


def test_interp_at_rotation_permutes_grid():
    rng = np.random.default_rng(1)
    n = (8, 8, 5)
    lv = _lvec([0, 0, 0], np.diag([4.0, 4.0, 3.0]))
    buf = rng.standard_normal(n)
    R = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    out = H.interp_at(buf, lv, lv, n, rot=R, rot_center=np.zeros(3))
    i, j, k = np.meshgrid(*[np.arange(m) for m in n], indexing="ij")
    assert np.allclose(out, buf[(-j) % 8, i, k], atol=1e-12)
    # no rotation, shifted output grid = shifted values
    lv2 = lv.copy()
    lv2[0] = lv[0] + H.steps(lv, n)[0] * 3
    assert np.allclose(H.interp_at(buf, lv, lv2, n), np.roll(buf, -3, axis=0), atol=1e-12)


def test_interp_tip_at_identity_and_padding():
    rng = np.random.default_rng(2)
    n = (6, 6, 6)
    lv = _lvec([0, 0, 0], np.diag([3.0, 3.0, 3.0]))
    buf = rng.standard_normal(n)
    out = H.interp_tip_at(buf, lv, lv, n)
    # away from the half-way planes the same grid is the identity
    inner = [0, 1, 2, 4, 5]
    assert np.allclose(out[np.ix_(inner, inner, inner)], buf[np.ix_(inner, inner, inner)], atol=1e-12)
    # twice the cell at the same step: the tip sits at the corners, zero elsewhere
    lv2 = _lvec([5, 5, 5], np.diag([6.0, 6.0, 6.0]))
    big = H.interp_tip_at(buf, lv, lv2, (12, 12, 12))
    assert np.allclose(big[:2, :2, :2], buf[:2, :2, :2], atol=1e-12)
    assert np.allclose(big[-2:, -2:, -2:], buf[-2:, -2:, -2:], atol=1e-12)
    assert np.all(big[4:9, :, :] == 0)


def test_grad_of_sampled_sine():
    n = (16, 4, 5)
    h = 0.25
    i = np.arange(n[0])
    buf = np.broadcast_to(np.sin(2 * np.pi * i / n[0])[:, None, None], n).copy()
    g = H.grad(buf, (h, 1.0, 1.0), scale=(-1, -1, -1, 2))
    want = -np.cos(2 * np.pi * i / n[0]) * np.sin(2 * np.pi / n[0]) / h
    assert np.allclose(g[..., 0], want[:, None, None], atol=1e-12)
    assert np.allclose(g[..., 1:3], 0, atol=1e-12) and np.allclose(g[..., 3], 2 * buf)


def test_grad_potential_closed_form():
    n = (12, 3, 4)
    L = 3.0
    lv = _lvec([0.5, 0.0, 0.0], np.diag([L, 1.0, 1.5]))
    x = lv[0][0] + np.arange(n[0]) * L / n[0]
    pot = np.broadcast_to(np.cos(2 * np.pi * (x - lv[0][0]) / L)[:, None, None], n).copy()
    h = L / n[0]
    f = H.grad_potential(pot, lv, lv, n, h)
    want = np.sin(2 * np.pi * (x - lv[0][0]) / L) * np.sin(2 * np.pi * h / L) / h
    assert np.allclose(f[..., 0], want[:, None, None], atol=1e-12)
    assert np.allclose(f[..., 3], pot, atol=1e-12)


def test_tip_charges_uniform_field():
    n = (4, 4, 6)
    lv = _lvec([0, 0, 0], np.diag([2.0, 2.0, 3.0]))
    field = np.broadcast_to(np.array([0.3, -0.2, 1.5, 7.0]), n + (4,)).copy()
    Qs, QZs = [-10.0, 20.0, -10.0, 0.0], [0.1, 0.0, -0.1, 0.0]
    assert np.allclose(H.tip_charges(field, lv, Qs, QZs), 0.0, atol=1e-10)
    assert np.allclose(H.tip_charges(field, lv, [0.5, 0, 0, 0], [0.37, 0, 0, 0]), 0.5 * field, atol=1e-12)


def test_power_positive_preserves_sum():
    rng = np.random.default_rng(3)
    buf = rng.standard_normal((5, 6, 7))
    assert np.isclose(H.power_positive(buf, 1.0).sum(), buf.sum(), rtol=1e-12)
    out = H.power_positive(buf, 1.3, normalize=False)
    assert np.all(out[buf <= 0] == 0) and np.allclose(out[buf > 0], buf[buf > 0] ** 1.3)


@pytest.mark.parametrize("cell", [np.diag([8.0, 8.0, 8.0]), np.array([[6.0, 0, 0], [0.8, 5.9, 0], [0, 0, 6.2]])])
def test_tip_density_moments(cell):
    n = (48, 48, 48)
    sigma = 0.7
    dV = abs(np.linalg.det(cell)) / np.prod(n)
    s = H.tip_density(cell, n, sigma=sigma, multipole={"s": 0.8})
    assert np.isclose(s.sum() * dV, 0.8, rtol=1e-12)
    # the Cartesian position of each node (wrapped coordinates along the lattice)
    L = np.linalg.norm(cell, axis=1)
    cs = []
    for i in range(3):
        c = np.linspace(0, L[i] * (1 - 1 / n[i]), n[i])
        c[c >= L[i] / 2] -= L[i]
        c[c <= -L[i] / 2] += L[i]
        cs.append(c)
    X, Y, Z = np.meshgrid(*cs, indexing="ij")
    U = np.stack([cell[i] / L[i] for i in range(3)])
    r = X[..., None] * U[0] + Y[..., None] * U[1] + Z[..., None] * U[2]
    if np.allclose(cell, np.diag(np.diag(cell))):
        pz = H.tip_density(cell, n, sigma=sigma, multipole={"pz": 1.0})
        assert np.isclose((pz * r[..., 2]).sum() * dV, sigma, rtol=1e-5)
        assert abs(pz.sum() * dV) < 1e-6  # the unpaired +L/2 plane of the wrapped coordinates
        dz2 = H.tip_density(cell, n, sigma=sigma, multipole={"dz2": 1.0})
        assert np.isclose((dz2 * (r[..., 2] / sigma) ** 2).sum() * dV, 1.0, rtol=1e-5)
        assert abs(dz2.sum() * dV) < 1e-6
        t = 0.3
        pt = H.tip_density(cell, n, sigma=sigma, multipole={"pz": 1.0}, tilt=t)
        assert np.isclose((pt * r[..., 0]).sum() * dV, -np.sin(t) * sigma, rtol=1e-5)
        assert np.isclose((pt * r[..., 2]).sum() * dV, np.cos(t) * sigma, rtol=1e-5)
        none = H.tip_density(cell, n, sigma=sigma, multipole=None)
        assert np.allclose(none, s / 0.8, rtol=1e-12)
        assert np.all(H.tip_density(cell, n, sigma=sigma, multipole={"bogus": 3.0}) == 0)


def test_correlation_brute_force_and_orientation():
    rng = np.random.default_rng(4)
    n = (5, 4, 6)
    rho, arr = rng.standard_normal(n), rng.standard_normal(n)
    assert np.allclose(H.correlate_fft(rho, arr, 0.37, scale=-2.0), H.correlate_direct(rho, arr, 0.37, scale=-2.0), atol=1e-12)
    delta = np.zeros(n)
    delta[2, 1, 3] = 1.0
    E = H.correlate_fft(delta, arr, 1.0)
    assert np.allclose(E, np.roll(arr, (-2, -1, -3), axis=(0, 1, 2)), atol=1e-12)


@pytest.mark.parametrize("name,coefs", [("vdw_nodamp_ocl", (3.0,)), ("vdw_invr4_ocl", (2.1, 0.01)), ("vdw_invr8_ocl", (2.4, 0.02)),
                                        ("vdw_r2", (2.1, 0.01)), ("vdw_r4", (2.4, 0.02))])
def test_ocl_vdw_force_and_tail(name, coefs):
    kind = {"vdw_nodamp_ocl": -1, "vdw_r2": 1, "vdw_r4": 2, "vdw_invr4_ocl": 3, "vdw_invr8_ocl": 4}[name]
    adamp = {-1: 0.0, 1: 0.5, 2: 0.5, 3: 0.03, 4: 0.01}[kind]
    rs = np.linspace(0.8, 6.0, 60)
    E = np.zeros(len(rs))
    F = np.zeros(len(rs))
    h = 1e-5
    c = np.zeros(4)
    c[: len(coefs)] = coefs
    lib = O.lib()
    for k, r in enumerate(rs):
        Es = np.zeros(3)
        Fs = np.zeros(3)
        rr = np.array([r - h, r, r + h])
        lib.orc_radial(O.MODEL[name], float(adamp), 3, O._ptr(rr), O._ptr(c), O._ptr(Es), O._ptr(Fs))
        E[k], F[k] = Es[1], Fs[1]
        # the force on the probe is -dE/dr along +x
        assert np.isclose(F[k], -(Es[2] - Es[0]) / (2 * h), rtol=1e-5, atol=1e-9)
    # r^-6 tail: -c6 / r^6 with c6 = c6 (no damping) or 2 E0 R^6
    c6 = coefs[0] if kind == -1 else 2 * coefs[1] * coefs[0] ** 6
    far = rs > 5.0
    assert np.allclose(E[far], -c6 / rs[far] ** 6, rtol=2e-2 if kind in (3, 4) else 1e-5)
