"""FP64 numpy oracle of the Hartree / full-density-based force-field grids --
TEST INFRASTRUCTURE.

Only tests/ may import this module, as the checker of the GPU path.  Each
function restates one piece of ppafm/ocl (cited per function) with numpy in
FP64 on the same inputs; the reference runs these in FP32 OpenCL (pinned
instead by the closed forms and brute-force checks in
tests/test_hartree_oracle_cpu.py).

Grids are C-order (nx, ny, nz) arrays with a lattice lvec (4,3) = origin,
a, b, c; step vectors a/nx, b/ny, c/nz.
"""

import numpy as np


def steps(lvec, n):
    lvec = np.asarray(lvec, dtype=np.float64)
    return np.stack([lvec[i + 1] / n[i] for i in range(3)])


def points(lvec, n):
    """Cartesian points origin + i a' + j b' + k c' of a grid, (nx,ny,nz,3),
    summed in that order like the kernels."""
    s = steps(lvec, n)
    o = np.asarray(lvec, dtype=np.float64)[0]
    i, j, k = np.meshgrid(np.arange(n[0]), np.arange(n[1]), np.arange(n[2]), indexing="ij")
    return o + i[..., None] * s[0] + j[..., None] * s[1] + k[..., None] * s[2]


def trilinear(buf, lvec, p, zero_origin=False):
    """linearInterpB (FF.cl:33-71): periodic trilinear interpolation of buf
    (nx,ny,nz[,c]) at Cartesian points p (...,3)."""
    n = buf.shape[:3]
    T = np.linalg.inv(steps(lvec, n))  # coordinates = (p - o) @ T
    o = np.zeros(3) if zero_origin else np.asarray(lvec, dtype=np.float64)[0]
    u = (p - o) @ T
    fl = np.floor(u)
    d = np.minimum(u - fl, np.nextafter(1.0, 0.0))
    i0 = np.mod(fl.astype(np.int64), np.array(n))
    i1 = (i0 + 1) % np.array(n)
    out = 0.0
    for cx in (0, 1):
        for cy in (0, 1):
            for cz in (0, 1):
                ix = i1[..., 0] if cx else i0[..., 0]
                iy = i1[..., 1] if cy else i0[..., 1]
                iz = i1[..., 2] if cz else i0[..., 2]
                w = (d[..., 0] if cx else 1 - d[..., 0]) * (d[..., 1] if cy else 1 - d[..., 1]) * (d[..., 2] if cz else 1 - d[..., 2])
                v = buf[ix, iy, iz]
                out = out + (w[..., None] * v if v.ndim > w.ndim else w * v)
    return out


def interp_at(buf, lvec_in, lvec_out, n_out, rot=np.eye(3), rot_center=np.zeros(3)):
    """interp_at (FF.cl:1010-1046, DataGrid.interp_at field.py:528-575)."""
    p = points(lvec_out, n_out)
    rc = np.asarray(rot_center, dtype=np.float64)
    q = rc + (p - rc) @ np.asarray(rot, dtype=np.float64).T
    return trilinear(buf, lvec_in, q)


def interp_tip_at(buf, lvec_in, lvec_out, n_out):
    """interp_tip_at (FF.cl:1049-1093, TipDensity.interp_at field.py:607-660):
    wrapped out indices, zero outside half the input cell (coordinates
    inv(A) p, A rows a, b, c of lvec_in)."""
    n_out = tuple(int(v) for v in n_out)
    idx = [np.arange(m) for m in n_out]
    for a in range(3):
        idx[a] = np.where(idx[a] > n_out[a] // 2, idx[a] - n_out[a], idx[a])
    i, j, k = np.meshgrid(*idx, indexing="ij")
    s = steps(lvec_out, n_out)
    p = i[..., None] * s[0] + j[..., None] * s[1] + k[..., None] * s[2]
    frac = p @ np.linalg.inv(np.asarray(lvec_in, dtype=np.float64)[1:]).T
    out = trilinear(buf, lvec_in, p, zero_origin=True)
    return np.where((np.abs(frac) > 0.5).any(axis=-1), 0.0, out)


def grad(buf, step, scale=(1.0, 1.0, 1.0, 1.0)):
    """grad (FF.cl:920-961): periodic centred differences, (nx,ny,nz,4) C order."""
    out = np.empty(buf.shape + (4,))
    for a in range(3):
        out[..., a] = 0.5 * (np.roll(buf, -1, axis=a) - np.roll(buf, 1, axis=a)) / step[a]
    out[..., 3] = buf
    return out * np.asarray(scale, dtype=np.float64)


def grad_potential(pot, pot_lvec, lvec, n, h):
    """gradPotentialGrid (FF.cl:964-1007): (-dV/dx, -dV/dy, -dV/dz, V) on the
    grid (lvec, n) from interpolations at p -/+ h."""
    p = points(lvec, n)
    out = np.empty(tuple(n) + (4,))
    for a in range(3):
        e = np.zeros(3)
        e[a] = h
        out[..., a] = 0.5 * (trilinear(pot, pot_lvec, p - e) - trilinear(pot, pot_lvec, p + e)) / h
    out[..., 3] = trilinear(pot, pot_lvec, p)
    return out


def tip_charges(field, lvec, Qs, QZs):
    """The charge terms of evalLJC_Hartree (FF.cl:906-911) on field's grid."""
    p = points(lvec, field.shape[:3])
    out = np.zeros(field.shape)
    for q, dz in zip(Qs, QZs):
        if q != 0:
            out += q * trilinear(field, lvec, p + np.array([0.0, 0.0, dz]))
    return out


def power_positive(buf, p, normalize=True):
    """DataGrid.power_positive (field.py:377-445, FF.cl:1096-1176)."""
    scale = (buf.sum() / buf[buf > 0].sum()) ** p if normalize else 1.0
    return scale * np.maximum(buf, 0.0) ** p


HARMONICS = ("s", "px", "py", "pz", "dz2", "dx2", "dy2", "dxy", "dxz", "dyz")


def harmonic(kind, X, Y, Z, tilt=0.0):
    """getSphericalHarmonic (fieldFFT.py:58-96) after rotZX."""
    ca, sa = np.cos(tilt), np.sin(tilt)
    Z, X = ca * Z - sa * X, sa * Z + ca * X
    return {
        "s": 1.0, "px": X, "py": Y, "pz": Z,
        "dz2": 0.25 * (2 * Z**2 - X**2 - Y**2), "dx2": 0.25 * (2 * X**2 - Y**2 - Z**2), "dy2": 0.25 * (2 * Y**2 - X**2 - Z**2),
        "dxy": X * Y, "dxz": X * Z, "dyz": Y * Z,
    }.get(kind, 0.0)


def tip_density(lvec, shape, center=(0, 0, 0), sigma=0.71, multipole=None, tilt=0.0):
    """MultipoleTipDensity._make_tip_density (field.py:671-690) with
    getProbeDensity (fieldFFT.py:99-118), in FP64."""
    lvec = np.asarray(lvec, dtype=np.float64)
    L = np.linalg.norm(lvec, axis=1)
    cs = []
    for i in range(3):
        c = np.linspace(0, L[i] * (1 - 1 / shape[i]), shape[i]) - center[i]
        c[c >= L[i] / 2] -= L[i]
        c[c <= -L[i] / 2] += L[i]
        cs.append(c)
    X, Y, Z = np.meshgrid(*cs, indexing="ij")
    Lmat = np.stack([lvec[i] / np.linalg.norm(lvec[i]) for i in range(3)])
    mat = Lmat.T
    dd = L / np.asarray(shape)
    rx = X * mat[0, 0] + Y * mat[0, 1] + Z * mat[0, 2]
    ry = X * mat[1, 0] + Y * mat[1, 1] + Z * mat[1, 2]
    rz = X * mat[2, 0] + Y * mat[2, 1] + Z * mat[2, 2]
    radial = np.exp(-(rx**2 + ry**2 + rz**2) / (2 * sigma**2))
    radial /= radial.sum() * abs(np.linalg.det(mat)) * dd[0] * dd[1] * dd[2]
    if multipole is None:
        return radial
    rho = np.zeros(radial.shape)
    for kind, coef in multipole.items():
        rho += radial * coef * harmonic(kind, rx / sigma, ry / sigma, rz / sigma, tilt)
    return rho


def correlate_fft(rho, arr, cell_vol, scale=1.0):
    """FFTCrossCorrelation.correlate (field.py:804-843) with numpy FFTs."""
    return np.real(np.fft.ifftn(np.conj(np.fft.fftn(rho)) * np.fft.fftn(arr))) * (scale * cell_vol)


def correlate_direct(rho, arr, cell_vol, scale=1.0):
    """E[r] = scale * cell_vol * sum_s rho[s] arr[s + r] by brute force
    (the definition the FFT form computes; small grids only)."""
    n = arr.shape
    E = np.zeros(n)
    for i in range(n[0]):
        for j in range(n[1]):
            for k in range(n[2]):
                E[i, j, k] = (rho * np.roll(arr, (-i, -j, -k), axis=(0, 1, 2))).sum()
    return E * (scale * cell_vol)
