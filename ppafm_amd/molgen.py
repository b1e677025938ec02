"""Seeded synthetic inputs for the benchmark and parity configurations.

Planar molecules: a honeycomb lattice (1.40 A bonds) grown site by site from
the site nearest the origin, each step adding a uniformly chosen site of the
current frontier (a random-shape patch), centred at x = y = 0 in the z = 0
plane, with z jitter U[-0.3, 0.3] A and no xy jitter.  Elements C/N/O/H are
drawn with probabilities 0.7/0.1/0.1/0.1, charges U[-0.3, 0.3] e shifted to
net zero.  Every molecule owns a numpy ``default_rng`` seeded with
``1_000_000 * config + index``; the draws come in the order: (atom count,
when the config asks for one), growth choices, z jitter, elements, charges.

Scan geometries follow the reference conventions: tip apex positions,
z descending in steps of 0.1 A, CO probe 3 A below the apex (AFMulator
default tipR0, ocl/AFMulator.py:101).
"""

import math

import numpy as np

BOND = 1.40
ELEMENTS = np.array([6, 7, 8, 1], dtype=np.int32)  # C, N, O, H
ELEMENT_P = np.array([0.7, 0.1, 0.1, 0.1])


def _honeycomb_sites(radius):
    """Sites of a honeycomb lattice within ``radius`` of the origin and their
    nearest-neighbour lists."""
    a1 = np.array([math.sqrt(3) * BOND, 0.0])
    a2 = np.array([math.sqrt(3) / 2 * BOND, 1.5 * BOND])
    basis = [np.zeros(2), np.array([0.0, BOND])]
    m = int(radius / BOND) + 3
    pts = []
    for i in range(-m, m + 1):
        for j in range(-m, m + 1):
            for b in basis:
                p = i * a1 + j * a2 + b
                if p @ p <= radius * radius:
                    pts.append(p)
    pts = np.array(pts)
    d2 = ((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1)
    nbr = [np.nonzero((d2[i] > 0.5) & (d2[i] < (1.1 * BOND) ** 2))[0] for i in range(len(pts))]
    return pts, nbr


_SITE_CACHE = {}


def _sites_for(n):
    radius = BOND * (math.sqrt(n) * 1.2 + 4)
    key = round(radius, 3)
    if key not in _SITE_CACHE:
        _SITE_CACHE[key] = _honeycomb_sites(radius)
    return _SITE_CACHE[key]


def grow_patch(n, rng):
    """Random-shape patch of n honeycomb sites (xy, centred)."""
    pts, nbr = _sites_for(n)
    start = int(np.argmin((pts**2).sum(-1)))
    chosen = [start]
    in_patch = {start}
    frontier = []
    front_set = set()
    for j in nbr[start]:
        frontier.append(int(j))
        front_set.add(int(j))
    while len(chosen) < n:
        k = int(rng.integers(len(frontier)))
        s = frontier[k]
        frontier[k] = frontier[-1]
        frontier.pop()
        front_set.discard(s)
        chosen.append(s)
        in_patch.add(s)
        for j in nbr[s]:
            j = int(j)
            if j not in in_patch and j not in front_set:
                frontier.append(j)
                front_set.add(j)
    xy = pts[chosen].copy()
    xy -= xy.mean(axis=0)
    return xy


def molecule(n_atoms, seed):
    """Return xyzs (N,3), Zs (N,), qs (N,) for one seeded planar molecule."""
    rng = np.random.default_rng(seed)
    return _molecule_from_rng(n_atoms, rng)


def _molecule_from_rng(n_atoms, rng):
    xy = grow_patch(n_atoms, rng)
    z = rng.uniform(-0.3, 0.3, n_atoms)
    Zs = ELEMENTS[rng.choice(4, size=n_atoms, p=ELEMENT_P)]
    qs = rng.uniform(-0.3, 0.3, n_atoms)
    qs -= qs.mean()
    xyzs = np.concatenate([xy, z[:, None]], axis=1)
    return xyzs, Zs.astype(np.int32), qs


def config_molecule(config, index=0, n_atoms=None):
    """Molecule ``index`` of a configuration (seed = 1e6*config + index).

    n_atoms None means "draw it": U[8,120] as in the batched configuration."""
    rng = np.random.default_rng(1_000_000 * config + index)
    if n_atoms is None:
        n_atoms = int(rng.integers(8, 121))
    return _molecule_from_rng(n_atoms, rng)


# ------------------------------------------------------------------ configs

# Each config: grid origin, cell lengths (orthogonal), grid points, scan window
# (x, y ranges of the tip apex), apex z top, number of z steps, z step.
CONFIGS = {
    0: dict(natoms=36, origin=(-8.0, -8.0, -6.0), cell=(16.0, 16.0, 20.0), n=(160, 160, 200),
            scan_n=(128, 128), scan_x=(-6.35, 6.35), scan_y=(-6.35, 6.35), z_top=10.0, nz=40, dz=0.1),
    1: dict(natoms=120, origin=(-12.8, -12.8, -8.0), cell=(25.6, 25.6, 32.0), n=(256, 256, 320),
            scan_n=(256, 256), scan_x=(-10.0, 10.0), scan_y=(-10.0, 10.0), z_top=12.0, nz=60, dz=0.1),
    2: dict(natoms=None, origin=(-6.4, -6.4, -6.0), cell=(12.8, 12.8, 16.0), n=(128, 128, 160),
            scan_n=(128, 128), scan_x=(-4.8, 4.8), scan_y=(-4.8, 4.8), z_top=9.0, nz=30, dz=0.1),
    # c4: Au(111) slab + molecule (c4_system), tip tilted by tilt_deg about y
    4: dict(natoms=None, origin=(-20.0, -20.0, -10.0), cell=(40.0, 40.0, 30.0), n=(400, 400, 300),
            scan_n=(400, 400), scan_x=(-18.0, 18.0), scan_y=(-18.0, 18.0), z_top=12.0, nz=60, dz=0.1,
            tilt_deg=10.0, instances=16),
}


# ------------------------------------------------- large tilted case (c4)

AU_Z = 79
FCC_A = 4.08  # A, Au lattice constant
C4_CELLS = (23, 14)  # rectangular fcc(111) cells along x, y (2 atoms per layer each)
C4_LAYERS = 3
C4_TOP = -3.0  # z of the slab's top layer
C4_MOLECULE_ATOMS = 60


def fcc111_slab(m=C4_CELLS[0], n=C4_CELLS[1], layers=C4_LAYERS, a=FCC_A, z_top=C4_TOP):
    """fcc(111) slab on the rectangular cell (d, d*sqrt(3)) of the close-packed
    layer, d = a/sqrt(2) the nearest-neighbour distance; ABC stacking, layer
    spacing a/sqrt(3), top layer at z_top, centred on x = y = 0.  Returns
    positions (N, 3) and the supercell's lateral vectors (2, 3)."""
    d = a / math.sqrt(2.0)
    h = d * math.sqrt(3.0)
    dz = a / math.sqrt(3.0)
    base = np.array([[0.0, 0.0], [0.5 * d, 0.5 * h]])
    shift = np.array([0.5 * d, h / 6.0])  # A -> B -> C site offsets of the close-packed stacking
    i, j = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    cells = np.stack([i.ravel() * d, j.ravel() * h], 1)
    out = []
    for layer in range(layers):
        xy = (cells[:, None, :] + base[None] + (layers - 1 - layer) * shift).reshape(-1, 2)
        z = np.full((len(xy), 1), z_top - layer * dz)
        out.append(np.concatenate([xy, z], 1))
    xyz = np.concatenate(out)
    L = np.array([m * d, n * h])
    xyz[:, :2] = np.mod(xyz[:, :2] + 0.5 * L, L) - 0.5 * L  # wrap into the cell centred on the origin
    return xyz, np.array([[L[0], 0.0, 0.0], [0.0, L[1], 0.0]])


def c4_system(instance=0, images=1):
    """The c4 system of one instance: the Au(111) slab plus a seeded 60-atom
    molecule from the config-0 generator in the z = 0 plane (seed 4e6 +
    instance), with (2*images+1)^2 xy images of the slab supercell.
    Returns xyz (N, 3), Zs, qs (slab neutral) and the supercell vectors."""
    slab, sv = fcc111_slab()
    mol, mZ, mq = config_molecule(4, instance, C4_MOLECULE_ATOMS)
    xyz = np.concatenate([slab, mol])
    Zs = np.concatenate([np.full(len(slab), AU_Z, np.int32), mZ])
    qs = np.concatenate([np.zeros(len(slab)), mq])
    out_x, out_z, out_q = [], [], []
    for a in range(-images, images + 1):
        for b in range(-images, images + 1):
            out_x.append(xyz + a * sv[0] + b * sv[1])
            out_z.append(Zs)
            out_q.append(qs)
    return np.concatenate(out_x), np.concatenate(out_z), np.concatenate(out_q), sv


def config_lvec(cfg):
    o = np.array(cfg["origin"], dtype=np.float64)
    c = cfg["cell"]
    return np.array([o, [c[0], 0, 0], [0, c[1], 0], [0, 0, c[2]]], dtype=np.float64)


def scan_tips(cfg):
    """Tip apex positions (nx, ny, nz, 3) in stroke order (index 0 = top)."""
    nx, ny = cfg["scan_n"]
    xs = np.linspace(cfg["scan_x"][0], cfg["scan_x"][1], nx)
    ys = np.linspace(cfg["scan_y"][0], cfg["scan_y"][1], ny)
    zs = cfg["z_top"] - cfg["dz"] * np.arange(cfg["nz"])
    t = np.zeros((nx, ny, cfg["nz"], 3))
    t[..., 0] = xs[:, None, None]
    t[..., 1] = ys[None, :, None]
    t[..., 2] = zs[None, None, :]
    return t


# Probe and tip (recorded choices where ppafm's CPU and OpenCL defaults
# differ): O probe, monopole charge -0.1 e, tipR0 = (0, 0, 3) A, k_xy = 0.25 N/m, k_r = 30 N/m.
PROBE_Z = 8
PROBE_Q = -0.1
TIP_R0 = (0.0, 0.0, 3.0)
K_LAT = 0.25
K_RAD = 30.0
DF_AMPLITUDE = 0.8  # 8 z-steps of 0.1 A -> n = 8, 9 taps
F0 = 30300.0
K0 = 1800.0
