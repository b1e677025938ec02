"""Hand-derived trajectories of the OpenCL-mode FIRE loop (relaxStrokesTilted,
ppafm/ocl/cl/relax.cl:391-464 with update_FIRE, relax.cl:105-126).

Each case puts the probe in a field where every FIRE iterate has a closed
form, with the tip springs switched off (stiffness 0), so the force is the
grid force alone.  The expected positions, stored forces and iteration counts
below are derived by hand from the FIRE rule, not by running a loop:

* ``constant_field``: a uniform force F0 with |F0|^2 = 2e-8, comparable to
  F2SAFE = 1e-8, two iterations per z-step.  From v = 0 the first iteration
  only accelerates (x += F0 dt^2); the second mixes v with F0 through
  alpha * sqrt(vv / (ff + F2SAFE)) = alpha * dt * rho, rho = sqrt(ff / (ff +
  F2SAFE)).  Per z-step the probe moves F0 (0.75 - 0.25 a1 (1 - rho)) with a1
  = 0.1 * 0.99^(2 k + 1) at step k: alpha persists along the stroke
  (relax.cl:415-426) and F2SAFE sits inside the square root (relax.cl:121).
* ``stiff_field``: a restoring force -K (x - xc) along x, K = 500, so that
  every leap-frog move overshoots (K dt^2 > 2): v . f < 0 at every iteration
  after the first, which zeroes v, halves dt down to dtmin = 0.1 dtmax
  (relax.cl:116-118, 419) and resets alpha.  Each move multiplies the
  offset e = x - xc by (1 - K dt^2) with dt = 0.5, 0.25, 0.125, 0.0625 and
  then dtmin = 0.05 (K dtmin^2 = 1.25, factor -0.25).  dt persists into the
  next z-step (relax.cl:415-426): there its first move uses dt = min(dtmax,
  1.1 dtmin) = 0.055 (factor -0.5125), not dtmax.  The stored force is the
  grid force of the last evaluation, before the final move (relax.cl:446-456).
* ``zero_field``: no force at all.  vv = ff = 0, so only F2SAFE keeps
  sqrt(vv / (ff + F2SAFE)) finite; the probe stays exactly at its start and
  the loop stops after one evaluation (the test |f|^2 < F2CONV follows the
  move, relax.cl:444).
* ``weak_field``: |F0|^2 = 0.5e-8 < F2CONV: one iteration that still moves
  the probe by F0 dt^2 = 0.25 F0 before the convergence test.
* ``strong_field``: |F0|^2 = 1e-4 never converges: every z-step stops at the
  128-iteration cap (N_RELAX_STEP_MAX, relax.cl:90).
"""

import numpy as np

N = (16, 16, 16)
H = 0.1  # grid step (A); index coordinates = position / H
DT, DAMP = 0.5, 0.1  # relax_params (dt, damping) of the OpenCL scanner


def rot_y(deg):
    t = np.deg2rad(deg)
    return np.array([[np.cos(t), 0, np.sin(t)], [0, 1, 0], [-np.sin(t), 0, np.cos(t)]])


def _uniform(F0, E0=0.25):
    FE = np.zeros(N + (4,))
    FE[..., :3] = F0
    FE[..., 3] = E0
    return FE


def _case(FE, R, start, nz, max_iters, step=-0.1):
    """Common fields of a case: the field, the tip frame rows R (A, B, C), the
    stroke (tip start, step along C), springs off, equilibrium offset 3 A
    along -C."""
    return dict(FE=FE, R=np.asarray(R, float), tip0=np.asarray(start, float), nz=nz, max_iters=max_iters, step=step,
                dpos0=np.array([0.0, 0.0, -3.0, 3.0]))


def probe_start(c):
    """Probe at the first tip position: tip + R^T (0, 0, -3)."""
    return c["tip0"] + c["R"].T @ c["dpos0"][:3]


def dtip(c):
    return c["R"][2] * c["step"]


def constant_field(f2safe):
    u = np.array([0.6, 0.0, 0.8])
    F0 = np.sqrt(2e-8) * u
    R = rot_y(10.0)
    c = _case(_uniform(F0), R, (0.8, 0.8, 4.0), nz=3, max_iters=2)
    rho = np.sqrt(2e-8 / (2e-8 + f2safe))
    x = probe_start(c)
    paths = []
    for k in range(c["nz"]):
        a1 = DAMP * 0.99 ** (2 * k + 1)
        x = x + F0 * (0.75 - 0.25 * a1 * (1 - rho))
        paths.append(x.copy())
        x = x + dtip(c)
    fe = np.append(R @ F0, 0.25)
    c.update(paths=np.array(paths), forces=np.tile(fe, (c["nz"], 1)), iters=np.full(c["nz"], 2))
    return c


def stiff_field(dtmin):
    K, xc, e0 = 500.0, 0.75, 1e-6
    P = np.arange(N[0]) * H
    FE = np.zeros(N + (4,))
    FE[..., 0] = (-K * (P - xc))[:, None, None]
    gE, E0 = 0.3, -0.2
    FE[..., 3] = (E0 + gE * P)[:, None, None]
    R = np.eye(3)
    c = _case(FE, R, (xc + e0, 0.8, 4.0), nz=2, max_iters=6)
    # step 0: dt = 0.5 (first move, v = 0), then v.f < 0 every time:
    # dt = 0.25, 0.125, 0.0625, max(dtmin, 0.03125) = dtmin, dtmin
    # step 1: dt = min(0.5, 1.1 dtmin), then dtmin five times
    dts = [[0.5, 0.25, 0.125, 0.0625, dtmin, dtmin], [min(DT, 1.1 * dtmin)] + [dtmin] * 5]
    e = e0
    x0 = probe_start(c)
    paths, forces = [], []
    for k in range(2):
        for dt in dts[k]:
            e_prev = e
            e = e * (1 - K * dt * dt)
        paths.append(x0 + np.array([e - e0, 0.0, 0.0]) + k * dtip(c))
        forces.append([-K * e_prev, 0.0, 0.0, E0 + gE * (xc + e_prev)])
    c.update(paths=np.array(paths), forces=np.array(forces), iters=np.full(2, 6))
    return c


def zero_field():
    c = _case(_uniform(np.zeros(3), E0=0.5), rot_y(-7.0), (0.8, 0.8, 4.0), nz=2, max_iters=128)
    x = probe_start(c)
    c.update(paths=np.array([x, x + dtip(c)]), forces=np.tile([0, 0, 0, 0.5], (2, 1)), iters=np.ones(2, int))
    return c


def weak_field():
    F0 = np.sqrt(0.5e-8) * np.array([0.0, 0.6, -0.8])
    R = rot_y(10.0)
    c = _case(_uniform(F0), R, (0.8, 0.8, 4.0), nz=2, max_iters=128)
    x = probe_start(c)
    p0 = x + 0.25 * F0
    c.update(paths=np.array([p0, p0 + dtip(c) + 0.25 * F0]), forces=np.tile(np.append(R @ F0, 0.25), (2, 1)),
             iters=np.ones(2, int))
    return c


def strong_field():
    F0 = 1e-2 * np.array([0.0, 0.0, 1.0])
    c = _case(_uniform(F0), np.eye(3), (0.8, 0.8, 4.0), nz=2, max_iters=128)
    c.update(paths=None, forces=np.tile(np.append(F0, 0.25), (2, 1)), iters=np.full(2, 128))
    return c
