"""The OpenCL-mode FIRE of the device kernel (mode 1 of relax_dev.hip, the
path of RelaxedScanner, AFMulator and the tilted batches) on the
hand-derived trajectories of tests/fire_cases.py: dt and alpha kept along
the stroke, dtmin = 0.1 dtmax, F2SAFE inside the square root, the test after
the move, the 128-iteration cap and the stored force of the last evaluation
(relax.cl:90-126, 391-464).  Positions come back in FP64; forces are stored
as FP32 (as the reference's float4 output)."""

import numpy as np
import pytest

import fire_cases as FC
from ppafm_amd import device as D

pytestmark = pytest.mark.gpu

F32 = float(np.float32(1e-8))  # F2CONV, F2SAFE (1e-8f)
DTMIN_FRAC = float(np.float32(0.1))


def run(c):
    n = FC.N
    cell = np.diag([n[0] * FC.H, n[1] * FC.H, n[2] * FC.H])
    grid = D.DeviceArray.from_host(np.ascontiguousarray(c["FE"]), np.float64)
    rp = D.relax_params(mode=1, alg=1, max_iters=c["max_iters"], conv_f2=F32, dt=FC.DT, damping=FC.DAMP,
                        dtmin_frac=DTMIN_FRAC, f2safe=F32)
    tip = D.tip_desc(c["dpos0"][:3], (0.0, 0.0, 0.0), 0.0, c["dpos0"][3], rot=c["R"])
    scan = D.scan_desc(1, 1, c["nz"], p0=c["tip0"], dstep=FC.dtip(c))
    pos, force, iters, status = D.relax(n, cell, grid, rp, tip, scan, pos=True, force=True, iters=True, pos_f32=False,
                                        status=True)
    D.synchronize()
    run.status = status.download()[0, 0]
    return force.download()[0, 0], pos.download()[0, 0], iters.download()[0, 0]


CASES = {
    "constant": lambda: FC.constant_field(F32),
    "stiff": lambda: FC.stiff_field(FC.DT * DTMIN_FRAC),
    "zero": FC.zero_field,
    "weak": FC.weak_field,
    "strong": FC.strong_field,
}


@pytest.mark.parametrize("name", list(CASES))
def test_device_fire_follows_hand_derivation(name):
    print(f"case: hand-derived OpenCL-mode FIRE, {name} field", flush=True)
    c = CASES[name]()
    fe, paths, it = run(c)
    assert np.array_equal(it, c["iters"]), (it, c["iters"])
    if c["paths"] is not None:
        err = np.abs(paths - c["paths"]).max()
        print(f"{name}: max |dr| {err:.3g} A")
        assert err <= (1e-9 if name == "stiff" else 1e-12)
    else:
        assert np.isfinite(paths).all()
    scale = max(np.abs(c["forces"]).max(), 1e-30)
    assert np.abs(fe - c["forces"]).max() <= 1e-6 * scale + 1e-12
    # per-position status: the constant (|F|^2 = 2e-8 above the 1e-8 stop
    # rule), stiff (growing oscillation) and strong fields never converge, so
    # the cap ends every step; the zero and weak fields converge at once
    capped = name in ("constant", "stiff", "strong")
    assert np.array_equal(run.status, np.full(c["nz"], 1 if capped else 0, np.uint8)), run.status


@pytest.mark.parametrize("mode", [0, 1])
def test_status_flags_non_finite_grid(mode):
    """A grid of NaN forces: every step runs to the cap (no stop-rule test
    passes on NaN) and ends at a NaN position, flagged with both bits."""
    print(f"case: NaN grid, mode {mode}, 2 columns of 3 steps", flush=True)
    n = FC.N
    cell = np.diag([n[0] * FC.H, n[1] * FC.H, n[2] * FC.H])
    FE = np.full(n + (4,), np.nan)
    grid = D.DeviceArray.from_host(FE, np.float64)
    rp = D.relax_params(mode=mode, alg=1, max_iters=20, conv_f2=F32, dt=FC.DT, damping=FC.DAMP, dtmin_frac=DTMIN_FRAC,
                        f2safe=F32 if mode == 1 else 0.0)
    tip = D.tip_desc((0.0, 0.0, -3.0), (0.25, 0.25, 0.0), 30.0, 3.0)
    scan = D.scan_desc(2, 1, 3, p0=(0.8, 0.8, 4.0), da=(0.1, 0.0, 0.0), dstep=(0.0, 0.0, -0.1))
    pos, force, iters, status = D.relax(n, cell, grid, rp, tip, scan, pos=True, force=True, iters=True, pos_f32=False,
                                        status=True)
    D.synchronize()
    assert np.isnan(pos.download()).all()
    assert (iters.download() == 20).all()
    assert (status.download() == 3).all()
