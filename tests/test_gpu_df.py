"""The Python mirrors common.Fz2df and common.Fz2df_tilt (ppafm/common.py:
421-445; the convolutions run on the GPU) against the oracle's fz2df
(np.convolve valid mode with the Giessibl weights, pinned by the reference's
golden weight vectors and a brute-force convolution): the tilted variant is
the per-axis conversion with the step length |d| projected on d."""

import numpy as np
import pytest

import oracle as O
from ppafm_amd import common as C

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("amp,dz,nz", [(0.8, 0.1, 40), (1.0, 0.2, 23), (0.35, 0.05, 31)])
def test_fz2df_matches_oracle(amp, dz, nz):
    print(f"case: Fz2df, amplitude {amp}, dz {dz}, {nz} steps", flush=True)
    rng = np.random.default_rng(int(nz))
    F = rng.normal(0, 0.1, (6, 5, nz))
    got = C.Fz2df(F, dz, 1800.0, 30300.0, amplitude=amp)
    ref = O.fz2df(F, dz, 1800.0, 30300.0, amplitude=amp)
    assert got.shape == ref.shape
    assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("d", [(0.0, 0.0, -0.1), (0.0173, 0.0, -0.0985), (0.02, -0.01, -0.09)])
def test_fz2df_tilt_matches_oracle(d):
    print(f"case: Fz2df_tilt, step {d}", flush=True)
    rng = np.random.default_rng(3)
    F = rng.normal(0, 0.1, (5, 4, 33, 3))
    got = C.Fz2df_tilt(F, d, 1800.0, 30300.0, amplitude=0.8)
    dr = float(np.linalg.norm(d))
    ref = sum(O.fz2df(np.ascontiguousarray(F[..., k]), dr, 1800.0, 30300.0, amplitude=0.8) * d[k] for k in range(3))
    assert got.shape == ref.shape
    assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    if d[0] == d[1] == 0.0:
        # an untilted step of length 0.1 pointing down is -0.1 times Fz2df of Fz
        ref_z = -0.1 * C.Fz2df(F[..., 2], 0.1, 1800.0, 30300.0, amplitude=0.8)
        assert np.abs(got - ref_z).max() <= 1e-12 * max(1.0, np.abs(ref_z).max())
