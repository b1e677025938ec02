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


def _corr(F, w, scale):
    """Brute-force valid correlation along the last axis in float64."""
    nzout = F.shape[-1] - len(w) + 1
    return scale * sum(F[..., k:k + nzout] * w[k] for k in range(len(w)))


@pytest.mark.parametrize("ncol,nz,nw", [(1000, 30, 9), (37, 300, 10), (5, 2500, 9), (3, 1, 1)])
def test_convolve_z_strided_records(ncol, nz, nw):
    """The batch's layout: Fz is the third float of 16-byte force records,
    column after column. Covers a ragged last block of columns, one column per
    block and single-weight windows."""
    from ppafm_amd import device as D
    print(f"case: convolve_z records, {ncol} columns x {nz}, {nw} weights", flush=True)
    rng = np.random.default_rng(nz + nw)
    rec = rng.normal(0, 0.1, (ncol, nz, 4)).astype(np.float32)
    w = rng.normal(0, 1.0, nw)
    dev = D.DeviceArray.from_host(rec)
    got = D.convolve_z_dev(dev, ncol, nz, w, scale=0.7, in_stride=nz * 4, in_step=4, in_offset=2).download()
    ref = _corr(rec[..., 2].astype(np.float64), w, 0.7)
    assert got.shape == ref.shape
    assert np.abs(got - ref).max() <= 1e-6 * max(1.0, np.abs(ref).max())


def test_convolve_z_long_columns_agree():
    """Columns of 4096 doubles are staged in LDS, longer ones are not: the
    outputs they share come from the same inputs and must be the same bits."""
    from ppafm_amd import device as D
    print("case: convolve_z, 4096 and 4097 doubles per column", flush=True)
    rng = np.random.default_rng(11)
    w = rng.normal(0, 1.0, 9)
    F = rng.normal(0, 0.1, (3, 4097))
    long = D.convolve_z_dev(D.DeviceArray.from_host(F), 3, 4097, w, scale=1.3, f64=True).download()
    short = D.convolve_z_dev(D.DeviceArray.from_host(np.ascontiguousarray(F[:, :4096])), 3, 4096, w, scale=1.3,
                             f64=True).download()
    assert np.array_equal(long[:, :-1], short)
    ref = _corr(F, w, 1.3)
    assert np.abs(long - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
