"""FP32 FFT autocorrelation on the GPU vs the FP64 FFT oracle and the exact
integer counts.  The FP32 path is not exact; the tolerances are stated here:

  per-offset count error vs FP64:   |c32 - c64| <= 2.5e-8 * M * log2(P) + 0.05
  per-bin S2 error vs exact:        |S2_32(r) - S2_exact(r)| <= 1e-5 (absolute) on
                                    the small random shapes, 1e-6 on the 128^3 field

with M the occupied voxel count (C(0), the largest correlation value) and P the
padded cell count (the same count bound as the full-size configurations').  The
FP64 values round to the exact counts, so the FP32 values are held to the same
bound against the exact integers.  Requires the GPU (marker gpu)."""
import math

import numpy as np
import pytest

from oracle import oracle as o
from rustmspt_amd import s2, synth

pytestmark = pytest.mark.gpu


def fp64_counts(occ, offs):
    pad = o.padded_fft_dims(occ.shape)
    corr = o.fft_autocorr(occ, pad)
    idx = tuple(np.mod(offs[:, a], pad[a]) for a in range(3))
    return corr[idx], pad


@pytest.mark.parametrize("shape,r_max,p,pitch", [
    ((8, 6, 5), 4, 0.4, 1.0), ((7, 11, 13), 9, 0.5, 1.0), ((1, 2, 9), 6, 0.4, 1.0), ((2, 37, 3), 12, 0.5, 0.5),
    ((1, 1, 40), 20, 0.5, 1.0), ((4, 1, 1), 3, 0.5, 1.0), ((24, 20, 18), 10, 0.3, 1.0), ((33, 31, 70), 24, 0.3, 1.0),
    ((48, 48, 48), 12, 0.25, 1.0), ((20, 5, 9), 7, 0.4, 1.5),
])
def test_fft32_counts_vs_fp64_and_exact(ctx, shape, r_max, p, pitch):
    rng = np.random.default_rng(sum(shape) + r_max)
    occ = (rng.random(shape) < p).astype(np.uint8)
    vol = s2.Volume.from_bytes(ctx, occ)
    offs, rs = s2.shell_order_offsets(shape, r_max, pitch)
    c32 = vol.fft32_counts(r_max, pitch).astype(np.float64)
    exact = o.hits_list(o.pack(occ), shape, offs).astype(np.float64)
    c64, pad = fp64_counts(occ, offs)
    M = float(occ.sum())
    tol = 2.5e-8 * M * math.log2(float(np.prod(pad))) + 0.05
    assert np.max(np.abs(c32 - c64)) <= tol
    assert np.array_equal(np.round(c64), exact)  # the FP64 reference itself rounds to the exact counts
    assert np.max(np.abs(c32 - exact)) <= tol + 1e-6
    # curve: same fold as exact, FP32 correlation values
    e = vol.exact_s2(r_max, pitch)
    f = vol.fft32_s2(r_max, pitch)
    assert f[0] == e[0]
    assert np.max(np.abs(f - e)) <= 1e-5


def test_fft32_grf_per_bin_error(ctx):
    occ = synth.grf_volume(128, 5.0, 0.3, seed=17)
    vol = s2.Volume.from_bytes(ctx, occ)
    e = vol.exact_s2(32)
    f = vol.fft32_s2(32)
    err = np.abs(f - e)
    assert err[0] == 0.0 and np.max(err) <= 1e-6, err
    counts = vol.fft32_counts(32).astype(np.float64)
    exact = vol.exact_counts(32).astype(np.float64)
    M = float(occ.sum())
    pad = s2.fft_padded_dims(occ.shape, 32)
    assert np.max(np.abs(counts - exact)) <= 2.5e-8 * M * math.log2(float(np.prod(pad))) + 0.05


def test_fft32_empty_and_solid(ctx):
    v = s2.Volume.from_bytes(ctx, np.zeros((12, 10, 9), np.uint8))
    assert np.all(v.fft32_s2(5) == 0.0)
    w = s2.Volume.from_bytes(ctx, np.ones((16, 16, 16), np.uint8))
    assert np.max(np.abs(w.fft32_s2(8) - 1.0)) <= 1e-5
