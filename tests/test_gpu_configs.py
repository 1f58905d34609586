"""The measured configurations at their full sizes, checked through properties
that do not need a full CPU recount: sampled offsets against the oracle, the
identity C(d) = C(-d), hits <= valid, the curve as the oracle's ordered fold of
the device counts, FP32 FFT against the FP64 FFT per radius bin, Monte Carlo
within 3 sigma.  Each case prints its per-bin figures (pytest -s).

  single large measure: 500 particles (icospheres + ellipsoids), 256^3, r_max=64
  CT-like volume:       768^3 Gaussian random field (correlation length 12,
                        volume fraction 0.35), r_max=128

FP32 tolerances (stated, as for the small cases): per-offset count error
<= 4e-7 * M * log2(P) + 1e-3 with M the occupied count and P the padded cell
count; per-bin |S2_fft32 - S2_exact| <= 1e-5.  Requires the GPU (marker gpu)."""
import math
import os

import numpy as np
import pytest
import scipy.fft as sfft

from oracle import oracle as o
from rustmspt_amd import s2, synth

pytestmark = pytest.mark.gpu

WORKERS = max(1, min(16, os.cpu_count() or 1))


def fp64_counts(occ, offs):
    """FP64 real-FFT autocorrelation on the reference's padded grid, read at
    the offsets (the FP64 FFT oracle)."""
    pad = [int(v) for v in o.padded_fft_dims(occ.shape)]
    f = sfft.rfftn(occ.astype(np.float64), s=pad, workers=WORKERS)
    f = (f.real * f.real + f.imag * f.imag).astype(np.complex128)
    corr = sfft.irfftn(f, s=pad, workers=WORKERS)
    idx = tuple(np.mod(offs[:, a], pad[a]) for a in range(3))
    return corr[idx], pad


def index_of(offs):
    """Dense lookup offset -> position in the list."""
    R = int(np.abs(offs).max())
    span = 2 * R + 1
    lut = np.full(span ** 3, -1, np.int64)
    key = ((offs[:, 0] + R) * span + (offs[:, 1] + R)) * span + (offs[:, 2] + R)
    lut[key] = np.arange(len(offs))
    neg = ((-offs[:, 0] + R) * span + (-offs[:, 1] + R)) * span + (-offs[:, 2] + R)
    return lut[neg]


def check_exact(vol, occ, r_max, nsample, seed):
    got = vol.exact_counts(r_max)
    offs, rs = s2.shell_order_offsets(occ.shape, r_max, 1.0)
    assert len(got) == len(offs)
    rng = np.random.default_rng(seed)
    pick = np.sort(rng.choice(len(offs), nsample, replace=False))
    words = o.pack(occ)
    assert np.array_equal(got[pick], o.hits_list(words, occ.shape, offs[pick]))
    twin = index_of(offs)
    assert np.all(twin >= 0) and np.array_equal(got, got[twin])
    valid = np.prod(np.array(occ.shape, np.int64)[None, :] - np.abs(offs), axis=1)
    assert np.all(got <= valid.astype(np.uint64))
    vf = occ.sum() / occ.size
    curve = vol.exact_s2(r_max)
    assert np.array_equal(curve, o.curve_from_hits(occ.shape, r_max, offs, rs, got, vf))
    return got, offs, rs, curve


def check_fft32(vol, occ, r_max, offs, rs, exact, curve, with_fp64):
    c32 = vol.fft32_counts(r_max).astype(np.float64)
    ex = exact.astype(np.float64)
    M = float(occ.sum())
    pad = [int(v) for v in o.padded_fft_dims(occ.shape)]
    tol = 4e-7 * M * math.log2(float(np.prod(pad))) + 1e-3
    f = vol.fft32_s2(r_max)
    d_s2 = np.abs(f - curve)
    if with_fp64:
        c64, _ = fp64_counts(occ, offs)
        assert np.array_equal(np.rint(c64), ex)  # the FP64 oracle rounds to the exact counts
    print(f"\n{occ.shape} r_max={r_max} M={M:.0f} tolerance={tol:.1f} counts")
    print(" r   max|c32-exact|" + ("   max|c32-c64|" if with_fp64 else "") + "   |S2_32-S2_exact|")
    for r in range(1, r_max + 1):
        sl = slice(int(rs[r]), int(rs[r + 1]))
        e_ex = float(np.max(np.abs(c32[sl] - ex[sl])))
        line = f"{r:3d} {e_ex:16.3f}"
        if with_fp64:
            e64 = float(np.max(np.abs(c32[sl] - c64[sl])))
            assert e64 <= tol, (r, e64, tol)
            line += f" {e64:14.3f}"
        assert e_ex <= tol + 0.5, (r, e_ex, tol)
        print(line + f" {d_s2[r]:18.3e}")
    assert f[0] == curve[0] and float(d_s2.max()) <= 1e-5, d_s2


def check_mc(vol, occ, r_max, offs, rs, exact, total, seed):
    per_r = max(200, total // r_max)
    hv = vol.mc_counts(r_max, 1.0, per_r, seed).astype(np.float64)
    valid = np.prod(np.array(occ.shape, np.float64)[None, :] - np.abs(offs), axis=1)
    z = []
    for r in range(1, r_max + 1):
        sl = slice(int(rs[r]), int(rs[r + 1]))
        p = float(exact[sl].astype(np.float64).sum() / valid[sl].sum())
        n = hv[r, 1]
        z.append((hv[r, 0] / n - p) / math.sqrt(p * (1 - p) / n))
    z = np.abs(np.array(z))
    print(f"MC {per_r} samples per radius: max |z| = {z.max():.2f}, beyond 2 sigma: {np.mean(z > 2):.3f}")
    # family-wise rule over r_max radii: all within 3.5 sigma, at most a quarter beyond 2 sigma
    assert np.all(z < 3.5), z
    assert np.mean(z > 2.0) <= 0.25, z


@pytest.fixture(scope="module")
def large(ctx):
    mesh, bbox, _ = synth.config_large(seed=2, n=500, grid=256)
    vol = s2.voxelize(ctx, mesh, bbox, 1.0)
    occ = vol.to_bytes()
    yield vol, occ
    vol.free()


def test_single_large_measure_exact(ctx, large):
    vol, occ = large
    check_exact(vol, occ, 64, 2000, seed=1)


def test_single_large_measure_fft32_per_bin(ctx, large):
    vol, occ = large
    exact = vol.exact_counts(64)
    offs, rs = s2.shell_order_offsets(occ.shape, 64, 1.0)
    check_fft32(vol, occ, 64, offs, rs, exact, vol.exact_s2(64), with_fp64=True)


def test_single_large_measure_monte_carlo(ctx, large):
    vol, occ = large
    exact = vol.exact_counts(64)
    offs, rs = s2.shell_order_offsets(occ.shape, 64, 1.0)
    check_mc(vol, occ, 64, offs, rs, exact, 10_000_000, seed=7)


def test_ct_like_volume_768(ctx):
    occ = synth.grf_volume(768, 12.0, 0.35, seed=5)
    assert abs(occ.mean() - 0.35) < 1e-4
    vol = s2.Volume.from_bytes(ctx, occ)
    exact, offs, rs, curve = check_exact(vol, occ, 128, 300, seed=2)
    assert len(offs) == 8887792
    check_fft32(vol, occ, 128, offs, rs, exact, curve, with_fp64=False)
    vol.free()
