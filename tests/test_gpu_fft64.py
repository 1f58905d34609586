"""The exact path's second kernel: the FP64 FFT correlation rounded to integer
counts (calculate_s2_exact_fft, src/geometry/s2.rs:1219-1261), and the choice
between it and direct counting (plan_exact_cpu, src/geometry/s2.rs:733-766).
Both kernels must give the same integer counts and bit-identical curves, as
the reference's forced_kernels_are_selection_independent test asks
(src/geometry/s2.rs:1921-1935).  Requires the GPU (marker gpu)."""
import numpy as np
import pytest

from oracle import oracle as o
from rustmspt_amd import s2, synth
from rustmspt_amd._lib import S2HipError

pytestmark = pytest.mark.gpu


@pytest.fixture()
def method(ctx):
    yield ctx.set_exact_method
    ctx.set_exact_method("auto")


def both(ctx, method, occ, r_max, pitch=1.0):
    vol = s2.Volume.from_bytes(ctx, occ)
    out = {}
    for m in ("direct", "fft64"):
        method(m)
        out[m] = (vol.exact_counts(r_max, pitch), vol.exact_s2(r_max, pitch),
                  ctx.exact_tune_info(occ.shape, r_max, pitch if pitch >= 1e-9 else 1e-9)["kernel"])
    vol.free()
    return out


@pytest.mark.parametrize("shape,r_max,p,pitch", [
    ((9, 8, 1), 5, 0.4, 1.0), ((7, 6, 31), 6, 0.5, 1.0), ((6, 9, 32), 7, 0.3, 1.0), ((8, 5, 33), 6, 0.6, 1.0),
    ((4, 6, 127), 12, 0.4, 1.0), ((3, 4, 200), 20, 0.35, 1.0), ((24, 20, 18), 10, 0.3, 1.0),
    ((9, 2, 7), 5, 0.5, 0.5), ((20, 5, 9), 7, 0.4, 1.5), ((1, 1, 30), 12, 0.5, 0.7), ((2, 37, 3), 12, 0.5, 0.5),
    ((33, 31, 70), 24, 0.3, 1.0), ((40, 40, 40), 40, 0.25, 1.0), ((61, 45, 38), 19, 0.5, 1.0),
])
def test_fft64_equals_direct_and_oracle(ctx, method, shape, r_max, p, pitch):
    # lengths 2n-1 padded to 2,3,5-smooth sizes: radix 8/4/2 and 3, 5 stages all occur
    rng = np.random.default_rng(hash((shape, r_max, 64)) & 0xFFFF)
    occ = (rng.random(shape) < p).astype(np.uint8)
    out = both(ctx, method, occ, r_max, pitch)
    (cd, sd, kd), (cf, sf, kf) = out["direct"], out["fft64"]
    assert (kd, kf) == ("direct", "fft64")
    assert np.array_equal(cd, cf)
    assert sd.tobytes() == sf.tobytes()
    assert sf.tobytes() == o.exact_s2(occ, r_max, pitch).tobytes()


def test_fft64_solid_empty_single(ctx, method):
    for occ in (np.ones((17, 12, 40), np.uint8), np.zeros((17, 12, 40), np.uint8)):
        out = both(ctx, method, occ, 9)
        assert np.array_equal(out["direct"][0], out["fft64"][0])
        assert out["direct"][1].tobytes() == out["fft64"][1].tobytes()
    occ = np.zeros((15, 15, 15), np.uint8)
    occ[7, 7, 7] = 1
    out = both(ctx, method, occ, 6)
    assert np.array_equal(out["direct"][0], out["fft64"][0])


def test_fft64_256_cube_r64_equals_direct(ctx, method):
    # the benchmark's shape: 1,123,772 offsets, every count equal
    occ = synth.grf_volume(256, 6.0, 0.3, seed=12)
    out = both(ctx, method, occ, 64)
    assert len(out["fft64"][0]) == 1123772
    assert np.array_equal(out["direct"][0], out["fft64"][0])
    assert out["direct"][1].tobytes() == out["fft64"][1].tobytes()


def test_auto_picks_fft64_at_256_and_direct_at_small_radius(ctx, method):
    occ = synth.grf_volume(256, 6.0, 0.3, seed=13)
    vol = s2.Volume.from_bytes(ctx, occ)
    method("auto")
    c_auto = vol.exact_counts(64)
    assert ctx.exact_tune_info(occ.shape, 64, 1.0)["kernel"] == "fft64"
    method("direct")
    assert np.array_equal(vol.exact_counts(64), c_auto)
    # r_max = 2: a few hundred offsets; direct counting is modelled far cheaper
    method("auto")
    vol.exact_counts(2)
    assert ctx.exact_tune_info(occ.shape, 2, 1.0)["kernel"] == "direct"
    vol.free()


def test_fft64_refused_where_rounding_is_not_provably_exact(ctx, method):
    # a solid 620^3 volume: K^(3/2) = 1.5e13 puts the worst-case bound above
    # 1/4, so the forced FFT kernel is refused and auto falls back to direct
    n = 620
    row = np.zeros(s2.row_words(n), np.uint64)
    row[:n // 64] = np.uint64(0xFFFFFFFFFFFFFFFF)
    row[n // 64] = np.uint64((1 << (n % 64)) - 1)
    vol = s2.Volume.from_packed(ctx, np.tile(row, n * n), (n, n, n))
    method("fft64")
    with pytest.raises(S2HipError):
        vol.exact_counts(1)
    method("auto")
    c = vol.exact_counts(1)
    assert ctx.exact_tune_info((n, n, n), 1, 1.0)["kernel"] == "direct"
    offs, _ = s2.shell_order_offsets((n, n, n), 1, 1.0)
    assert np.array_equal(c, np.prod(n - np.abs(offs), axis=1))  # solid: every pair hits
    vol.free()


def test_budgeted_context_counts_directly(ctx, method):
    # the model's FFT choice keeps to the context's device budget: at 256^3
    # r40 the FP64 spectra (about 0.9 GB) do not fit 256 MiB, so a budgeted
    # context counts directly, as the reference's planner keeps to its budget
    # (src/geometry/s2.rs:733-766); without a budget the same shape takes the FFT
    occ = synth.grf_volume(256, 6.0, 0.3, seed=31)
    ref = o.exact_s2(occ, 40)
    method("auto")
    vol = s2.Volume.from_bytes(ctx, occ)
    assert np.array_equal(vol.exact_s2(40), ref)
    assert ctx.exact_tune_info(occ.shape, 40, 1.0)["kernel"] == "fft64"
    vol.free()
    with s2.Context(0, 256 << 20) as small:
        v2 = s2.Volume.from_bytes(small, occ)
        assert np.array_equal(v2.exact_s2(40), ref)
        assert small.exact_tune_info(occ.shape, 40, 1.0)["kernel"] == "direct"
        assert small.bytes_in_use <= 256 << 20
        v2.free()


# Every length the two-stage register kernels take (96 = 8 x 12 ... 512 = 16 x 32), on each
# axis in turn: axis 0 runs the x pass, axis 1 the y passes, axis 2 the z passes.  Untimed
# plans (autotune off) take the two-stage kernels wherever the length allows.
TWO_STAGE_CASES = [((60, 80, 103), 24, (96, 120, 128)), ((100, 140, 200), 32, (160, 192, 256)),
                   ((200, 316, 250), 64, (320, 384, 320)), ((280, 340, 360), 48, (360, 400, 512))]


@pytest.mark.parametrize("case", range(len(TWO_STAGE_CASES)))
@pytest.mark.parametrize("rot", range(3))
def test_two_stage_lengths_equal_direct(ctx, method, case, rot):
    shape, r_max, lengths = TWO_STAGE_CASES[case]
    shape = tuple(np.roll(shape, rot))
    assert tuple(s2.fft_padded_dims(shape, r_max)) == tuple(np.roll(lengths, rot))
    occ = synth.grf_volume(max(shape), 4.0, 0.3, seed=40 + case)[:shape[0], :shape[1], :shape[2]].copy()
    ctx.set_autotune(False)
    try:
        out = both(ctx, method, occ, r_max)
        vol = s2.Volume.from_bytes(ctx, occ)
        c32 = vol.fft32_counts(r_max, 1.0).astype(np.float64)
        vol.free()
    finally:
        ctx.set_autotune(True)
    cd, cf = out["direct"][0], out["fft64"][0]
    assert out["fft64"][2] == "fft64"
    assert np.array_equal(cd, cf)
    assert out["direct"][1].tobytes() == out["fft64"][1].tobytes()
    # FP32 on the same kernels: within the stated per-bin bound (tests/test_gpu_fft.py)
    tol = 2.5e-8 * float(cd.max()) * np.log2(float(np.prod(s2.fft_padded_dims(shape, r_max)))) + 0.05
    assert np.max(np.abs(c32 - cd.astype(np.float64))) <= tol
