"""Exact S2 on the GPU vs the oracle and brute force: integer counts and f64
curves bit-exact.  Requires the GPU (marker gpu)."""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as o
from rustmspt_amd import s2

pytestmark = pytest.mark.gpu
G = Path(__file__).resolve().parent / "golden"


@pytest.fixture(autouse=True, scope="module")
def direct_kernel(ctx):
    # these tests cover the direct counting kernel; tests/test_gpu_fft64.py
    # covers the FP64 FFT kernel and the choice between them
    ctx.set_exact_method("direct")
    yield
    ctx.set_exact_method("auto")


def hits_oracle(occ, r_max, pitch=1.0):
    offs, rs = o.domain_offsets(occ.shape, r_max, pitch)
    return offs, rs, o.hits_list(o.pack(occ), occ.shape, offs)


def test_golden_tiny_grids(ctx):
    # the reference's own tiny grids (src/geometry/s2.rs:1750, 1836; tests/gpu_s2_boundaries_tests.rs:49)
    for case in json.loads((G / "tiny_pair_counts.json").read_text()):
        dims = case["dims"]
        occ = np.array(case["occ"], np.uint8).reshape(dims)
        want = {(r[0], r[1], r[2]): r[3] for r in case["counts"]}
        r_max = int(np.ceil(np.sqrt(sum((n - 1) ** 2 for n in dims)))) + 1
        vol = s2.Volume.from_bytes(ctx, occ)
        got = vol.exact_counts(r_max, 1.0)
        offs, _ = s2.shell_order_offsets(dims, r_max, 1.0)
        assert len(offs) > 0
        for d, h in zip(offs.tolist(), got.tolist()):
            if tuple(d) in want:
                assert h == want[tuple(d)], (dims, d)
        assert np.array_equal(vol.exact_s2(r_max), o.exact_s2(occ, r_max)), dims
        vol.free()


@pytest.mark.parametrize("shape,r_max,p,pitch", [
    ((9, 8, 1), 5, 0.4, 1.0), ((7, 6, 31), 6, 0.5, 1.0), ((6, 9, 32), 7, 0.3, 1.0), ((8, 5, 33), 6, 0.6, 1.0),
    ((5, 7, 63), 9, 0.2, 1.0), ((6, 6, 64), 9, 0.5, 1.0), ((7, 5, 65), 8, 0.5, 1.0), ((4, 6, 127), 12, 0.4, 1.0),
    ((5, 4, 129), 12, 0.5, 1.0), ((3, 4, 200), 20, 0.35, 1.0), ((4, 3, 300), 40, 0.5, 1.0), ((24, 20, 18), 10, 0.3, 1.0),
    ((9, 2, 7), 5, 0.5, 0.5), ((20, 5, 9), 7, 0.4, 1.5), ((1, 1, 30), 12, 0.5, 0.7), ((16, 16, 16), 6, 0.45, 1.5),
    ((2, 37, 3), 12, 0.5, 0.5), ((33, 31, 70), 24, 0.3, 1.0), ((40, 40, 40), 40, 0.25, 1.0),
])
def test_random_volumes_counts_and_curve(ctx, shape, r_max, p, pitch):
    rng = np.random.default_rng(hash((shape, r_max)) & 0xFFFF)
    occ = (rng.random(shape) < p).astype(np.uint8)
    vol = s2.Volume.from_bytes(ctx, occ)
    offs, rs, want = hits_oracle(occ, r_max, pitch)
    got = vol.exact_counts(r_max, pitch)
    assert np.array_equal(got, want)
    assert np.array_equal(vol.exact_s2(r_max, pitch), o.exact_s2(occ, r_max, pitch))
    vol.free()


def test_packed_upload_roundtrip_and_popcount(ctx):
    rng = np.random.default_rng(2)
    occ = (rng.random((13, 11, 150)) < 0.3).astype(np.uint8)
    words = o.pack(occ)
    v1 = s2.Volume.from_packed(ctx, words, occ.shape)
    v2 = s2.Volume.from_bytes(ctx, occ)
    assert np.array_equal(v1.read_packed(), words) and np.array_equal(v2.read_packed(), words)
    assert v1.occupied == v2.occupied == int(occ.sum())
    assert np.array_equal(v2.to_bytes(), occ)


def test_solid_and_empty(ctx):
    solid = np.ones((32, 32, 32), np.uint8)
    v = s2.Volume.from_bytes(ctx, solid)
    assert np.all(v.exact_s2(16) == 1.0)
    e = s2.Volume.from_bytes(ctx, np.zeros((10, 9, 70), np.uint8))
    assert np.all(e.exact_s2(6) == 0.0) and np.all(e.exact_counts(6) == 0)


def test_single_voxel_known_answer(ctx):
    occ = np.zeros((4, 4, 4), np.uint8)
    occ[0, 0, 0] = 1
    v = s2.Volume.from_bytes(ctx, occ)
    assert v.exact_s2(2).tolist() == [1 / 64, 0.0, 0.0]


def test_128_cube_r32_full_parity(ctx):
    rng = np.random.default_rng(128)
    occ = (rng.random((128, 128, 128)) < 0.2).astype(np.uint8)
    vol = s2.Volume.from_bytes(ctx, occ)
    offs, rs, want = hits_oracle(occ, 32)
    assert np.array_equal(vol.exact_counts(32), want)
    assert np.array_equal(vol.exact_s2(32), o.curve_from_hits(occ.shape, 32, offs, rs, want, occ.sum() / occ.size))


def test_256_cube_r64_sampled_offsets_and_properties(ctx):
    from rustmspt_amd import synth
    occ = synth.grf_volume(256, 6.0, 0.3, seed=11)
    vol = s2.Volume.from_bytes(ctx, occ)
    got = vol.exact_counts(64)
    offs, rs = s2.shell_order_offsets(occ.shape, 64, 1.0)
    assert len(got) == 1123772
    rng = np.random.default_rng(0)
    pick = rng.choice(len(offs), 1500, replace=False)
    words = o.pack(occ)
    assert np.array_equal(got[pick], o.hits_list(words, occ.shape, offs[pick]))
    # C(d) = C(-d) across the list, and hits never exceed valid
    index = {tuple(d): i for i, d in enumerate(offs[:200000].tolist())}
    for i, d in enumerate(offs[:200000].tolist()):
        j = index.get((-d[0], -d[1], -d[2]))
        if j is not None:
            assert got[i] == got[j]
    valid = np.prod(np.array(occ.shape)[None, :] - np.abs(offs), axis=1)
    assert np.all(got <= valid)
    # curve equals the oracle's ordered fold over the device counts
    curve = vol.exact_s2(64)
    vf = occ.sum() / occ.size
    assert np.array_equal(curve, o.curve_from_hits(occ.shape, 64, offs, rs, got, vf))


@pytest.mark.parametrize("band,waves", [(1, 64), (2, 1000), (3, 16384), (6, 100000), (40, 7)])
def test_counts_independent_of_launch_shape(ctx, band, waves):
    rng = np.random.default_rng(band * 1000 + waves)
    occ = (rng.random((40, 37, 70)) < 0.3).astype(np.uint8)
    vol = s2.Volume.from_bytes(ctx, occ)
    offs, rs, want = hits_oracle(occ, 14)
    ctx.exact_tune_set(occ.shape, 14, 1.0, band, waves)
    assert np.array_equal(vol.exact_counts(14), want)
    info = ctx.exact_tune_info(occ.shape, 14, 1.0)
    assert (info["band"], info["waves"], info["picked_by"]) == (band, waves, "cost model")


def test_autotuner_times_candidates_on_first_call(ctx):
    from rustmspt_amd import synth
    occ = synth.grf_volume(96, 4.0, 0.3, seed=3)[:95, :93, :91].copy()  # a shape no other test uses
    vol = s2.Volume.from_bytes(ctx, occ)
    before = ctx.exact_tune_info(occ.shape, 23, 1.0)
    assert before["picked_by"] == "not run"
    got = vol.exact_counts(23)
    info = ctx.exact_tune_info(occ.shape, 23, 1.0)
    assert info["picked_by"] == "device timing" and info["timed_candidates"] >= 2
    # the timed shape is reused by a second context without timing again
    with s2.Context(0) as other:
        assert other.exact_tune_info(occ.shape, 23, 1.0)["picked_by"] == "device timing"
    offs, rs, want = hits_oracle(occ, 23)
    assert np.array_equal(got, want)
    assert np.array_equal(vol.exact_counts(23), want)  # second call reuses the tuned plan


def test_valu_probe_is_plausible(ctx):
    # 256 CUs x 64 lanes x 2.4 GHz = 39.3e12 lane-ops/s at one 64-lane VALU op per 4 cycles per SIMD
    rate = ctx.valu_probe()
    assert 5e12 < rate < 200e12, rate


def test_u32_lane_counters_cannot_wrap():
    # a solid 4096 x 2048 x 1024 volume at r_max = 1 with one x-chunk per
    # column block asked for (waves = 1): a single item would add
    # 4096 * 2047 * 1023 > 2^32 pairs into one lane's u32 counter, so the
    # plan splits x until every item's bound fits; counts are the analytic
    # valid-pair numbers of a solid grid
    dims = (4096, 2048, 1024)
    rw = s2.row_words(dims[2])
    row = np.zeros(rw, np.uint64)
    row[: dims[2] // 64] = np.uint64(0xFFFFFFFFFFFFFFFF)
    with s2.Context(0) as c:
        c.set_exact_method("direct")
        c.set_autotune(False)
        vol = s2.Volume.from_packed(c, np.tile(row, dims[0] * dims[1]), dims)
        c.exact_tune_set(dims, 1, 1.0, 1, 1)
        got = vol.exact_counts(1)
        offs, _ = s2.shell_order_offsets(dims, 1, 1.0)
        want = np.prod(np.array(dims, np.int64)[None, :] - np.abs(offs), axis=1)
        assert np.array_equal(got, want.astype(np.uint64)), (got, want)
        assert c.exact_tune_info(dims, 1, 1.0)["items"] >= 2
        vol.free()
