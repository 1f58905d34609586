"""Volume ingest on the device: classification and bit-packing of z-major sample
volumes, bit-exact against the oracle's restatement (decode, classify, pack in
the S2 layout), the boundary-mode background, RAW-folder and TIFF streaming,
and the reference's sample CT stack known answer.  Requires the GPU (marker
gpu)."""
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as o
from rustmspt_amd import s2
from rustmspt_amd import volume as V

from tiff_writer import LZW, write_tiff

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def packed(lv):
    words = lv.volume.read_packed()
    lv.volume.free()
    return words


@pytest.mark.parametrize("W", [70, 96])   # rows of 8-byte multiples take the wide-load kernel
@pytest.mark.parametrize("dt", ["u1", "i1", "<u2", ">u2", "<i2", ">i2", "<u4", ">u4", "<i4", ">i4"])
def test_pack_matches_oracle_all_sample_types(ctx, dt, W):
    rng = np.random.default_rng(sum(map(ord, dt)) + W)
    D, H = 130, 33   # z words 64 + 64 + 2
    info = np.iinfo(np.dtype(dt))
    vol = rng.integers(max(info.min, -1000), min(info.max, 1000), (D, H, W), endpoint=True).astype(dt)
    lo, hi = -37, 412
    lv = V.from_samples(ctx, vol, V.Phase.threshold(lo, hi))
    want, n = o.classify_pack(vol, "range", lo, hi)
    assert lv.volume.dims == [W, H, D] and lv.occupied == n
    assert np.array_equal(packed(lv), want)
    bg = int(vol[0, 0, 0])
    lv = V.from_samples(ctx, vol, V.Phase.foreground(bg))
    want, n = o.classify_pack(vol, "foreground", background=bg)
    assert lv.background == bg and lv.occupied == n and np.array_equal(packed(lv), want)


@pytest.mark.parametrize("dt", ["u1", "i1", "<u2", ">i2", ">u4", "<i4"])
def test_pack_rule_edges(ctx, dt):
    """Thresholds beyond the sample range, empty ranges, int64 extremes and
    backgrounds no sample can take: the wide kernel clamps 8/16-bit rules to
    32-bit comparisons, which must keep every outcome."""
    D, H, W = 70, 6, 48
    info = np.iinfo(np.dtype(dt))
    rng = np.random.default_rng(W + sum(map(ord, dt)))
    vol = rng.integers(info.min, info.max, (D, H, W), endpoint=True).astype(dt)
    vol[0, 0, :4] = [info.min, info.max, 0, -1 if info.min < 0 else 1]
    i64 = np.iinfo(np.int64)
    ranges = [(5, 3), (i64.min, i64.max), (i64.min, int(info.min) - 1), (int(info.max) + 1, i64.max),
              (int(info.min), int(info.min)), (int(info.max), int(info.max)), (-1, 0), (-200, -100),
              (70000, 80000), (int(info.min) - 5, int(info.min) + 3), (int(info.max) - 3, int(info.max) + 9)]
    for lo, hi in ranges:
        lv = V.from_samples(ctx, vol, V.Phase.threshold(lo, hi))
        want, n = o.classify_pack(vol, "range", lo, hi)
        assert lv.occupied == n and np.array_equal(packed(lv), want), (lo, hi)
    for bg in [int(info.min) - 1, int(info.max) + 1, int(info.min), int(info.max), 0, i64.min, i64.max]:
        lv = V.from_samples(ctx, vol, V.Phase.foreground(bg))
        want, n = o.classify_pack(vol, "foreground", background=bg)
        assert lv.occupied == n and np.array_equal(packed(lv), want), bg


@pytest.mark.parametrize("dt", ["u1", "<u2", "<i4"])
def test_pack_deep_stacks(ctx, dt):
    """2,200 slices: 36 row words, more than one LDS chunk of the wide kernel
    for 8-bit (15 words a chunk) and 16-bit (31) samples; rows and columns not
    multiples of the kernel's tile."""
    D, H, W = 2200, 7, 136
    rng = np.random.default_rng(D + H + W)
    vol = (rng.random((D, H, W)) < 0.4).astype(dt) * 7
    lv = V.from_samples(ctx, vol, V.Phase.foreground(0))
    want, n = o.classify_pack(vol, "foreground", background=0)
    assert lv.volume.dims == [W, H, D] and lv.occupied == n and np.array_equal(packed(lv), want)


@pytest.mark.parametrize("shape", [(1, 1, 1), (1, 5, 7), (2, 3, 4), (7, 1, 9), (6, 5, 1), (65, 9, 3), (9, 8, 200)])
@pytest.mark.parametrize("dt", ["u1", "i2", "u4"])
def test_background_mode_matches_oracle(ctx, shape, dt):
    rng = np.random.default_rng(sum(shape))
    vol = rng.integers(0, 4, shape).astype(dt)   # few values: many ties
    want_bg = o.boundary_mode(vol)
    lv = V.from_samples(ctx, vol, V.Phase.foreground())
    want, n = o.classify_pack(vol, "foreground", background=want_bg)
    assert lv.background == want_bg and lv.occupied == n and np.array_equal(packed(lv), want)


def test_background_ties_pick_smallest(ctx):
    vol = np.zeros((3, 3, 3), np.int16)
    vol[0], vol[1], vol[2] = -4, 9, 9   # 9 boundary values of -4, 17 of 9
    assert V.from_samples(ctx, vol, V.Phase.foreground()).background == 9
    vol = np.zeros((1, 1, 2), np.uint32)
    vol[0, 0] = [70000, 3]              # one each: the smallest
    assert V.from_samples(ctx, vol, V.Phase.foreground()).background == 3


def test_raw_folder_and_tiff_load_equal_samples(ctx, tmp_path):
    rng = np.random.default_rng(11)
    vol = rng.integers(0, 3000, (70, 40, 50)).astype("<u2")
    folder = tmp_path / "raw"
    folder.mkdir()
    for z in range(vol.shape[0]):
        (folder / f"slice_{z:04d}.raw").write_bytes(vol[z].tobytes())
    phase = V.Phase.threshold(1000, 2000)
    want, n = o.classify_pack(vol, "range", 1000, 2000)
    src = V.VolumeSource(str(folder), "raw", width=50, height=40, bits=16)
    lv = V.load(ctx, src, phase)
    assert (lv.info.width, lv.info.height, lv.info.depth) == (50, 40, 70) and lv.occupied == n
    assert np.array_equal(packed(lv), want)
    # a slice range
    lv = V.load(ctx, V.VolumeSource(str(folder), "raw", 5, 68, width=50, height=40, bits=16), phase)
    assert np.array_equal(packed(lv), o.classify_pack(vol[5:69], "range", 1000, 2000)[0])
    # the same volume big-endian in a compressed multi-page TIFF
    write_tiff(tmp_path / "v.tif", list(vol.astype(">u2")), order=">", compression=LZW, predictor=2)
    lv = V.load(ctx, V.VolumeSource(str(tmp_path / "v.tif"), "tiff"), phase)
    assert lv.info.big_endian and np.array_equal(packed(lv), want)


def test_streaming_through_several_staging_buffers(ctx, tmp_path):
    # 1024 x 1024 uint8 slices: 64 per 64 MiB staging buffer, 150 slices = three buffers
    rng = np.random.default_rng(3)
    folder = tmp_path / "big"
    folder.mkdir()
    vol = np.empty((150, 1024, 1024), np.uint8)
    for z in range(150):
        vol[z] = rng.integers(0, 2, (1024, 1024), dtype=np.uint8) * (z % 5 + 1)
        (folder / f"{z:05d}.raw").write_bytes(vol[z].tobytes())
    lv = V.load(ctx, V.VolumeSource(str(folder), "raw", width=1024, height=1024, bits=8), V.Phase.foreground(0))
    want, n = o.classify_pack(vol, "foreground", background=0)
    assert lv.occupied == n and np.array_equal(packed(lv), want)


def test_ct_stack_known_answer(ctx, tmp_path):
    # docs/en-us/examples/crop.md:111: background 0, 768151 foreground voxels
    z = np.load(GOLDEN / "ct_stack.npz")
    folder = tmp_path / "ct_stack"
    folder.mkdir()
    for name, sl in zip(z["names"], z["samples"]):
        (folder / str(name)).write_bytes(sl.astype("<u2").tobytes())
    cfg = {"type": "raw", "path": str(folder), "slice_start": -1, "slice_end": -1,
           "raw": {"width": 744, "height": 789, "bits": 16, "signed": False, "byte_order": "little"}}
    lv = V.load(ctx, V.VolumeSource.from_config(cfg), V.Phase.foreground())
    assert lv.background == 0 and lv.occupied == 768151
    assert lv.volume.dims == [744, 789, 3]
    want, _ = o.classify_pack(z["samples"], "foreground", background=0)
    assert np.array_equal(lv.volume.read_packed(), want)
    # the ingested grid is an ordinary resident volume: exact S2 equals the oracle's
    r = 3
    got = lv.volume.exact_s2(r)
    occ = np.transpose(z["samples"] != 0, (2, 1, 0))
    assert np.array_equal(got, o.exact_s2_packed(want, occ.shape, int(occ.sum()), r))
    lv.volume.free()


def test_ingest_errors(ctx, tmp_path):
    with pytest.raises(Exception, match="sample bytes"):
        import ctypes as C
        from rustmspt_amd import _lib as L
        info = L.VolumeInfo()
        info.width, info.height, info.depth, info.bits = 4, 4, 4, 8
        buf = np.zeros(10, np.uint8)
        h = C.c_void_p()
        ph = V.Phase.threshold(1)._c()
        L.check(L.lib().s2h_vol_from_samples(ctx.handle, buf.ctypes.data_as(C.c_void_p), buf.nbytes, C.byref(info),
                                             C.byref(ph), C.byref(h), None, None))
    with pytest.raises(Exception, match="RAW slice size mismatch"):
        f = tmp_path / "bad"
        f.mkdir()
        (f / "a.raw").write_bytes(b"\0" * 10)
        V.load(ctx, V.VolumeSource(str(f), "raw", width=4, height=4, bits=8), V.Phase.threshold(1))
