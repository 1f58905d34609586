"""Volume file readers (host side of the ingest, no device): RAW slice folders
and TIFF stacks decoded by libs2hip.so against the reference's own known
answers (tests/io_tests.rs), an independent TIFF encoder (tests/tiff_writer.py)
and the reference's sample CT stack (tests/golden/ct_stack.npz)."""
import os
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as o
from rustmspt_amd import volume as V
from rustmspt_amd._lib import S2HipError

from tiff_writer import DEFLATE, LZW, NONE, PACKBITS, write_tiff

GOLDEN = Path(__file__).resolve().parent / "golden"


def write_raw_folder(folder: Path, slices, names=None):
    folder.mkdir(parents=True, exist_ok=True)
    for i, s in enumerate(slices):
        name = names[i] if names else f"{i:03d}.raw"
        (folder / name).write_bytes(bytes(s) if isinstance(s, (bytes, list)) else np.ascontiguousarray(s).tobytes())


def test_raw_folder_loads_with_inclusive_range(tmp_path):
    # tests/io_tests.rs:53-79
    f = tmp_path / "raw_slices"
    write_raw_folder(f, [[1, 2, 3, 4], [5, 6, 7, 8], [9, 10, 11, 12]])
    src = V.VolumeSource(str(f), "raw", 1, 2, width=2, height=2, bits=8)
    info = V.probe(src)
    assert (info.width, info.height, info.depth, info.bits, info.signed) == (2, 2, 2, 8, False)
    assert V.read(src).reshape(-1).tolist() == [5, 6, 7, 8, 9, 10, 11, 12]


def test_raw_folder_supports_big_endian(tmp_path):
    # tests/io_tests.rs:82-104
    f = tmp_path / "raw_be"
    write_raw_folder(f, [[0x01, 0x02, 0x03, 0x04]])
    src = V.VolumeSource(str(f), "raw", -1, -1, width=2, height=1, bits=16, byte_order="big")
    got = V.read(src)
    assert got.shape == (1, 1, 2) and got.astype(np.int64).reshape(-1).tolist() == [0x0102, 0x0304]
    assert got.astype(np.int64).reshape(-1).tolist() == o.decode_raw_slice(bytes([1, 2, 3, 4]), 16, False, True).tolist()


@pytest.mark.parametrize("bits,signed,order", [(8, False, "little"), (8, True, "big"), (16, False, "little"),
                                                (16, True, "big"), (32, False, "big"), (32, True, "little")])
def test_raw_sample_types_match_decode(tmp_path, bits, signed, order):
    rng = np.random.default_rng(bits + signed)
    W, H, D = 13, 7, 5
    raw = [rng.integers(0, 256, W * H * bits // 8, dtype=np.uint8).tobytes() for _ in range(D)]
    write_raw_folder(tmp_path / "v", raw, names=[f"s{i:02d}.bin" for i in range(D)])
    src = V.VolumeSource(str(tmp_path / "v"), "raw", width=W, height=H, bits=bits, signed=signed, byte_order=order)
    got = V.read(src).astype(np.int64).reshape(D, -1)
    for z in range(D):
        assert np.array_equal(got[z], o.decode_raw_slice(raw[z], bits, signed, order == "big"))


def test_raw_slices_sorted_by_name_and_errors(tmp_path):
    f = tmp_path / "unsorted"
    write_raw_folder(f, [[3] * 4, [1] * 4, [2] * 4], names=["c.raw", "a.raw", "b.raw"])
    (f / "sub").mkdir()  # folders are skipped
    got = V.read(V.VolumeSource(str(f), "raw", width=2, height=2, bits=8))
    assert got[:, 0, 0].tolist() == [1, 2, 3]
    with pytest.raises(S2HipError, match="Invalid slice range"):
        V.read(V.VolumeSource(str(f), "raw", 2, 1, width=2, height=2, bits=8))
    with pytest.raises(S2HipError, match="Invalid slice range"):
        V.read(V.VolumeSource(str(f), "raw", 0, 3, width=2, height=2, bits=8))
    with pytest.raises(S2HipError, match="RAW slice size mismatch"):
        V.read(V.VolumeSource(str(f), "raw", width=3, height=2, bits=8))
    with pytest.raises(S2HipError, match="RAW bits must be 8/16/32"):
        V.probe(V.VolumeSource(str(f), "raw", width=2, height=2, bits=12))
    with pytest.raises(S2HipError, match="width/height must be > 0"):
        V.probe(V.VolumeSource(str(f), "raw", width=0, height=2, bits=8))
    (tmp_path / "empty").mkdir()
    with pytest.raises(S2HipError, match="no files"):
        V.probe(V.VolumeSource(str(tmp_path / "empty"), "raw", width=2, height=2, bits=8))
    with pytest.raises(ValueError, match="Invalid byte_order"):
        V.read(V.VolumeSource(str(f), "raw", width=2, height=2, bits=8, byte_order="middle"))


def test_volume_source_from_crop_config():
    # the crop input block (src/config/crop.rs:4-22, data/input/crop_config.yaml)
    cfg = {"type": "raw", "path": "data/input/ct_stack/", "slice_start": -1, "slice_end": -1,
           "raw": {"width": 744, "height": 789, "bits": 16, "signed": False, "byte_order": "little"}}
    s = V.VolumeSource.from_config(cfg)
    assert (s.kind, s.width, s.height, s.bits, s.signed, s.byte_order) == ("raw", 744, 789, 16, False, "little")
    with pytest.raises(ValueError, match="input.raw is required"):
        V.VolumeSource.from_config({"type": "raw", "path": "x"})
    with pytest.raises(ValueError, match="Expected raw or tiff"):
        V.VolumeSource.from_config({"type": "png", "path": "x"})
    assert V.VolumeSource.from_config({"type": "TIF", "path": "x"}).kind == "tif"


def test_tiff_file_and_folder_roundtrip(tmp_path):
    # tests/io_tests.rs:107-147 with this repository's TIFF writer
    vol = np.array([1, 2, 3, 4, 11, 12, 13, 14, 21, 22, 23, 24], np.uint16).reshape(3, 2, 2)
    write_tiff(tmp_path / "stack.tiff", list(vol))
    got = V.read(V.VolumeSource(str(tmp_path / "stack.tiff"), "tiff"))
    info = V.probe(V.VolumeSource(str(tmp_path / "stack.tiff"), "tiff"))
    assert (info.width, info.height, info.depth, info.bits, info.signed) == (2, 2, 3, 16, False)
    assert np.array_equal(got, vol)
    for ext in ("tiff", "tif"):
        folder = tmp_path / f"slices_{ext}"
        folder.mkdir()
        for z in range(3):
            write_tiff(folder / f"ct_{z:04d}.{ext}", [vol[z]])
        (folder / "notes.txt").write_text("not a slice")
        assert np.array_equal(V.read(V.VolumeSource(str(folder), "tiff")), vol)
        assert np.array_equal(V.read(V.VolumeSource(str(folder), "tiff", 1, 2)), vol[1:])


@pytest.mark.parametrize("order", ["<", ">"])
@pytest.mark.parametrize("compression", [NONE, LZW, DEFLATE, PACKBITS])
@pytest.mark.parametrize("predictor", [1, 2])
@pytest.mark.parametrize("layout", ["strips", "tiles", "bigtiff"])
def test_tiff_encodings(tmp_path, order, compression, predictor, layout):
    rng = np.random.default_rng(compression + predictor)
    pages = []
    for dt in ("u2",):
        for _ in range(4):
            pages.append((rng.integers(0, 60000, (29, 41)) // 97).astype(dt))
    kw = dict(order=order, compression=compression, predictor=predictor)
    if layout == "strips":
        kw["rows_per_strip"] = 6
    elif layout == "tiles":
        kw["tile"] = (16, 32)
    else:
        kw["bigtiff"] = True
    p = tmp_path / "v.tif"
    write_tiff(p, pages, **kw)
    got = V.read(V.VolumeSource(str(p), "tiff", 1, 3))
    assert np.array_equal(got.astype(np.int64), np.stack(pages[1:4]).astype(np.int64))
    assert V.probe(V.VolumeSource(str(p), "tiff")).big_endian == (order == ">")


@pytest.mark.parametrize("dt", ["u1", "i1", "u2", "i2", "u4", "i4"])
def test_tiff_sample_types(tmp_path, dt):
    rng = np.random.default_rng(len(dt))
    info = np.iinfo(dt)
    pages = [rng.integers(info.min, info.max, (9, 11), endpoint=True).astype(dt) for _ in range(2)]
    write_tiff(tmp_path / "t.tif", pages, compression=LZW, predictor=2)
    got = V.read(V.VolumeSource(str(tmp_path / "t.tif"), "tiff"))
    assert got.dtype.kind == np.dtype(dt).kind and np.array_equal(got.astype(np.int64), np.stack(pages).astype(np.int64))


def test_tiff_lzw_long_strip(tmp_path):
    # one 64 KiB strip of mixed noise and runs: codes widen to 12 bits and the table is cleared
    rng = np.random.default_rng(5)
    page = rng.integers(0, 256, (256, 256)).astype(np.uint8)
    page[::3] = 7
    write_tiff(tmp_path / "l.tif", [page, page.T.copy()], compression=LZW)
    assert np.array_equal(V.read(V.VolumeSource(str(tmp_path / "l.tif"), "tiff")), np.stack([page, page.T]))


def test_tiff_mismatches_are_errors(tmp_path):
    folder = tmp_path / "mix"
    folder.mkdir()
    write_tiff(folder / "a.tif", [np.zeros((4, 5), np.uint16)])
    write_tiff(folder / "b.tif", [np.zeros((4, 6), np.uint16)])
    with pytest.raises(S2HipError, match="TIFF shape mismatch"):
        V.probe(V.VolumeSource(str(folder), "tiff"))
    write_tiff(folder / "b.tif", [np.zeros((4, 5), np.int16)])
    with pytest.raises(S2HipError, match="TIFF numeric type mismatch"):
        V.probe(V.VolumeSource(str(folder), "tiff"))
    (tmp_path / "none").mkdir()
    with pytest.raises(S2HipError, match="No TIFF files found"):
        V.probe(V.VolumeSource(str(tmp_path / "none"), "tiff"))
    with pytest.raises(S2HipError, match="neither file nor folder"):
        V.probe(V.VolumeSource(str(tmp_path / "missing"), "tiff"))
    (tmp_path / "bad.tif").write_bytes(b"not a tiff at all")
    with pytest.raises(S2HipError, match="not a TIFF file"):
        V.probe(V.VolumeSource(str(tmp_path / "bad.tif"), "tiff"))


def test_ct_stack_fixture_known_answer(tmp_path):
    # the reference's sample CT stack: three 744 x 789 uint16 LE slices; its docs state
    # background 0 and 768151 foreground voxels (docs/en-us/examples/crop.md:111)
    z = np.load(GOLDEN / "ct_stack.npz")
    vol = z["samples"]
    write_raw_folder(tmp_path / "ct", list(vol), names=[str(n) for n in z["names"]])
    src = V.VolumeSource(str(tmp_path / "ct"), "raw", width=744, height=789, bits=16)
    got = V.read(src)
    assert got.shape == (3, 789, 744) and np.array_equal(got, vol)
    assert o.boundary_mode(got) == 0
    assert int(np.count_nonzero(got != 0)) == 768151


def test_oracle_boundary_enumeration():
    # each boundary voxel exactly once, for degenerate extents too
    for D, H, W in [(1, 1, 1), (1, 3, 4), (2, 3, 4), (5, 1, 4), (5, 2, 1), (5, 4, 1), (6, 5, 7)]:
        vol = np.arange(D * H * W).reshape(D, H, W)
        vals = o.boundary_values(vol)
        z, y, x = np.unravel_index(vals, vol.shape)
        on = (z == 0) | (z == D - 1) | (y == 0) | (y == H - 1) | (x == 0) | (x == W - 1)
        expect = np.count_nonzero((np.indices(vol.shape)[0] == 0) | (np.indices(vol.shape)[0] == D - 1) |
                                  (np.indices(vol.shape)[1] == 0) | (np.indices(vol.shape)[1] == H - 1) |
                                  (np.indices(vol.shape)[2] == 0) | (np.indices(vol.shape)[2] == W - 1))
        assert on.all() and len(np.unique(vals)) == len(vals) == expect
    tie = np.zeros((3, 3, 3), np.int64)
    tie[0] = 5
    tie[2] = 5
    tie[1] = 2  # 18 fives on the z faces, 8 twos on the middle ring: mode 5
    assert o.boundary_mode(tie) == 5
    tie[1] = 2
    tie[2] = 2  # 9 fives, 17 twos
    assert o.boundary_mode(tie) == 2
    even = np.zeros((1, 1, 2), np.int64)
    even[0, 0] = [9, 4]
    assert o.boundary_mode(even) == 4  # equal counts: the smallest value
