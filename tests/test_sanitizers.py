"""The host-only C++ of the library under the sanitizers, on the CPU
(tests/sanitize): the STL parser and weld, granule split, clipped volume
fraction and the RAW/TIFF decoders under AddressSanitizer + UBSan over valid
and damaged inputs, with the valid ones checked against the production
library; the serving core (queue, admission, slot workers, pollers) with a
host-only backend under ThreadSanitizer and AddressSanitizer -- its stress
polls finished requests without waiting from several threads, the pattern in
which a request released by a poller was once still written by its worker."""
import ctypes as C
import shutil
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parent))
from tiff_writer import DEFLATE, LZW, NONE, PACKBITS, write_tiff  # noqa: E402

from rustmspt_amd import _lib as L  # noqa: E402
from rustmspt_amd import mesh as M  # noqa: E402
from rustmspt_amd import volume as V  # noqa: E402

HERE = Path(__file__).resolve().parent / "sanitize"
GOLDEN = Path(__file__).resolve().parent / "golden"
REPORTS = ("ERROR: AddressSanitizer", "runtime error:", "WARNING: ThreadSanitizer", "ERROR: LeakSanitizer")


@pytest.fixture(scope="module")
def built(tmp_path_factory):
    if not shutil.which("g++") or not Path("/opt/rocm/llvm/bin/clang++").exists():
        pytest.skip("g++ and ROCm's clang++ are needed for the sanitizer builds")
    out = tmp_path_factory.mktemp("sanitize")
    subprocess.run(["make", "-s", "-C", str(HERE), f"OUT={out}", "all"], check=True)
    return out


def run(exe, args, timeout=300):
    r = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, timeout=timeout)
    for rep in REPORTS:
        assert rep not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0, (r.returncode, r.stderr[-2000:])
    return r.stdout.splitlines()


def test_serving_core_under_tsan_and_asan(built):
    for exe in ("serve_tsan", "serve_asan"):
        assert run(built / exe, [])[-1] == "serve stress ok"


def stl_inputs(tmp):
    m = M.merge_meshes([M.icosphere_mesh(np.array([1.0, 2.0, 3.0]), 1.5, 2), M.box_mesh([4, 4, 4], [6, 7, 5])])
    M.save_stl(tmp / "bin.stl", m)
    lines = ["solid t"]
    for f in m.faces[:40]:
        lines += ["facet normal 0 0 0", "outer loop"]
        lines += [f"vertex {float(x)!r} {float(y)!r} {float(z)!r}" for x, y, z in m.vertices[f]]
        lines += ["endloop", "endfacet"]
    (tmp / "ascii.stl").write_text("\n".join(lines + ["endsolid t"]) + "\n")
    data = (tmp / "bin.stl").read_bytes()
    bad = {"truncated.stl": data[:len(data) // 2 + 7], "header_only.stl": data[:84], "short.stl": data[:40],
           "count_lies.stl": data[:80] + (10 ** 6).to_bytes(4, "little") + data[84:],
           "ascii_broken.stl": b"solid x\nfacet normal 0 0 0\nouter loop\nvertex 1 2\nvertex nan inf 3\n",
           "empty.stl": b""}
    for k, v in bad.items():
        (tmp / k).write_bytes(v)
    return ["bin.stl", "ascii.stl", str(GOLDEN / "particles.stl")], list(bad)


def test_host_units_under_asan_ubsan(built, tmp_path):
    good, bad = stl_inputs(tmp_path)
    args = []
    for f in good + bad:
        args += ["stl", tmp_path / f if not f.startswith("/") else f]
    rng = np.random.default_rng(3)
    # RAW folders: sample types and byte orders, plus a short slice
    raws = []
    for bits, signed, be in [(8, 0, 0), (16, 1, 1), (32, 0, 0), (16, 0, 0)]:
        d = tmp_path / f"raw{bits}{signed}{be}"
        d.mkdir()
        dt = np.dtype(f"{'>' if be else '<'}{'i' if signed else 'u'}{bits // 8}")
        for z in range(3):
            (d / f"{z:02d}.raw").write_bytes(rng.integers(0, 200, (5, 7)).astype(dt).tobytes())
        raws.append((d, 7, 5, bits, signed, be))
    short = tmp_path / "rawshort"
    short.mkdir()
    (short / "a.raw").write_bytes(bytes(35))
    (short / "b.raw").write_bytes(bytes(20))
    for d, w, h, bits, signed, be in raws + [(short, 7, 5, 8, 0, 0)]:
        args += ["raw", d, w, h, bits, signed, be]
    # TIFFs: encodings and layouts, then damaged copies
    tiffs = []
    pages = [rng.integers(0, 60000, (37, 29)).astype(np.uint16) for _ in range(3)]
    for k, (comp, pred, kw) in enumerate([(NONE, 1, {}), (LZW, 2, {"rows_per_strip": 5}), (DEFLATE, 1, {}),
                                          (PACKBITS, 1, {"tile": (16, 16)}), (LZW, 1, {"bigtiff": True})]):
        p = tmp_path / f"t{k}.tif"
        write_tiff(p, pages, order="<" if k % 2 else ">", compression=comp, predictor=pred, **kw)
        tiffs.append(p)
    damaged = []
    for k, p in enumerate(tiffs):
        data = p.read_bytes()
        for j, cut in enumerate([len(data) // 3, len(data) - 9, 12]):
            q = tmp_path / f"cut{k}_{j}.tif"
            q.write_bytes(data[:cut])
            damaged.append(q)
        q = tmp_path / f"flip{k}.tif"
        b = bytearray(data)
        for i in rng.integers(8, len(b), 40):
            b[i] ^= 0xFF
        q.write_bytes(bytes(b))
        damaged.append(q)
    for p in tiffs + damaged:
        args += ["tiff", p]
    out = run(built / "host_asan", args)
    assert len(out) == len(good) + len(bad) + len(raws) + 1 + len(tiffs) + len(damaged)
    # the valid inputs agree with the production library
    for f, line in zip(good, out):
        path = tmp_path / f if not f.startswith("/") else Path(f)
        m = M.load_stl(path)
        g, ng = M.granules(m)
        assert line.split()[:4] == ["stl", str(len(m.vertices)), str(len(m.faces)), str(ng)], line
    assert all(line.startswith("error") for line in out[len(good):len(good) + 3])
    k0 = len(good) + len(bad)
    for (d, w, h, bits, signed, be), line in zip(raws, out[k0:k0 + len(raws)]):
        want = V.read(V.VolumeSource(str(d), "raw", width=w, height=h, bits=bits, signed=bool(signed),
                                     byte_order="big" if be else "little"))
        assert line.split()[1:] == [str(w), str(h), "3", str(bits), str(fnv64(want.tobytes()))], line
    assert out[k0 + len(raws)].startswith("error")
    t0 = k0 + len(raws) + 1
    for p, line in zip(tiffs, out[t0:t0 + len(tiffs)]):
        want = V.read(V.VolumeSource(str(p), "tiff"))
        assert line.split()[1:] == ["29", "37", "3", "16", str(fnv64(want.tobytes()))], line


def fnv64(b: bytes) -> int:
    h = 1469598103934665603
    for x in b:
        h = ((h ^ x) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h
