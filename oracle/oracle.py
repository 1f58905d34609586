"""ctypes bindings of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.  Each function
restates the reference function cited in s2_oracle.c.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_u64p = C.POINTER(C.c_uint64)
_i64p = C.POINTER(C.c_int64)
_u32p = C.POINTER(C.c_uint32)
_u8p = C.POINTER(C.c_uint8)
_f64p = C.POINTER(C.c_double)


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE), "liboracle.so"], check=True)
    return LIB_PATH


def _load():
    if not LIB_PATH.exists():
        build()
    lib = C.CDLL(str(LIB_PATH))
    sig = {
        "orc_ray_triangle": (C.c_int, [_f64p, _f64p, _f64p, _f64p, _f64p, _f64p]),
        "orc_occupancy_dims": (None, [_f64p, _f64p, C.c_double, _u64p]),
        "orc_granules": (C.c_uint32, [_u32p, C.c_uint64, C.c_uint64, _u32p]),
        "orc_voxelize": (C.c_uint64, [_f64p, C.c_uint64, _u32p, C.c_uint64, _u32p, C.c_uint32, _f64p,
                                      C.c_double, _u64p, _u8p]),
        "orc_shell_offsets": (C.c_uint64, [C.c_double, C.c_double, _i64p, C.c_uint64]),
        "orc_domain_offsets": (C.c_uint64, [_u64p, C.c_uint32, C.c_double, _i64p, C.c_uint64, _u64p]),
        "orc_row_words": (C.c_uint64, [C.c_uint64]),
        "orc_pack": (None, [_u8p, _u64p, _u64p]),
        "orc_pair_counts_bits": (None, [_u64p, _u64p, _i64p, _u64p, _u64p]),
        "orc_pair_counts_bool": (None, [_u8p, _u64p, _i64p, _u64p, _u64p]),
        "orc_hits_list": (None, [_u64p, _u64p, _i64p, C.c_uint64, _u64p]),
        "orc_fill_missing": (None, [_f64p, _u8p, C.c_uint64, C.c_double]),
        "orc_curve_from_hits": (None, [_u64p, C.c_uint32, _i64p, _u64p, _u64p, C.c_double, _f64p]),
        "orc_exact_s2": (None, [_u64p, _u64p, C.c_uint32, C.c_double, C.c_uint64, _f64p]),
        "orc_smooth_fft_length": (C.c_uint64, [C.c_uint64]),
        "orc_padded_fft_dims": (None, [_u64p, _u64p]),
        "orc_exact_shell_work": (None, [_u64p, C.c_uint32, C.c_double, _u64p, _f64p, _u64p]),
        "orc_plan_exact_cpu": (C.c_int, [_u64p, C.c_uint32, C.c_double, C.c_uint64, C.c_double, _f64p, _u64p]),
        "orc_fft_autocorr": (None, [_u8p, _u64p, _u64p, _f64p]),
        "orc_fft_hits": (None, [_f64p, _u64p, _i64p, C.c_uint64, _u64p]),
        "orc_philox4x32_10": (None, [_u32p, _u32p, _u32p]),
        "orc_mc_counts": (None, [_u64p, _u64p, C.c_uint32, C.c_double, C.c_uint64, C.c_uint64, _u64p]),
        "orc_mesh_mc_counts": (None, [_f64p, C.c_uint64, _u32p, C.c_uint64, _f64p, _f64p, C.c_uint32, C.c_uint64,
                                      C.c_uint64, _u64p]),
        "orc_num_threads": (C.c_int, []),
        "orc_set_threads": (None, [C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def _dims(dims):
    return np.ascontiguousarray(np.asarray(dims, dtype=np.uint64))


# ------------------------------------------------------------------ geometry

def ray_triangle(o, d, a, b, c):
    t = C.c_double(0.0)
    arrs = [np.ascontiguousarray(np.asarray(v, dtype=np.float64)) for v in (o, d, a, b, c)]
    hit = lib().orc_ray_triangle(*[_p(x, _f64p) for x in arrs], C.byref(t))
    return t.value if hit else None


def occupancy_dims(bmin, bmax, pitch):
    out = np.zeros(3, np.uint64)
    lib().orc_occupancy_dims(_p(np.asarray(bmin, np.float64), _f64p), _p(np.asarray(bmax, np.float64), _f64p),
                             float(pitch), _p(out, _u64p))
    return [int(v) for v in out]


def granules(tri, nverts):
    tri = np.ascontiguousarray(tri, dtype=np.uint32).reshape(-1)
    comp = np.zeros(max(len(tri) // 3, 1), np.uint32)
    n = lib().orc_granules(_p(tri, _u32p), len(tri) // 3, int(nverts), _p(comp, _u32p))
    return comp[: len(tri) // 3], int(n)


def voxelize(verts, tri, bmin, pitch, dims):
    """build_bbox_occupancy: returns the bool grid (uint8, shape dims)."""
    verts = np.ascontiguousarray(verts, dtype=np.float64).reshape(-1)
    tri = np.ascontiguousarray(tri, dtype=np.uint32).reshape(-1)
    comp, n = granules(tri, len(verts) // 3)
    d = _dims(dims)
    occ = np.zeros(int(d[0] * d[1] * d[2]), np.uint8)
    lib().orc_voxelize(_p(verts, _f64p), len(verts) // 3, _p(tri, _u32p), len(tri) // 3,
                       _p(np.ascontiguousarray(comp), _u32p), n, _p(np.asarray(bmin, np.float64), _f64p),
                       float(pitch), _p(d, _u64p), _p(occ, _u8p))
    return occ.reshape([int(v) for v in dims])


# ------------------------------------------------------------------ shells and counts

def shell_offsets(distance_vox, half_width_vox):
    n = lib().orc_shell_offsets(float(distance_vox), float(half_width_vox), None, 0)
    out = np.zeros((max(n, 1), 3), np.int64)
    lib().orc_shell_offsets(float(distance_vox), float(half_width_vox), _p(out, _i64p), n)
    return out[:n]


def domain_offsets(dims, r_max, pitch):
    d = _dims(dims)
    n = lib().orc_domain_offsets(_p(d, _u64p), int(r_max), float(pitch), None, 0, None)
    out = np.zeros((max(n, 1), 3), np.int64)
    rs = np.zeros(int(r_max) + 2, np.uint64)
    lib().orc_domain_offsets(_p(d, _u64p), int(r_max), float(pitch), _p(out, _i64p), n, _p(rs, _u64p))
    return out[:n], rs


def row_words(nz):
    return int(lib().orc_row_words(int(nz)))


def pack(occ):
    occ = np.ascontiguousarray(occ, dtype=np.uint8)
    dims = _dims(occ.shape)
    words = np.zeros(int(dims[0] * dims[1]) * row_words(int(dims[2])), np.uint64)
    lib().orc_pack(_p(occ.reshape(-1), _u8p), _p(dims, _u64p), _p(words, _u64p))
    return words


def pair_counts_bits(words, dims, off):
    h, v = C.c_uint64(), C.c_uint64()
    o = np.asarray(off, np.int64)
    lib().orc_pair_counts_bits(_p(words, _u64p), _p(_dims(dims), _u64p), _p(o, _i64p), C.byref(h), C.byref(v))
    return h.value, v.value


def pair_counts_bool(occ, off):
    occ = np.ascontiguousarray(occ, dtype=np.uint8)
    h, v = C.c_uint64(), C.c_uint64()
    o = np.asarray(off, np.int64)
    lib().orc_pair_counts_bool(_p(occ.reshape(-1), _u8p), _p(_dims(occ.shape), _u64p), _p(o, _i64p),
                               C.byref(h), C.byref(v))
    return h.value, v.value


def hits_list(words, dims, offs):
    offs = np.ascontiguousarray(offs, np.int64)
    out = np.zeros(max(len(offs), 1), np.uint64)
    lib().orc_hits_list(_p(words, _u64p), _p(_dims(dims), _u64p), _p(offs, _i64p), len(offs), _p(out, _u64p))
    return out[: len(offs)]


def fill_missing(values, has_support, vf):
    v = np.ascontiguousarray(values, np.float64).copy()
    s = np.ascontiguousarray(has_support, np.uint8)
    lib().orc_fill_missing(_p(v, _f64p), _p(s, _u8p), len(v), float(vf))
    return v


def curve_from_hits(dims, r_max, offs, rs, hits, vf):
    out = np.zeros(int(r_max) + 1, np.float64)
    lib().orc_curve_from_hits(_p(_dims(dims), _u64p), int(r_max), _p(np.ascontiguousarray(offs, np.int64), _i64p),
                              _p(np.ascontiguousarray(rs, np.uint64), _u64p),
                              _p(np.ascontiguousarray(hits, np.uint64), _u64p), float(vf), _p(out, _f64p))
    return out


def exact_s2(occ, r_max, pitch=1.0):
    """VoxelS2::calculate_exact_with(Direct) on a bool grid."""
    occ = np.ascontiguousarray(occ, dtype=np.uint8)
    words = pack(occ)
    out = np.zeros(int(r_max) + 1, np.float64)
    lib().orc_exact_s2(_p(words, _u64p), _p(_dims(occ.shape), _u64p), int(r_max), float(pitch),
                       int(occ.sum()), _p(out, _f64p))
    return out


def exact_s2_packed(words, dims, occupied, r_max, pitch=1.0):
    out = np.zeros(int(r_max) + 1, np.float64)
    lib().orc_exact_s2(_p(words, _u64p), _p(_dims(dims), _u64p), int(r_max), float(pitch), int(occupied),
                       _p(out, _f64p))
    return out


# ------------------------------------------------------------------ planner / FFT

def smooth_fft_length(n):
    return int(lib().orc_smooth_fft_length(int(n)))


def padded_fft_dims(dims):
    out = np.zeros(3, np.uint64)
    lib().orc_padded_fft_dims(_p(_dims(dims), _u64p), _p(out, _u64p))
    return [int(v) for v in out]


def exact_shell_work(dims, r_max, pitch):
    k, w, m = C.c_uint64(), C.c_double(), C.c_uint64()
    lib().orc_exact_shell_work(_p(_dims(dims), _u64p), int(r_max), float(pitch), C.byref(k), C.byref(w), C.byref(m))
    return k.value, w.value, m.value


def plan_exact_cpu(dims, r_max, pitch, workers, budget=768 * 1024 * 1024):
    out = np.zeros(5, np.float64)
    k = C.c_uint64()
    m = lib().orc_plan_exact_cpu(_p(_dims(dims), _u64p), int(r_max), float(pitch), int(workers), float(budget),
                                 _p(out, _f64p), C.byref(k))
    return {"method": "fft" if m == 0 else "direct", "fft_seconds": out[0], "direct_seconds": out[1],
            "fft_bytes": out[2], "direct_bytes": out[3], "pair_work": out[4], "offsets": k.value}


def fft_autocorr(occ, padded=None):
    occ = np.ascontiguousarray(occ, dtype=np.uint8)
    padded = padded or padded_fft_dims(occ.shape)
    p = _dims(padded)
    corr = np.zeros(int(p[0] * p[1] * p[2]), np.float64)
    lib().orc_fft_autocorr(_p(occ.reshape(-1), _u8p), _p(_dims(occ.shape), _u64p), _p(p, _u64p), _p(corr, _f64p))
    return corr.reshape([int(v) for v in padded])


def fft_hits(corr, offs):
    offs = np.ascontiguousarray(offs, np.int64)
    out = np.zeros(max(len(offs), 1), np.uint64)
    lib().orc_fft_hits(_p(np.ascontiguousarray(corr, np.float64).reshape(-1), _f64p),
                       _p(_dims(corr.shape), _u64p), _p(offs, _i64p), len(offs), _p(out, _u64p))
    return out[: len(offs)]


# ------------------------------------------------------------------ Philox / MC

def philox4x32_10(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    o = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(_p(c, _u32p), _p(k, _u32p), _p(o, _u32p))
    return [int(v) for v in o]


def mc_counts(words, dims, r_max, pitch, samples, seed):
    hv = np.zeros(2 * (int(r_max) + 1), np.uint64)
    lib().orc_mc_counts(_p(words, _u64p), _p(_dims(dims), _u64p), int(r_max), float(pitch), int(samples),
                        int(seed) & 0xFFFFFFFFFFFFFFFF, _p(hv, _u64p))
    return hv.reshape(-1, 2)


def mesh_mc_counts(verts, tri, bmin, bmax, r_max, samples, seed):
    """Mesh Monte Carlo (hits, valid) per radius, rows 0..r_max (row 0 zero)."""
    v = np.ascontiguousarray(verts, np.float64).reshape(-1)
    t = np.ascontiguousarray(tri, np.uint32).reshape(-1)
    out = np.zeros(2 * (int(r_max) + 1), np.uint64)
    lib().orc_mesh_mc_counts(_p(v, _f64p), len(v) // 3, _p(t, _u32p), len(t) // 3,
                             _p(np.ascontiguousarray(bmin, np.float64), _f64p),
                             _p(np.ascontiguousarray(bmax, np.float64), _f64p), int(r_max), int(samples),
                             int(seed) & 0xFFFFFFFFFFFFFFFF, _p(out, _u64p))
    return out.reshape(-1, 2)


def set_threads(n):
    lib().orc_set_threads(int(n))


def num_threads():
    return int(lib().orc_num_threads())


# ------------------------------------------------------------------ brute force (tiny grids)

def brute_pair_counts(occ, off):
    """Exhaustive (hits, valid) for one offset on a tiny bool grid."""
    occ = np.asarray(occ, dtype=bool)
    n = occ.shape
    d = [int(v) for v in off]
    sl_a, sl_b = [], []
    for a in range(3):
        if abs(d[a]) >= n[a]:
            return 0, 0
        if d[a] >= 0:
            sl_a.append(slice(0, n[a] - d[a]))
            sl_b.append(slice(d[a], n[a]))
        else:
            sl_a.append(slice(-d[a], n[a]))
            sl_b.append(slice(0, n[a] + d[a]))
    A = occ[tuple(sl_a)]
    B = occ[tuple(sl_b)]
    return int(np.count_nonzero(A & B)), int(A.size)


# ------------------------------------------------------------------ volume ingest

_RAW_DTYPES = {(8, False): "u1", (8, True): "i1", (16, False): "u2", (16, True): "i2", (32, False): "u4",
               (32, True): "i4"}


def decode_raw_slice(data: bytes, bits: int, signed: bool, big_endian: bool) -> np.ndarray:
    """decode_raw_slice (src/io/volume.rs:287-332): samples of the slice's own
    width and sign in the given byte order; misaligned lengths and other widths
    are errors."""
    if (bits, bool(signed)) not in _RAW_DTYPES:
        raise ValueError("RAW bits must be 8/16/32")
    dt = np.dtype(_RAW_DTYPES[(bits, bool(signed))]).newbyteorder(">" if big_endian else "<")
    if len(data) % dt.itemsize:
        raise ValueError(f"RAW byte length is not aligned to {bits}-bit samples")
    return np.frombuffer(data, dtype=dt).astype(np.int64)


def resolve_slice_range(total: int, start: int, end: int):
    """resolve_slice_range (src/io/volume.rs:258-285): inclusive, -1 = first/last."""
    if total == 0:
        raise ValueError("Input folder has no files")
    s = 0 if start < 0 else start
    e = total - 1 if end < 0 else end
    if s >= total or e >= total or s > e:
        raise ValueError(f"Invalid slice range [{start}, {end}] for {total} files")
    return s, e


def boundary_values(vol_zyx: np.ndarray) -> np.ndarray:
    """for_each_boundary_value (src/pipeline/crop.rs:312-330), written as a
    loop over z and y in the reference's order (each boundary voxel once)."""
    D, H, W = vol_zyx.shape
    out = []
    for z in range(D):
        if z == 0 or z + 1 == D:
            out.append(vol_zyx[z].reshape(-1))
            continue
        for y in range(H):
            if y == 0 or y + 1 == H:
                out.append(vol_zyx[z, y])
            else:
                out.append(vol_zyx[z, y, :1])
                if W > 1:
                    out.append(vol_zyx[z, y, W - 1:])
    return np.concatenate(out).astype(np.int64) if out else np.zeros(0, np.int64)


def boundary_mode(vol_zyx: np.ndarray) -> int:
    """detect_background_mode (src/pipeline/crop.rs:338-373): the most frequent
    boundary value, the smallest one on ties; 0 for an empty volume."""
    if vol_zyx.size == 0:
        return 0
    vals, counts = np.unique(boundary_values(vol_zyx), return_counts=True)  # ascending values
    return int(vals[int(np.argmax(counts))])  # argmax takes the first (smallest) maximum


def classify_pack(vol_zyx: np.ndarray, mode: str, lo=0, hi=0, background=0):
    """Occupancy of a z-major sample volume in the S2 layout (x-major, z
    fastest; S2 (x, y, z) = file (x, y, slice)), packed as OccupancyBits.
    mode: "range" (lo <= v <= hi) or "foreground" (v != background)."""
    v = np.asarray(vol_zyx).astype(np.int64)
    occ = (v >= lo) & (v <= hi) if mode == "range" else v != background
    return pack(np.transpose(occ, (2, 1, 0)).astype(np.uint8)), int(np.count_nonzero(occ))


# ------------------------------------------------------------------ STL weld

def _round_half_away(v: np.ndarray) -> np.ndarray:
    t = np.trunc(v)
    frac = v - t  # exact
    return t + np.where(np.abs(frac) >= 0.5, np.sign(v), 0.0)


def weld_key(v: np.ndarray) -> np.ndarray:
    """quantize_key (src/io/stl.rs:55-62): (v * 1e6).round() as i64 per axis --
    round half away from zero, then Rust's saturating float-to-int cast (NaN -> 0)."""
    r = _round_half_away(np.asarray(v, np.float64) * 1000000.0)
    out = np.zeros(r.shape, np.int64)
    ok = ~np.isnan(r)
    hi, lo = r >= 9.2233720368547758e18, r <= -9.2233720368547758e18
    mid = ok & ~hi & ~lo
    out[mid] = r[mid].astype(np.int64)
    out[hi] = np.iinfo(np.int64).max
    out[lo] = np.iinfo(np.int64).min
    return out


def weld(corners: np.ndarray, part=None):
    """dedup_vertex (src/io/stl.rs:64-80) over corners in order: the first corner
    with a key keeps the next vertex index; later corners with that key reuse it.
    part (one id per corner) keeps files apart, as load_stl_or_merge_folder
    welds each file separately.  Returns (vertices (V, 3), tri (N,))."""
    c = np.asarray(corners, np.float64).reshape(-1, 3)
    keys = weld_key(c)
    p = np.zeros(len(c), np.int64) if part is None else np.asarray(part, np.int64)
    index = {}
    tri = np.empty(len(c), np.uint32)
    verts = []
    for i in range(len(c)):
        k = (int(p[i]), int(keys[i, 0]), int(keys[i, 1]), int(keys[i, 2]))
        j = index.get(k)
        if j is None:
            j = index[k] = len(verts)
            verts.append(c[i])
        tri[i] = j
    return np.array(verts, np.float64).reshape(-1, 3), tri
