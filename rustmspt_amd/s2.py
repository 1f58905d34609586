"""Host-side mirror of RustMSPT's S2 entry points (src/geometry/s2.rs) on the HIP
engine.  Every computation goes through libs2hip.so; this module only marshals
arguments and applies the reference's argument rules (pitch floors, empty-grid
zeros, method names).

Reference API            -> here
  VoxelS2::new            -> VoxelS2Hip(ctx, mesh, bbox, pitch)
  VoxelS2::calculate[_seeded]/calculate_exact_with -> VoxelS2Hip.calculate*(...)
  calculate_s2[_seeded]   -> calculate_s2(ctx, mesh, bbox, r_max, pitch, method, samples, seed)
  try_calculate_s2_gpu_exact[_limited] -> try_calculate_s2_hip_exact(...)
  shell_offset_iter + offset_in_domain -> shell_order_offsets(...)
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Optional

import numpy as np

from . import _lib as L
from .mesh import BoundingBox, Mesh, granules


def _dims(dims):
    return np.ascontiguousarray(np.asarray(dims, dtype=np.uint64).reshape(3))


def occupancy_dims(bbox: BoundingBox, pitch: float):
    """src/geometry/s2.rs:110-117."""
    out = []
    for a in range(3):
        n = math.ceil((bbox.max[a] - bbox.min[a]) / pitch)
        out.append(max(1, int(n) if n > 0 else 0))
    return out


def row_words(nz: int) -> int:
    return (int(nz) + 63) // 64 + 1


class Context:
    """One device context (s2h_create).  hbm_budget_bytes=0: no budget."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        h = C.c_void_p()
        L.check(L.lib().s2h_create(int(device), int(hbm_budget_bytes), C.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            L.lib().s2h_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def device_name(self) -> str:
        buf = C.create_string_buffer(256)
        L.check(L.lib().s2h_device_name(self.handle, buf, 256))
        return buf.value.decode()

    def profile(self, enable: bool = True) -> None:
        L.check(L.lib().s2h_profile(self.handle, 1 if enable else 0))

    def profile_read(self):
        """Last call's device times (ms): counts, gather+fold, total, launches."""
        out = np.zeros(4, np.float64)
        L.check(L.lib().s2h_profile_read(self.handle, L.ptr(out, L._f64p), 4))
        return {"count_ms": out[0], "fold_ms": out[1], "device_ms": out[2], "launches": int(out[3])}

    def valu_probe(self) -> float:
        """Measured int32 VALU lane-operations per second (s2h_valu_probe)."""
        v = C.c_double()
        L.check(L.lib().s2h_valu_probe(self.handle, C.byref(v)))
        return float(v.value)

    def exact_tune_info(self, dims, r_max: int, pitch: float = 1.0) -> dict:
        """Launch shape of the exact count kernel for this plan (see s2h_exact_tune_info)."""
        out = np.zeros(8, np.int64)
        L.check(L.lib().s2h_exact_tune_info(self.handle, L.ptr(_dims(dims), L._u64p), int(r_max), float(pitch),
                                            L.ptr(out, L._i64p), 8))
        return {"band": int(out[0]), "waves": int(out[1]),
                "picked_by": {1: "device timing", -1: "cost model", 0: "not run"}[int(out[2])],
                "timed_candidates": int(out[3]), "items": int(out[4]), "groups": int(out[5]),
                "model_wave_instructions": int(out[6]),
                "kernel": {0: "none", 1: "direct", 2: "fft64"}[int(out[7])]}

    def set_autotune(self, enable: bool) -> None:
        L.check(L.lib().s2h_set_autotune(self.handle, 1 if enable else 0))

    EXACT_METHODS = {"auto": 0, "direct": 1, "fft64": 2}

    def set_exact_method(self, method: str) -> None:
        """Exact kernel of exact_counts / exact_s2: "direct", "fft64" (FP64 FFT
        rounded to the integer counts; only where its round-off bound allows) or
        "auto" (the faster modelled one).  Mirrors ExactCpuMethod and
        plan_exact_cpu (src/geometry/s2.rs:565-569, 733-766)."""
        L.check(L.lib().s2h_exact_method_set(self.handle, self.EXACT_METHODS[method]))

    def exact_tune_set(self, dims, r_max: int, pitch: float, band: int, waves: int) -> None:
        L.check(L.lib().s2h_exact_tune_set(self.handle, L.ptr(_dims(dims), L._u64p), int(r_max), float(pitch),
                                           int(band), int(waves)))

    @property
    def bytes_in_use(self) -> int:
        v = C.c_uint64()
        L.check(L.lib().s2h_bytes_in_use(self.handle, C.byref(v)))
        return int(v.value)


class Volume:
    """A device-resident packed occupancy volume (s2h_vol)."""

    def __init__(self, ctx: Context, handle, owned: bool = True):
        self.ctx = ctx
        self.handle = handle
        self.owned = owned  # False: a view of a volume owned elsewhere (a seed's), never freed here
        d = np.zeros(3, np.uint64)
        L.check(L.lib().s2h_vol_dims(handle, L.ptr(d, L._u64p)))
        self.dims = [int(v) for v in d]

    @staticmethod
    def from_bytes(ctx: Context, occ: np.ndarray) -> "Volume":
        occ = np.ascontiguousarray(occ, dtype=np.uint8)
        h = C.c_void_p()
        L.check(L.lib().s2h_vol_from_bytes(ctx.handle, L.ptr(occ.reshape(-1), L._u8p), L.ptr(_dims(occ.shape), L._u64p),
                                           C.byref(h)))
        return Volume(ctx, h)

    @staticmethod
    def from_packed(ctx: Context, words: np.ndarray, dims) -> "Volume":
        words = np.ascontiguousarray(words, dtype=np.uint64).reshape(-1)
        h = C.c_void_p()
        L.check(L.lib().s2h_vol_from_packed(ctx.handle, L.ptr(words, L._u64p), len(words), L.ptr(_dims(dims), L._u64p),
                                            C.byref(h)))
        return Volume(ctx, h)

    @property
    def nwords(self) -> int:
        return self.dims[0] * self.dims[1] * row_words(self.dims[2])

    def read_packed(self) -> np.ndarray:
        out = np.zeros(self.nwords, np.uint64)
        L.check(L.lib().s2h_vol_read_packed(self.ctx.handle, self.handle, L.ptr(out, L._u64p), len(out)))
        return out

    def to_bytes(self) -> np.ndarray:
        """Unpack to a bool grid (x-major, z fastest) on the host."""
        w = self.read_packed().reshape(self.dims[0], self.dims[1], row_words(self.dims[2]))
        bits = np.unpackbits(w.view(np.uint8).reshape(self.dims[0], self.dims[1], -1), axis=2, bitorder="little")
        return bits[:, :, : self.dims[2]].astype(np.uint8)

    @property
    def occupied(self) -> int:
        v = C.c_uint64()
        L.check(L.lib().s2h_vol_occupied(self.ctx.handle, self.handle, C.byref(v)))
        return int(v.value)

    def free(self):
        if self.handle and self.owned:
            L.lib().s2h_vol_free(self.ctx.handle, self.handle)
        self.handle = None

    def __del__(self):
        try:
            if self.ctx.handle:
                self.free()
        except Exception:
            pass

    # ---------------------------------------------------------------- kernels
    def exact_counts(self, r_max: int, pitch: float = 1.0) -> np.ndarray:
        n = shell_count(self.dims, r_max, pitch)
        out = np.zeros(max(n, 1), np.uint64)
        L.check(L.lib().s2h_exact_counts(self.ctx.handle, self.handle, int(r_max), float(pitch),
                                         L.ptr(out, L._u64p), n))
        return out[:n]

    def exact_s2(self, r_max: int, pitch: float = 1.0) -> np.ndarray:
        out = np.zeros(int(r_max) + 1, np.float64)
        L.check(L.lib().s2h_exact_s2(self.ctx.handle, self.handle, int(r_max), float(pitch), L.ptr(out, L._f64p),
                                     len(out)))
        return out

    def fft32_counts(self, r_max: int, pitch: float = 1.0) -> np.ndarray:
        n = shell_count(self.dims, r_max, pitch)
        out = np.zeros(max(n, 1), np.float32)
        L.check(L.lib().s2h_fft32_counts(self.ctx.handle, self.handle, int(r_max), float(pitch),
                                         L.ptr(out, L._f32p), n))
        return out[:n]

    def fft32_s2(self, r_max: int, pitch: float = 1.0) -> np.ndarray:
        out = np.zeros(int(r_max) + 1, np.float64)
        L.check(L.lib().s2h_fft32_s2(self.ctx.handle, self.handle, int(r_max), float(pitch), L.ptr(out, L._f64p),
                                     len(out)))
        return out

    def mc_counts(self, r_max: int, pitch: float, samples: int, seed: int) -> np.ndarray:
        out = np.zeros(2 * (int(r_max) + 1), np.uint64)
        L.check(L.lib().s2h_mc_counts(self.ctx.handle, self.handle, int(r_max), float(pitch), int(samples),
                                      int(seed) & 0xFFFFFFFFFFFFFFFF, L.ptr(out, L._u64p), len(out)))
        return out.reshape(-1, 2)

    def mc_s2(self, r_max: int, pitch: float, samples: int, seed: int) -> np.ndarray:
        out = np.zeros(int(r_max) + 1, np.float64)
        L.check(L.lib().s2h_mc_s2(self.ctx.handle, self.handle, int(r_max), float(pitch), int(samples),
                                  int(seed) & 0xFFFFFFFFFFFFFFFF, L.ptr(out, L._f64p), len(out)))
        return out


def shell_count(dims, r_max: int, pitch: float) -> int:
    n = C.c_uint64()
    L.check(L.lib().s2h_shell_count(L.ptr(_dims(dims), L._u64p), int(r_max), float(pitch), C.byref(n)))
    return int(n.value)


def shell_order_offsets(dims, r_max: int, pitch: float):
    """In-domain offsets of radii 1..r_max in the reference's summation order and
    radius_start (offsets of radius r are [radius_start[r], radius_start[r+1]))."""
    n = shell_count(dims, r_max, pitch)
    offs = np.zeros((max(n, 1), 3), np.int64)
    rs = np.zeros(int(r_max) + 2, np.uint64)
    L.check(L.lib().s2h_shell_offsets(L.ptr(_dims(dims), L._u64p), int(r_max), float(pitch), L.ptr(offs, L._i64p), n,
                                      L.ptr(rs, L._u64p), len(rs)))
    return offs[:n], rs


def fft_padded_dims(dims, r_max: int, pitch: float = 1.0):
    """Padded lengths of the FFT kernels for this shape (s2h_fft_padded_dims)."""
    out = np.zeros(3, np.uint64)
    L.check(L.lib().s2h_fft_padded_dims(L.ptr(_dims(dims), L._u64p), int(r_max), float(pitch), L.ptr(out, L._u64p)))
    return [int(v) for v in out]


def fill_missing(values, has_support, vf):
    v = np.ascontiguousarray(values, np.float64).copy()
    s = np.ascontiguousarray(has_support, np.uint8)
    L.check(L.lib().s2h_fill_missing(L.ptr(v, L._f64p), L.ptr(s, L._u8p), len(v), float(vf)))
    return v


def voxelize(ctx: Context, mesh: Mesh, bbox: BoundingBox, pitch: float) -> Volume:
    """build_bbox_occupancy on the device (src/geometry/s2.rs:160-201)."""
    dims = occupancy_dims(bbox, pitch)
    gid, ng = granules(mesh)
    h = C.c_void_p()
    occ = C.c_uint64()
    bmin = np.asarray(bbox.min, np.float64)
    L.check(L.lib().s2h_voxelize(ctx.handle, L.ptr(mesh.vertices.reshape(-1), L._f64p), len(mesh.vertices),
                                 L.ptr(mesh.faces.reshape(-1), L._u32p), len(mesh.faces), L.ptr(gid, L._u32p), ng,
                                 L.ptr(bmin, L._f64p), float(pitch), L.ptr(_dims(dims), L._u64p), C.byref(h),
                                 C.byref(occ)))
    return Volume(ctx, h)


class VoxelS2Hip:
    """VoxelS2 (src/geometry/s2.rs:1371-1541) with a device-resident grid."""

    def __init__(self, ctx: Context, mesh: Mesh, bbox: BoundingBox, pitch: float):
        self.ctx = ctx
        self.pitch = max(float(pitch), 1e-9)
        self.volume = voxelize(ctx, mesh, bbox, self.pitch)
        self.dims = self.volume.dims

    @staticmethod
    def from_volume(volume: Volume, pitch: float) -> "VoxelS2Hip":
        obj = VoxelS2Hip.__new__(VoxelS2Hip)
        obj.ctx = volume.ctx
        obj.pitch = max(float(pitch), 1e-9)
        obj.volume = volume
        obj.dims = volume.dims
        return obj

    def calculate_exact(self, r_max: int) -> np.ndarray:
        return self.volume.exact_s2(r_max, self.pitch)

    def calculate(self, r_max: int, method: str, samples: int) -> np.ndarray:
        return self.calculate_seeded(r_max, method, samples, None)

    def calculate_seeded(self, r_max: int, method: str, samples: int, seed: Optional[int]) -> np.ndarray:
        """exact -> exact counts; any other method -> voxel Monte Carlo.  The
        reference draws an unseeded stream when seed is None; the device stream
        is counter-based, so None uses seed 0."""
        if method == "exact":
            return self.volume.exact_s2(r_max, self.pitch)
        if method == "fft32":
            return self.volume.fft32_s2(r_max, self.pitch)
        return self.volume.mc_s2(r_max, self.pitch, max(int(samples), 200), 0 if seed is None else int(seed))


def _mesh_args(mesh: Mesh, bbox: BoundingBox):
    v = np.ascontiguousarray(mesh.vertices, np.float64).reshape(-1, 3)
    f = np.ascontiguousarray(mesh.faces, np.uint32).reshape(-1, 3)
    lo = np.ascontiguousarray(bbox.min, np.float64)
    hi = np.ascontiguousarray(bbox.max, np.float64)
    return v, f, lo, hi


def geometric_volume_fraction(mesh: Mesh, bbox: BoundingBox) -> float:
    """volume_fraction_in_bbox (src/geometry/volume.rs:437-469): every granule's
    in-box volume (its closed-mesh volume when it lies wholly inside the box,
    else the volume of the granule clipped by the box planes with capped cuts),
    summed in granule order, over the box volume, clamped to [0, 1]
    (s2h_volume_fraction_in_bbox, host C++)."""
    v, f, lo, hi = _mesh_args(mesh, bbox)
    out = C.c_double()
    L.check(L.lib().s2h_volume_fraction_in_bbox(L.ptr(v.reshape(-1), L._f64p), len(v), L.ptr(f.reshape(-1), L._u32p),
                                                len(f), L.ptr(lo, L._f64p), L.ptr(hi, L._f64p), C.byref(out)))
    return float(out.value)


def mesh_volume_in_bbox(mesh: Mesh, bbox: BoundingBox) -> float:
    """particle_volume_in_bbox (src/geometry/volume.rs:424-430) of one closed
    mesh (s2h_mesh_volume_in_bbox)."""
    v, f, lo, hi = _mesh_args(mesh, bbox)
    out = C.c_double()
    L.check(L.lib().s2h_mesh_volume_in_bbox(L.ptr(v.reshape(-1), L._f64p), len(v), L.ptr(f.reshape(-1), L._u32p),
                                            len(f), L.ptr(lo, L._f64p), L.ptr(hi, L._f64p), C.byref(out)))
    return float(out.value)


def ordered_sum(ctx: Context, terms, starts) -> np.ndarray:
    """Left-to-right f64 sums of the segments [starts[g], starts[g+1]) of terms,
    folded on the device by the exact curve's fold (s2h_ordered_sum)."""
    t = np.ascontiguousarray(terms, np.float64)
    st = np.ascontiguousarray(starts, np.uint64)
    out = np.zeros(max(len(st) - 1, 0), np.float64)
    L.check(L.lib().s2h_ordered_sum(ctx.handle, L.ptr(t, L._f64p), len(t), L.ptr(st, L._u64p), len(out),
                                    L.ptr(out, L._f64p)))
    return out


def mesh_mc_counts(ctx: Context, mesh: Mesh, bbox: BoundingBox, r_max: int, samples: int, seed: int) -> np.ndarray:
    """Continuous mesh Monte Carlo (hits, valid) per radius (s2h_mesh_mc_counts)."""
    out = np.zeros(2 * (int(r_max) + 1), np.uint64)
    v = np.ascontiguousarray(mesh.vertices, np.float64)
    f = np.ascontiguousarray(mesh.faces, np.uint32)
    L.check(L.lib().s2h_mesh_mc_counts(ctx.handle, L.ptr(v.reshape(-1), L._f64p), len(v), L.ptr(f.reshape(-1), L._u32p),
                                       len(f), L.ptr(np.asarray(bbox.min, np.float64), L._f64p),
                                       L.ptr(np.asarray(bbox.max, np.float64), L._f64p), int(r_max), int(samples),
                                       int(seed) & 0xFFFFFFFFFFFFFFFF, L.ptr(out, L._u64p), len(out)))
    return out.reshape(-1, 2)


def mesh_mc_s2(ctx: Context, mesh: Mesh, bbox: BoundingBox, r_max: int, samples: int, seed: int) -> np.ndarray:
    """calculate_s2_mesh_mc_seeded (src/geometry/s2.rs:1273-1320): S2(0) is the
    geometric volume fraction in the box, S2(r) = hits / valid."""
    vf = geometric_volume_fraction(mesh, bbox)
    out = np.zeros(int(r_max) + 1, np.float64)
    v = np.ascontiguousarray(mesh.vertices, np.float64)
    f = np.ascontiguousarray(mesh.faces, np.uint32)
    L.check(L.lib().s2h_mesh_mc_s2(ctx.handle, L.ptr(v.reshape(-1), L._f64p), len(v), L.ptr(f.reshape(-1), L._u32p),
                                   len(f), L.ptr(np.asarray(bbox.min, np.float64), L._f64p),
                                   L.ptr(np.asarray(bbox.max, np.float64), L._f64p), int(r_max), int(samples),
                                   int(seed) & 0xFFFFFFFFFFFFFFFF, float(vf), L.ptr(out, L._f64p), len(out)))
    return out


def calculate_s2(ctx: Context, mesh: Mesh, bbox: BoundingBox, r_max: int, voxel_pitch: float, method: str,
                 samples: int, seed: Optional[int] = None) -> np.ndarray:
    """calculate_s2_seeded (src/geometry/s2.rs:1345-1368) on the HIP engine: a
    non-exact method with voxel_pitch <= 0 is the continuous mesh Monte Carlo
    (seed None uses seed 0: the device stream is counter-based)."""
    if method != "exact" and voxel_pitch <= 0.0:
        return mesh_mc_s2(ctx, mesh, bbox, r_max, samples, 0 if seed is None else int(seed))
    pitch = voxel_pitch
    if voxel_pitch <= 0.0:
        print("[Warning] exact S2 requested with voxel_pitch <= 0; falling back to voxel_pitch=1.0")
        pitch = 1.0
    grid = VoxelS2Hip(ctx, mesh, bbox, pitch)
    try:
        return grid.calculate_seeded(r_max, method, samples, seed)
    finally:
        grid.volume.free()


def try_calculate_s2_hip_exact(ctx: Context, mesh: Mesh, bbox: BoundingBox, r_max: int, voxel_pitch: float,
                               limit_mb: Optional[int] = None) -> np.ndarray:
    """try_calculate_s2_gpu_exact_limited (src/geometry/s2.rs:1627-1678): checked
    grid normalisation, explicit errors (the caller applies the fallback policy)."""
    if r_max < 0 or r_max > 0xFFFFFFFF:
        raise L.S2HipError(1, "GPU exact radius count exceeds supported range")
    if not math.isfinite(voxel_pitch):
        raise L.S2HipError(1, "S2 voxel pitch must be finite")
    pitch = 1.0 if voxel_pitch <= 0.0 else voxel_pitch
    for a in range(3):
        size = bbox.max[a] - bbox.min[a]
        n = max(math.ceil(size / pitch), 1.0) if math.isfinite(size / pitch) else float("inf")
        if not math.isfinite(size) or size <= 0.0 or not math.isfinite(n) or n > 65535:
            raise L.S2HipError(5, "GPU exact grid dimensions exceed supported range")
    if limit_mb is not None:
        dims = occupancy_dims(bbox, pitch)
        need = dims[0] * dims[1] * row_words(dims[2]) * 8 + len(mesh.faces) * 200
        if need > limit_mb * 1024 * 1024:
            raise L.S2HipError(3, f"GPU exact working set {need} bytes exceeds budget {limit_mb * 1024 * 1024} bytes")
    return calculate_s2(ctx, mesh, bbox, r_max, pitch, "exact", 0)


def l2_norm(a, b) -> float:
    """src/geometry/s2.rs:1549-1559: over the common prefix (sequential sum)."""
    n = min(len(a), len(b))
    s = 0.0
    for i in range(n):
        d = float(a[i]) - float(b[i])
        s += d * d
    return math.sqrt(s)
