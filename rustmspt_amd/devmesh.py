"""Device-resident meshes (s2h_mesh_*): STL weld and granule split on the
device, and voxelization from the resident mesh with no host preparation.

Results equal the host path (mesh.load_stl / load_stl_or_merge_folder +
mesh.granules + s2.voxelize) bit for bit; what changes is where the work runs.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path
from typing import Sequence

import numpy as np

from . import _lib as L
from .mesh import BoundingBox, Mesh
from .s2 import Context, Volume, _dims, occupancy_dims


def stl_paths(path) -> list:
    """A file, or every *.stl of a folder in the host loader's order (sorted)."""
    p = Path(path)
    if p.is_dir():
        paths = sorted(q for q in p.iterdir() if q.suffix.lower() == ".stl")
        if not paths:
            raise ValueError(f"No STL files found in directory: {p}")
        return paths
    return [p]


class DeviceMesh:
    def __init__(self, ctx: Context, handle):
        self.ctx = ctx
        self.handle = handle
        nv, nt, ng = C.c_uint64(), C.c_uint64(), C.c_uint32()
        bmin, bmax = np.zeros(3), np.zeros(3)
        L.check(L.lib().s2h_mesh_info(handle, C.byref(nv), C.byref(nt), C.byref(ng), L.ptr(bmin, L._f64p),
                                      L.ptr(bmax, L._f64p)))
        self.nverts, self.ntris, self.ngranules = int(nv.value), int(nt.value), int(ng.value)
        self.bbox = BoundingBox(tuple(bmin.tolist()), tuple(bmax.tolist())) if self.nverts else None

    @staticmethod
    def load_stl(ctx: Context, path) -> "DeviceMesh":
        """load_stl_or_merge_folder on the device (src/io/stl.rs:282-302)."""
        names = [str(p).encode() for p in stl_paths(path)]
        arr = (C.c_char_p * len(names))(*names)
        h = C.c_void_p()
        L.check(L.lib().s2h_mesh_load_stl(ctx.handle, arr, len(names), C.byref(h)))
        return DeviceMesh(ctx, h)

    @staticmethod
    def from_stl_bytes(ctx: Context, files: Sequence[bytes]) -> "DeviceMesh":
        bufs = [C.create_string_buffer(bytes(b), len(b)) for b in files]
        ptrs = (C.c_void_p * len(bufs))(*[C.cast(b, C.c_void_p) for b in bufs])
        sizes = np.array([len(b) for b in files], np.uint64)
        h = C.c_void_p()
        L.check(L.lib().s2h_mesh_from_stl(ctx.handle, ptrs, L.ptr(sizes, L._u64p), len(files), C.byref(h)))
        return DeviceMesh(ctx, h)

    @staticmethod
    def from_mesh(ctx: Context, mesh: Mesh) -> "DeviceMesh":
        v = np.ascontiguousarray(mesh.vertices, np.float64)
        f = np.ascontiguousarray(mesh.faces, np.uint32)
        h = C.c_void_p()
        L.check(L.lib().s2h_mesh_from_arrays(ctx.handle, L.ptr(v.reshape(-1), L._f64p), len(v),
                                             L.ptr(f.reshape(-1), L._u32p), len(f), C.byref(h)))
        return DeviceMesh(ctx, h)

    def read(self):
        """(Mesh, granule id per face) copied back to the host."""
        xyz = np.zeros((self.nverts, 3), np.float64)
        tri = np.zeros((self.ntris, 3), np.uint32)
        gid = np.zeros(self.ntris, np.uint32)
        L.check(L.lib().s2h_mesh_read(self.ctx.handle, self.handle, L.ptr(xyz.reshape(-1), L._f64p), xyz.size,
                                      L.ptr(tri.reshape(-1), L._u32p), tri.size, L.ptr(gid, L._u32p), gid.size))
        return Mesh(xyz, tri), gid

    def voxelize(self, bbox: BoundingBox, pitch: float) -> Volume:
        dims = occupancy_dims(bbox, pitch)
        h = C.c_void_p()
        occ = C.c_uint64()
        bmin = np.asarray(bbox.min, np.float64)
        L.check(L.lib().s2h_mesh_voxelize(self.ctx.handle, self.handle, L.ptr(bmin, L._f64p), float(pitch),
                                          L.ptr(_dims(dims), L._u64p), C.byref(h), C.byref(occ)))
        return Volume(self.ctx, h)

    def set_vertices(self, vstart: int, xyz) -> None:
        """Move vertices [vstart, vstart + len(xyz)) (a rigid particle move)."""
        v = np.ascontiguousarray(np.asarray(xyz, np.float64).reshape(-1, 3))
        L.check(L.lib().s2h_mesh_set_vertices(self.ctx.handle, self.handle, int(vstart), L.ptr(v.reshape(-1), L._f64p),
                                              len(v)))
        bmin, bmax = np.zeros(3), np.zeros(3)
        L.check(L.lib().s2h_mesh_info(self.handle, None, None, None, L.ptr(bmin, L._f64p), L.ptr(bmax, L._f64p)))
        self.bbox = BoundingBox(tuple(bmin.tolist()), tuple(bmax.tolist())) if self.nverts else None

    def mesh_mc_counts(self, bbox: BoundingBox, r_max: int, samples: int, seed: int) -> np.ndarray:
        """Continuous mesh MC (hits, valid) per radius on the resident mesh."""
        out = np.zeros(2 * (r_max + 1), np.uint64)
        bmin = np.asarray(bbox.min, np.float64)
        bmax = np.asarray(bbox.max, np.float64)
        L.check(L.lib().s2h_mesh_mc_counts_resident(self.ctx.handle, self.handle, L.ptr(bmin, L._f64p),
                                                    L.ptr(bmax, L._f64p), int(r_max), int(samples),
                                                    int(seed) & 0xFFFFFFFFFFFFFFFF, L.ptr(out, L._u64p), len(out)))
        return out.reshape(-1, 2)

    def free(self):
        if self.handle:
            L.lib().s2h_mesh_free(self.ctx.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            if self.ctx.handle:
                self.free()
        except Exception:
            pass
