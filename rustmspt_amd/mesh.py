"""Triangle meshes for the S2 path: STL ingest (through the library's C++ loader,
which applies the reference's f32 -> f64 conversion and 1e-6 weld), binary STL
output, granule split, and the reference's test geometry (icosphere, box).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import _lib as L


@dataclass
class Mesh:
    """Welded triangle mesh: vertices (n, 3) f64, faces (m, 3) u32 (src/types.rs:124-150)."""

    vertices: np.ndarray
    faces: np.ndarray

    def __post_init__(self):
        self.vertices = np.ascontiguousarray(np.asarray(self.vertices, dtype=np.float64).reshape(-1, 3))
        self.faces = np.ascontiguousarray(np.asarray(self.faces, dtype=np.uint32).reshape(-1, 3))

    def is_empty(self) -> bool:
        return len(self.vertices) == 0 or len(self.faces) == 0


@dataclass
class BoundingBox:
    min: tuple
    max: tuple

    def size(self):
        return tuple(self.max[i] - self.min[i] for i in range(3))

    @staticmethod
    def from_size(size):
        return BoundingBox((0.0, 0.0, 0.0), tuple(float(v) for v in size))


def mesh_bbox(mesh: Mesh):
    """src/geometry/bbox.rs:8-25 (NaN coordinates are ignored as f64::min/max do)."""
    if len(mesh.vertices) == 0:
        return None
    return BoundingBox(tuple(np.fmin.reduce(mesh.vertices, axis=0).tolist()),
                       tuple(np.fmax.reduce(mesh.vertices, axis=0).tolist()))


def load_stl(path) -> Mesh:
    """src/io/stl.rs:214-239 load_stl (ASCII sniff, binary fallback, 1e-6 weld)."""
    xyz = L._f64p()
    tri = L._u32p()
    nv = C.c_uint64()
    nt = C.c_uint64()
    L.check(L.lib().s2h_stl_load(str(path).encode(), C.byref(xyz), C.byref(nv), C.byref(tri), C.byref(nt)))
    try:
        v = np.ctypeslib.as_array(xyz, shape=(max(nv.value, 1) * 3,))[: nv.value * 3].copy()
        f = np.ctypeslib.as_array(tri, shape=(max(nt.value, 1) * 3,))[: nt.value * 3].copy()
    finally:
        L.lib().s2h_free(C.cast(xyz, C.c_void_p))
        L.lib().s2h_free(C.cast(tri, C.c_void_p))
    return Mesh(v.reshape(-1, 3), f.reshape(-1, 3))


def load_stl_or_merge_folder(path) -> Mesh:
    """src/io/stl.rs:282-302: a file, or every *.stl of a folder merged in order."""
    p = Path(path)
    if p.is_dir():
        paths = sorted(q for q in p.iterdir() if q.suffix.lower() == ".stl")
        if not paths:
            raise ValueError(f"No STL files found in directory: {p}")
        return merge_meshes([load_stl(q) for q in paths])
    return load_stl(p)


def save_stl(path, mesh: Mesh, solid_name: str = "rustmspt") -> None:
    """src/io/stl.rs:311-353 save_stl: binary, f32 vertices, zero normals."""
    if mesh.is_empty():
        raise ValueError("Cannot write empty mesh")
    p = Path(path)
    p.parent.mkdir(parents=True, exist_ok=True)
    header = solid_name.encode()[:80].ljust(80, b"\0")
    rec = np.zeros(len(mesh.faces), dtype=[("n", "<f4", 3), ("v", "<f4", (3, 3)), ("attr", "<u2")])
    rec["v"] = mesh.vertices[mesh.faces].astype(np.float32)
    with open(p, "wb") as fh:
        fh.write(header)
        fh.write(np.uint32(len(mesh.faces)).tobytes())
        fh.write(rec.tobytes())


def granules(mesh: Mesh):
    """split_mesh_into_granules (src/geometry/mesh_ops.rs:48): granule id per face."""
    out = np.zeros(max(len(mesh.faces), 1), np.uint32)
    n = C.c_uint32()
    L.check(L.lib().s2h_mesh_granules(L.ptr(mesh.faces.reshape(-1), L._u32p), len(mesh.faces), len(mesh.vertices),
                                      L.ptr(out, L._u32p), C.byref(n)))
    return out[: len(mesh.faces)], int(n.value)


def merge_meshes(meshes) -> Mesh:
    """src/geometry/mesh_ops.rs:25-40."""
    verts, faces, off = [], [], 0
    for m in meshes:
        verts.append(m.vertices)
        faces.append(m.faces.astype(np.int64) + off)
        off += len(m.vertices)
    if not verts:
        return Mesh(np.zeros((0, 3)), np.zeros((0, 3), np.uint32))
    return Mesh(np.concatenate(verts), np.concatenate(faces).astype(np.uint32))


def box_mesh(bmin, bmax) -> Mesh:
    """src/geometry/mesh_ops.rs:329-~360 box_mesh (8 vertices, 12 faces, same winding)."""
    (x0, y0, z0), (x1, y1, z1) = bmin, bmax
    v = [(x0, y0, z0), (x1, y0, z0), (x1, y1, z0), (x0, y1, z0),
         (x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1)]
    f = [(0, 1, 2), (0, 2, 3), (4, 6, 5), (4, 7, 6), (0, 4, 5), (0, 5, 1),
         (1, 5, 6), (1, 6, 2), (2, 6, 7), (2, 7, 3), (3, 7, 4), (3, 4, 0)]
    return Mesh(np.array(v, np.float64), np.array(f, np.uint32))


_ICO_FACES = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
              (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
              (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]


def unit_icosphere(level: int):
    """Vertices on the unit sphere and faces of src/geometry/mesh_ops.rs:374-~445
    icosphere_mesh (subdivision by shared-edge midpoints, then normalisation)."""
    phi = (1.0 + 5.0 ** 0.5) / 2.0
    verts = [(-1.0, phi, 0.0), (1.0, phi, 0.0), (-1.0, -phi, 0.0), (1.0, -phi, 0.0),
             (0.0, -1.0, phi), (0.0, 1.0, phi), (0.0, -1.0, -phi), (0.0, 1.0, -phi),
             (phi, 0.0, -1.0), (phi, 0.0, 1.0), (-phi, 0.0, -1.0), (-phi, 0.0, 1.0)]
    faces = list(_ICO_FACES)
    for _ in range(level):
        mids = {}
        nxt = []
        for tri in faces:
            mid = []
            for i, j in ((tri[0], tri[1]), (tri[1], tri[2]), (tri[2], tri[0])):
                key = (i, j) if i < j else (j, i)
                if key not in mids:
                    a, b = verts[i], verts[j]
                    verts.append(((a[0] + b[0]) * 0.5, (a[1] + b[1]) * 0.5, (a[2] + b[2]) * 0.5))
                    mids[key] = len(verts) - 1
                mid.append(mids[key])
            nxt += [(tri[0], mid[0], mid[2]), (mid[0], tri[1], mid[1]), (mid[2], mid[1], tri[2]),
                    (mid[0], mid[1], mid[2])]
        faces = nxt
    v = np.array(verts, np.float64)
    n = np.sqrt(v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1] + v[:, 2] * v[:, 2])
    v = v * (1.0 / n)[:, None]
    return v, np.array(faces, np.uint32)


def icosphere_mesh(center, radius: float, level: int) -> Mesh:
    v, f = unit_icosphere(level)
    return Mesh(np.asarray(center, np.float64)[None, :] + v * float(radius), f)
