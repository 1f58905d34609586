"""Batched evaluation of simulated-annealing candidates (one particle move each)
against resident seed packings: s2h_seed_create / s2h_batch_eval /
s2h_seed_accept.

Mirrors the per-candidate work of run_sa_island (src/pipeline/optimize.rs:719-771:
coverage.replace, evaluate_voxel_grid, then accept or restore) for many
candidates at once; each returned curve is bit-identical to a full exact
evaluation of the candidate packing.  A candidate is (seed, particle, new
vertices): the seed keeps every particle's geometry on the device, so a batch
uploads only the moved particles' new vertex positions.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Sequence

import numpy as np

from . import _lib as L
from .mesh import BoundingBox, Mesh, granules, merge_meshes
from .s2 import Context, Volume, occupancy_dims


class Seed:
    """A resident seed packing built from per-particle meshes (particle i owns
    the i-th mesh's vertices and triangles)."""

    def __init__(self, ctx: Context, particle_meshes: Sequence[Mesh], bbox: BoundingBox, pitch: float, r_max: int):
        self.ctx = ctx
        self.pitch = float(pitch)
        self.r_max = int(r_max)
        self.bbox = bbox
        self.dims = occupancy_dims(bbox, self.pitch)
        self.meshes = [Mesh(np.ascontiguousarray(m.vertices, np.float64), m.faces.copy()) for m in particle_meshes]
        mesh = merge_meshes(self.meshes)
        gid, ng = granules(mesh)
        vstart = np.zeros(len(self.meshes) + 1, np.uint32)
        tstart = np.zeros(len(self.meshes) + 1, np.uint32)
        vstart[1:] = np.cumsum([len(m.vertices) for m in self.meshes])
        tstart[1:] = np.cumsum([len(m.faces) for m in self.meshes])
        h = C.c_void_p()
        L.check(L.lib().s2h_seed_create(ctx.handle, L.ptr(mesh.vertices.reshape(-1), L._f64p), len(mesh.vertices),
                                        L.ptr(mesh.faces.reshape(-1), L._u32p), len(mesh.faces), L.ptr(gid, L._u32p),
                                        ng, L.ptr(vstart, L._u32p), L.ptr(tstart, L._u32p), len(self.meshes),
                                        L.ptr(np.asarray(bbox.min, np.float64), L._f64p), self.pitch,
                                        L.ptr(np.asarray(self.dims, np.uint64), L._u64p), self.r_max, C.byref(h)))
        self.handle = h

    def volume(self) -> Volume:
        v = C.c_void_p()
        L.check(L.lib().s2h_seed_volume(self.handle, C.byref(v)))
        return Volume(self.ctx, v, owned=False)

    def exact_s2(self) -> np.ndarray:
        return self.volume().exact_s2(self.r_max, self.pitch)

    def moved_mesh(self, index: int, new_vertices) -> Mesh:
        """The whole packing with particle `index` at new_vertices."""
        parts = list(self.meshes)
        parts[index] = Mesh(np.asarray(new_vertices, np.float64), self.meshes[index].faces)
        return merge_meshes(parts)

    def accept(self, index: int, new_vertices) -> np.ndarray:
        """Commit a move on the device (coverage, rows, counts, geometry) and
        return the new seed's exact curve."""
        newv = np.ascontiguousarray(np.asarray(new_vertices, np.float64).reshape(-1, 3))
        if newv.shape != self.meshes[index].vertices.shape:
            raise ValueError("new vertices must match the particle's vertex count")
        out = np.zeros(self.r_max + 1, np.float64)
        L.check(L.lib().s2h_seed_accept(self.ctx.handle, self.handle, int(index), L.ptr(newv.reshape(-1), L._f64p),
                                        L.ptr(out, L._f64p), len(out)))
        self.meshes[index] = Mesh(newv.copy(), self.meshes[index].faces)
        return out

    def free(self):
        if self.handle:
            L.lib().s2h_seed_free(self.ctx.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            if self.ctx.handle:
                self.free()
        except Exception:
            pass


def batch_eval(ctx: Context, seeds: List[Seed], moves) -> List[np.ndarray]:
    """moves: iterable of (seed_index, particle_index, new_vertices).  Returns one
    exact S2 curve (r_max(seed)+1 values) per move."""
    keep = []
    cands = (L.Candidate * max(len(moves), 1))()
    total = 0
    for i, (si, pi, newv) in enumerate(moves):
        sd = seeds[si]
        v = np.ascontiguousarray(newv, np.float64)
        if v.shape != sd.meshes[pi].vertices.shape:
            raise ValueError("new vertices must match the particle's vertex count")
        keep.append(v)
        cands[i].seed = si
        cands[i].particle = pi
        cands[i].new_xyz = L.ptr(v.reshape(-1), L._f64p)
        total += sd.r_max + 1
    out = np.zeros(max(total, 1), np.float64)
    handles = (C.c_void_p * max(len(seeds), 1))(*[s.handle for s in seeds])
    L.check(L.lib().s2h_batch_eval(ctx.handle, handles, len(seeds), C.cast(cands, C.c_void_p), len(moves),
                                   L.ptr(out, L._f64p), total))
    curves, off = [], 0
    for si, _, _ in moves:
        n = seeds[si].r_max + 1
        curves.append(out[off:off + n].copy())
        off += n
    return curves
