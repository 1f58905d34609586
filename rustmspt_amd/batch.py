"""Batched evaluation of simulated-annealing candidates (one particle move each)
against resident seed packings: s2h_seed_create / s2h_batch_eval.

Mirrors the per-candidate work of run_sa_island (src/pipeline/optimize.rs:719-739:
coverage.replace, evaluate_voxel_grid, then accept or restore) for many
candidates at once; each returned curve is bit-identical to a full exact
evaluation of the candidate packing.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Sequence

import numpy as np

from . import _lib as L
from .mesh import BoundingBox, Mesh, granules, merge_meshes
from .s2 import Context, Volume, occupancy_dims


class Seed:
    """A resident seed packing built from per-particle meshes."""

    def __init__(self, ctx: Context, particle_meshes: Sequence[Mesh], bbox: BoundingBox, pitch: float, r_max: int):
        self.ctx = ctx
        self.pitch = float(pitch)
        self.r_max = int(r_max)
        self.bbox = bbox
        self.dims = occupancy_dims(bbox, self.pitch)
        self.meshes = [Mesh(m.vertices.copy(), m.faces.copy()) for m in particle_meshes]
        mesh = merge_meshes(self.meshes)
        gid, ng = granules(mesh)
        h = C.c_void_p()
        L.check(L.lib().s2h_seed_create(ctx.handle, L.ptr(mesh.vertices.reshape(-1), L._f64p), len(mesh.vertices),
                                        L.ptr(mesh.faces.reshape(-1), L._u32p), len(mesh.faces), L.ptr(gid, L._u32p),
                                        ng, L.ptr(np.asarray(bbox.min, np.float64), L._f64p), self.pitch,
                                        L.ptr(np.asarray(self.dims, np.uint64), L._u64p), self.r_max, C.byref(h)))
        self.handle = h
        # granule ids of each particle, local to the particle
        self._local_gran = []
        for m in self.meshes:
            g, _ = granules(m)
            self._local_gran.append(np.ascontiguousarray(g, np.uint32))

    def volume(self) -> Volume:
        v = C.c_void_p()
        L.check(L.lib().s2h_seed_volume(self.handle, C.byref(v)))
        vol = Volume(self.ctx, v)
        vol.handle = None  # owned by the seed: never freed through this view
        vol._view = v
        return vol

    def exact_s2(self) -> np.ndarray:
        view = self.volume()
        out = np.zeros(self.r_max + 1, np.float64)
        L.check(L.lib().s2h_exact_s2(self.ctx.handle, view._view, self.r_max, self.pitch, L.ptr(out, L._f64p),
                                     len(out)))
        return out

    def triangles(self, index: int, vertices=None) -> np.ndarray:
        m = self.meshes[index]
        v = m.vertices if vertices is None else np.asarray(vertices, np.float64)
        return np.ascontiguousarray(v[m.faces].reshape(-1))

    def moved_mesh(self, index: int, new_vertices) -> Mesh:
        """The whole packing with particle `index` at new_vertices."""
        parts = list(self.meshes)
        parts[index] = Mesh(np.asarray(new_vertices, np.float64), self.meshes[index].faces)
        return merge_meshes(parts)

    def accept(self, index: int, new_vertices) -> None:
        """Record a move on the host copy (the device seed is rebuilt by the caller)."""
        self.meshes[index] = Mesh(np.asarray(new_vertices, np.float64), self.meshes[index].faces)

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
        old = sd.triangles(pi)
        new = sd.triangles(pi, newv)
        g = sd._local_gran[pi]
        keep += [old, new, g]
        cands[i].seed = si
        cands[i].ntris = len(sd.meshes[pi].faces)
        cands[i].ngranules = int(g.max()) + 1 if len(g) else 0
        cands[i].old_tri = L.ptr(old, L._f64p)
        cands[i].new_tri = L.ptr(new, L._f64p)
        cands[i].granule = L.ptr(g, L._u32p)
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
