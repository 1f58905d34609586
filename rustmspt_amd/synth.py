"""Seeded synthetic S2 workloads (voxel units: bbox [0, N]^3, pitch 1).

Random numbers come from numpy's Philox bit generator keyed by (seed, stream), so
every workload is reproducible.  Particles are level-3 icospheres (1280
triangles, src/geometry/mesh_ops.rs:374) or icospheres scaled to ellipsoids and
rotated by a uniform random rotation (Shoemake's quaternion method); centres are
placed by rejection sampling on bounding spheres.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .mesh import BoundingBox, Mesh, merge_meshes, unit_icosphere


def rng_for(seed: int, stream: int = 0) -> np.random.Generator:
    key = np.array([int(seed) & (2 ** 64 - 1), int(stream) & (2 ** 64 - 1)], np.uint64)
    return np.random.Generator(np.random.Philox(key=key))


def random_rotation(rng: np.random.Generator) -> np.ndarray:
    """Uniform SO(3) rotation matrix (Shoemake 1992)."""
    u1, u2, u3 = rng.random(3)
    q = np.array([math.sqrt(1 - u1) * math.sin(2 * math.pi * u2), math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
                  math.sqrt(u1) * math.sin(2 * math.pi * u3), math.sqrt(u1) * math.cos(2 * math.pi * u3)])
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


_UNIT = {}


def _unit(level: int):
    if level not in _UNIT:
        _UNIT[level] = unit_icosphere(level)
    return _UNIT[level]


@dataclass
class Particle:
    center: np.ndarray
    axes: np.ndarray        # semi-axes (a, b, c)
    rotation: np.ndarray    # 3x3
    level: int = 3

    @property
    def bound(self) -> float:
        return float(np.max(self.axes))

    def mesh(self) -> Mesh:
        v, f = _unit(self.level)
        pts = (v * self.axes[None, :]) @ self.rotation.T + self.center[None, :]
        return Mesh(pts, f)

    def moved(self, shift, rotation) -> "Particle":
        return Particle(self.center + np.asarray(shift, float), self.axes.copy(), rotation @ self.rotation, self.level)


def place_particles(n: int, grid: int, rng: np.random.Generator, make, gap: float = 1.0, overlap: bool = False,
                    max_tries: int = 200000):
    """Rejection-sample n particles (make(rng) -> Particle at the origin) inside
    [r, grid - r]^3, bounding spheres at least `gap` apart unless overlap."""
    placed = []
    centers = np.zeros((0, 3))
    radii = np.zeros(0)
    tries = 0
    while len(placed) < n and tries < max_tries:
        tries += 1
        p = make(rng)
        r = p.bound
        if 2 * r + 2 >= grid:
            continue
        c = rng.uniform(r + 0.5, grid - r - 0.5, 3)
        if not overlap and len(placed):
            d = np.linalg.norm(centers - c, axis=1)
            if np.any(d < radii + r + gap):
                continue
        p.center = c
        placed.append(p)
        centers = np.vstack([centers, c])
        radii = np.append(radii, r)
    return placed


def sphere_maker(rmin: float, rmax: float, level: int = 3):
    def make(rng):
        r = rng.uniform(rmin, rmax)
        return Particle(np.zeros(3), np.array([r, r, r]), np.eye(3), level)
    return make


def ellipsoid_maker(r_eq_min: float, r_eq_max: float, level: int = 3):
    """Axis ratios (1, k1, k2) with k uniform in [1, 3], scaled to an equivalent
    radius r_eq (same volume), random orientation."""
    def make(rng):
        r_eq = rng.uniform(r_eq_min, r_eq_max)
        k = rng.uniform(1.0, 3.0, 2)
        a = r_eq / (k[0] * k[1]) ** (1.0 / 3.0)
        return Particle(np.zeros(3), np.array([a, a * k[0], a * k[1]]), random_rotation(rng), level)
    return make


def mixed_maker(r_eq_min: float, r_eq_max: float, level: int = 3):
    s, e = sphere_maker(r_eq_min, r_eq_max, level), ellipsoid_maker(r_eq_min, r_eq_max, level)

    def make(rng):
        return s(rng) if rng.random() < 0.5 else e(rng)
    return make


def packing_mesh(particles) -> Mesh:
    return merge_meshes([p.mesh() for p in particles])


def config_plain(seed: int = 1, n: int = 64, grid: int = 128):
    """64 non-overlapping level-3 icospheres, radii U[0.05, 0.10] of the unit box
    scaled to the grid (U[6.4, 12.8] voxels at 128^3)."""
    rng = rng_for(seed, 1)
    parts = place_particles(n, grid, rng, sphere_maker(0.05 * grid, 0.10 * grid))
    return packing_mesh(parts), BoundingBox.from_size((grid, grid, grid)), parts


def config_large(seed: int = 2, n: int = 500, grid: int = 256, r_eq=(0.040, 0.062)):
    """500 particles, 50/50 icospheres and ellipsoids (axis ratios U[1,3],
    uniform orientation), equivalent radius U[0.040, 0.062]*grid, bounding
    spheres non-overlapping."""
    rng = rng_for(seed, 2)
    parts = place_particles(n, grid, rng, mixed_maker(r_eq[0] * grid, r_eq[1] * grid), gap=0.5)
    return packing_mesh(parts), BoundingBox.from_size((grid, grid, grid)), parts


def config_batched(seed: int = 3, ncand: int = 256, grids=(96, 128, 192, 256)):
    """Batched optimize step: one seed packing per grid (U{150..400} ellipsoids,
    axis ratios U[1,3], equivalent radius for VF ~ 0.2, overlap allowed: the
    occupancy is a union) and ncand one-move candidates, each moving one particle
    of a uniformly drawn seed by a translation U[-2, 2] voxels per axis and a
    uniform random rotation.  Returns (seeds, moves): seeds is a list of
    (grid, particles) with r_max = grid / 4, moves a list of
    (seed index, particle index, new vertices)."""
    rng = rng_for(seed, 77)
    seeds = []
    for g in grids:
        n = int(rng.integers(150, 401))
        r_eq = (0.2 * g ** 3 / n * 3 / (4 * np.pi)) ** (1 / 3)
        seeds.append((g, place_particles(n, g, rng, ellipsoid_maker(0.8 * r_eq, 1.2 * r_eq), overlap=True)))
    moves = []
    for _ in range(ncand):
        si = int(rng.integers(0, len(seeds)))
        pi = int(rng.integers(0, len(seeds[si][1])))
        p = seeds[si][1][pi]
        moves.append((si, pi, p.moved(rng.uniform(-2, 2, 3), random_rotation(rng)).mesh().vertices))
    return seeds, moves


def grf_volume(n: int, corr_len: float, vf: float, seed: int) -> np.ndarray:
    """Gaussian random field (white noise filtered by a Gaussian kernel of the
    given correlation length, periodic FFT filter) thresholded at the (1-vf)
    quantile; returns a uint8 grid, x-major, z fastest."""
    import os

    import scipy.fft as sfft

    workers = max(1, min(16, os.cpu_count() or 1))
    rng = rng_for(seed, 5)
    noise = rng.standard_normal((n, n, n), dtype=np.float32)
    k = [np.fft.fftfreq(n).astype(np.float32)] * 2 + [np.fft.rfftfreq(n).astype(np.float32)]
    f = sfft.rfftn(noise, workers=workers)
    del noise
    s2 = (2 * math.pi * corr_len) ** 2 / 2.0
    f *= np.exp(-s2 * k[0][:, None, None] ** 2)
    f *= np.exp(-s2 * k[1][None, :, None] ** 2)
    f *= np.exp(-s2 * k[2][None, None, :] ** 2)
    field = sfft.irfftn(f, s=(n, n, n), axes=(0, 1, 2), workers=workers).astype(np.float32, copy=False)
    del f
    thr = np.quantile(field, 1.0 - vf)
    return (field > thr).astype(np.uint8)
