"""`optimize` on the HIP engine: the S2 backend choice, the evaluator and the
simulated-annealing island loop whose S2 evaluations run as batched exact
delta counts against a resident seed packing.

Reference interface -> here
  S2Method::resolve / name            (src/pipeline/optimize_execution.rs:14-39)   S2Method
  resolve_mode                        (42-55)                                       config.configured_mode
  select_s2_backend                   (58-134)                                      select_s2_backend
  OptimizeS2::{new, evaluate}         (142-343)                                     OptimizeS2Hip
  stage_rng / fixed_eval_seed         (src/pipeline/optimize.rs:80-101)             stage_key / fixed_eval_seed
  run_sa_island                       (359-833)                                     run_sa_island
  OptimizePipeline::run_in_pool       (870-1062)                                    OptimizePipeline.run

The island loop keeps the reference's schedule: one particle per iteration
moved by a translation (60 %), a pull toward another particle (30 %) or a
random jump (10 %), each with a rotation about the centroid; the boundary
constraint modes 1-3 (check_boundary_constraints_mode, src/geometry/bbox.rs:72-131);
S2 of the candidate, L2 loss to the target, Metropolis acceptance; the
cooling and the adaptive acceptance window; best tracking and the history
lines.  Differences, by design:
- Random numbers: every iteration draws from its own Philox4x32 stream keyed
  by (stage key, iteration), so a seeded run is reproducible; the reference's
  ChaCha12 stream is not reproduced.
- Batching: the iterations from the current one up to the next acceptance
  change nothing but the temperature schedule, which is known in advance when
  no move is accepted.  The loop therefore proposes a run of such iterations,
  evaluates all their candidates in one s2h_batch_eval (exact delta counts,
  one graph replay), then walks them in order: the first accepted candidate is
  committed on the device (s2h_seed_accept) and the rest are dropped.  With
  per-iteration streams this is exactly the sequential loop: the trajectory
  does not depend on the batch length (batch=1 is the sequential loop).
- Collisions: the reference rejects moves whose triangle meshes intersect
  another particle or come closer than min_neighbor_distance (parry3d's exact
  narrow phase).  That geometry test is not part of this engine; a caller
  passes `blocked(candidate_vertices, index, particles)` to supply one, and
  by default overlapping particles are allowed (the occupancy is their union).
- Pruning and several islands with migration are not part of this engine:
  prune_enabled is reported and skipped, islands > 1 is refused.
"""
from __future__ import annotations

import math
import os
import sys
import time
from dataclasses import dataclass
from pathlib import Path
from typing import Callable, List, Optional

import numpy as np

from . import batch as B
from . import config as cfg
from . import report
from . import s2
from .mesh import BoundingBox, Mesh, granules, load_stl, merge_meshes, mesh_bbox, save_stl

MASK64 = (1 << 64) - 1
PRUNE_STAGE, TARGET_STAGE, INPUT_STAGE, FINAL_STAGE, ISLAND_STAGE = 1, 2, 3, 4, 100


def stage_key(seed: Optional[int], stage: int) -> int:
    """The stage's stream key (stage_rng's seed mix, src/pipeline/optimize.rs:90-96);
    an unseeded run takes a key from the OS."""
    if seed is None:
        return int.from_bytes(os.urandom(8), "little")
    return (seed ^ ((stage * 0xD1B54A32D192ED03) & MASK64)) & MASK64


def fixed_eval_seed(seed: Optional[int], stage: int) -> Optional[int]:
    """src/pipeline/optimize.rs:98-101."""
    if seed is None:
        return None
    return (seed ^ ((stage * 0x9E3779B97F4A7C15) & MASK64) ^ 0x5BD1E995) & MASK64


class IterRng:
    """Uniform draws of one iteration: Philox4x32-10 keyed by (stage key, iteration)."""

    def __init__(self, key: int, iteration: int):
        # an explicit uint64 array: a list holding an int above 2^63 would
        # become float64 and lose the key's low bits
        self.g = np.random.Generator(np.random.Philox(key=np.array([key & MASK64, iteration & MASK64], np.uint64)))

    def uniform(self, lo: float, hi: float) -> float:   # gen_range(lo..hi)
        return lo + (hi - lo) * float(self.g.random())

    def index(self, n: int) -> int:                     # gen_range(0..n)
        return int(self.g.integers(0, n))

    def u64(self) -> int:
        return int(self.g.integers(0, 1 << 64, dtype=np.uint64))

    def bernoulli(self, p: float) -> bool:              # gen_bool(p)
        return float(self.g.random()) < p


# ------------------------------------------------------------------ rigid moves (src/geometry/mesh_ops.rs)

def centroid(v: np.ndarray) -> np.ndarray:
    """mesh_centroid (mesh_ops.rs:5-16): vertex mean."""
    return v.sum(axis=0) * (1.0 / len(v)) if len(v) else np.zeros(3)


def move_to_center(v: np.ndarray, target) -> np.ndarray:
    """move_mesh_to_target_center (mesh_ops.rs:213-217)."""
    return v + (np.asarray(target, np.float64) - centroid(v))


def rotate_about_center(v: np.ndarray, axis, angle: float) -> np.ndarray:
    """rotate_mesh_around_center (mesh_ops.rs:308-326): Rodrigues about the centroid."""
    axis = np.asarray(axis, np.float64)
    n = math.sqrt(float(axis @ axis))
    if n <= 1e-12:
        return v
    k = axis * (1.0 / n)
    c = centroid(v)
    p = v - c
    ct, st = math.cos(angle), math.sin(angle)
    return c + (p * ct + np.cross(k, p) * st + np.outer(p @ k, k) * (1.0 - ct))


def wrap_centroid(v: np.ndarray, box: BoundingBox) -> np.ndarray:
    """wrap_mesh_centroid_to_box (mesh_ops.rs:225-246)."""
    c = centroid(v)
    lo, size = np.asarray(box.min, np.float64), np.asarray(box.max, np.float64) - np.asarray(box.min, np.float64)
    w = np.where(size > 0, lo + np.mod(c - lo, np.where(size > 0, size, 1.0)), c)
    return v + (w - c)


def boundary_ok(v: np.ndarray, box: BoundingBox, mode: int, d1: float, d2: float) -> bool:
    """check_boundary_constraints_mode (src/geometry/bbox.rs:72-131)."""
    if len(v) == 0:
        return False
    lo = np.asarray(box.min, np.float64)
    size = np.asarray(box.max, np.float64) - lo
    mn, mx = v.min(axis=0) - lo, v.max(axis=0) - lo
    if np.all(mn >= 0.0) and np.all(mx <= size):
        return not (np.any(mn < d1) or np.any(mx > size - d1))
    if mode == 1:
        return False
    for i in range(3):
        if mn[i] < 0.0 and (abs(mn[i]) < d2 or mx[i] < d2):
            return False
        if mx[i] > size[i] and ((size[i] - mn[i]) < d2 or (mx[i] - size[i]) < d2):
            return False
        if mn[i] >= 0.0 and mx[i] <= size[i] and (mn[i] < d1 or mx[i] > size[i] - d1):
            return False
    return True


def parse_rotation_mode(mode: Optional[str], vector) -> Optional[object]:
    """parse_rotation_mode (src/pipeline/rotation.rs:17-47): None, an axis, or "any"."""
    raw = (mode or "any").strip().lower()
    axes = {"x": (1.0, 0.0, 0.0), "y": (0.0, 1.0, 0.0), "z": (0.0, 0.0, 1.0)}
    if raw == "none":
        return None
    if raw in axes:
        return np.array(axes[raw])
    if raw == "vector":
        if vector is None:
            raise cfg.ConfigError("optimization.rotation_axis_vector is required when rotation_mode='vector'")
        if len(vector) != 3:
            raise cfg.ConfigError("optimization.rotation_axis_vector must have length 3")
        a = np.asarray(vector, np.float64)
        if math.sqrt(float(a @ a)) <= 1e-12:
            raise cfg.ConfigError("optimization.rotation_axis_vector must be non-zero")
        return a
    if raw == "any":
        return "any"
    raise cfg.ConfigError(f"optimization.rotation_mode must be one of: none, x, y, z, vector, any (got '{raw}')")


def sample_axis(rng: IterRng, mode):
    """sample_rotation_axis (src/pipeline/rotation.rs:56-65)."""
    if mode is None:
        return None
    if isinstance(mode, str):
        return np.array([rng.uniform(-1.0, 1.0), rng.uniform(-1.0, 1.0), rng.uniform(-1.0, 1.0)])
    return mode


# ------------------------------------------------------------------ backend and evaluator

class S2Method:
    """S2Method::resolve / name (src/pipeline/optimize_execution.rs:14-39)."""
    VOXEL_EXACT, VOXEL_MC, MESH_MC = "voxel_exact", "voxel_mc", "mesh_mc"

    @staticmethod
    def resolve(p: cfg.OptimizationParams) -> str:
        if p.mc_method == "exact":
            return S2Method.VOXEL_EXACT
        return S2Method.MESH_MC if p.voxel_pitch <= 0.0 else S2Method.VOXEL_MC


def select_s2_backend(p: cfg.OptimizationParams, requested: cfg.AccelerationMode, workload: int, faces: int,
                      probe: Callable[[], cfg.BackendSelection]) -> cfg.BackendSelection:
    """select_s2_backend (src/pipeline/optimize_execution.rs:58-134) with a hip
    arm.  cpu never probes; auto below gpu_min_voxels never probes.  The
    reference keeps voxel methods on the CPU because its wgpu evaluator is
    mesh-MC only; the hip evaluator runs every method, so for `hip` (and `auto`
    resolving to it) the voxel methods are eligible and only the budget and
    the probe decide.  `gpu` keeps the reference's rules (voxel methods stay on
    the CPU; mesh MC needs backend wgpu / f32)."""
    acc = p.acceleration
    if requested == cfg.AccelerationMode.CPU:
        return cfg.BackendSelection("cpu")
    if requested == cfg.AccelerationMode.AUTO and workload < acc.gpu_min_voxels:
        return cfg.BackendSelection("cpu", f"workload {workload} voxels below gpu_min_voxels threshold "
                                           f"{acc.gpu_min_voxels}")
    method = S2Method.resolve(p)
    reason = None
    if requested == cfg.AccelerationMode.GPU:
        if method != S2Method.MESH_MC:
            reason = f"{method} has no compatible optimizer GPU evaluator; retaining its CPU definition"
        elif acc.backend != "wgpu" or acc.gpu_precision != "f32" or acc.gpu_prefer_power:
            raise cfg.ConfigError("optimizer GPU supports only backend=wgpu, gpu_precision=f32 and "
                                  "gpu_prefer_power=false")
        else:
            reason = "the wgpu backend is not part of this engine; use acceleration.mode: hip"
    elif acc.gpu_memory_limit_mb is not None:
        need = hip_evaluation_bytes(p, workload, faces)
        budget = acc.gpu_memory_limit_mb * 1024 * 1024
        if need > budget:
            reason = f"optimizer HIP evaluation requires {need} bytes, exceeding budget {budget}"
    sel = cfg.BackendSelection("cpu", reason) if reason else probe()
    if sel.backend != "hip" and not acc.cpu_fallback:
        raise cfg.GpuError(f"optimizer CPU fallback is disabled: {sel.fallback_reason or 'GPU unavailable'}")
    return sel


def hip_evaluation_bytes(p: cfg.OptimizationParams, workload: int, faces: int) -> int:
    """Device working set of the island's evaluator: the seed (packed volume,
    per-voxel coverage counts, counts, geometry) or the resident mesh."""
    if S2Method.resolve(p) == S2Method.MESH_MC:
        return faces * 144 + 16 * (p.r_max + 1) + (1 << 20)
    r = p.r_max / max(p.voxel_pitch, 1e-9) + 1.5
    return workload * (4 + 1) + int(4.19 * r ** 3) * 48 + faces * 200 + (16 << 20)


@dataclass
class IslandResult:
    best_loss: float
    best_s2: np.ndarray
    best_particles: List[np.ndarray]
    accepted: int
    iterations: int
    evaluations: int
    s2_time: float
    collision_time: float
    batches: int


class OptimizeS2Hip:
    """OptimizeS2 (src/pipeline/optimize_execution.rs:142-343) on the HIP engine."""

    def __init__(self, p: cfg.OptimizationParams, bbox: BoundingBox, mesh: Mesh, env=None):
        if not math.isfinite(p.voxel_pitch):
            raise cfg.ConfigError("optimization.voxel_pitch must be finite")
        self.params = p
        self.requested = cfg.configured_mode(p.acceleration, os.environ if env is None else env)
        self.method = S2Method.resolve(p)
        self.pitch = 1.0 if (self.method == S2Method.VOXEL_EXACT and p.voxel_pitch <= 0.0) else p.voxel_pitch
        est_pitch = 1.0 if self.method == S2Method.MESH_MC else max(self.pitch, 1e-9)
        size = np.asarray(bbox.max, np.float64) - np.asarray(bbox.min, np.float64)
        workload = 1
        for side in size:
            workload *= max(int(math.ceil(side / est_pitch)), 1)
        self.ctx = None

        def probe():
            reason = cfg.hip_device_probe()
            return cfg.BackendSelection("cpu", reason) if reason else cfg.BackendSelection("hip")

        sel = select_s2_backend(p, self.requested, workload, len(mesh.faces), probe)
        self.backend = sel.backend
        self.description = (f"S2 execution: requested={self.requested.value}, effective={self.backend}, "
                            f"method={self.method}, voxel_pitch={self.pitch}, reason={sel.fallback_reason or 'none'}")
        if self.backend == "hip":
            budget = (p.acceleration.gpu_memory_limit_mb or 0) * 1024 * 1024
            self.ctx = s2.Context(0, budget)

    def require_hip(self):
        if self.backend != "hip":
            from .pipeline import CpuPathUnavailable
            raise CpuPathUnavailable("optimize: the policy selected the CPU backend, which is the host crate's "
                                     "own implementation and not part of this engine")

    def evaluate(self, mesh: Mesh, bbox: BoundingBox, r_max: int, samples: int, seed: Optional[int]) -> np.ndarray:
        """OptimizeS2::evaluate -> calculate_s2_seeded on the merged mesh."""
        self.require_hip()
        return s2.calculate_s2(self.ctx, mesh, bbox, r_max, self.pitch if self.method != S2Method.MESH_MC
                               else self.params.voxel_pitch,
                               "exact" if self.method == S2Method.VOXEL_EXACT else "monte_carlo", samples,
                               0 if seed is None else seed)

    def close(self):
        if self.ctx is not None:
            self.ctx.close()
            self.ctx = None


# ------------------------------------------------------------------ the island loop

class _Schedule:
    """The temperature schedule and acceptance window of run_sa_island."""

    def __init__(self, p: cfg.OptimizationParams):
        t0 = max(p.initial_temperature, 1e-8)
        self.t0 = t0
        self.temperature = t0
        self.cooling = min(max(p.cooling_rate, 0.8), 0.99999)
        self.window = max(p.adaptive_temp_window if p.adaptive_temp_window is not None
                          else min(max(p.max_iterations // 40, 20), 100), 5)
        self.low = min(max(p.target_acceptance_low if p.target_acceptance_low is not None else 0.20, 0.0), 1.0)
        high = min(max(p.target_acceptance_high if p.target_acceptance_high is not None else 0.45, 0.0), 1.0)
        self.high = min(max(self.low + 0.05, 0.0), 1.0) if high <= self.low else high
        self.heat = max(p.adaptive_heat_factor if p.adaptive_heat_factor is not None else 1.08, 1.0)
        self.cool = min(max(p.adaptive_cool_factor if p.adaptive_cool_factor is not None else 0.94, 0.01), 1.0)
        self.ceiling = t0 * max(p.adaptive_temp_ceiling_factor if p.adaptive_temp_ceiling_factor is not None
                                else 5.0, 1.0)
        self.floor = 1e-9
        self.trials = 0
        self.accepts = 0

    def state(self):
        return (self.temperature, self.trials, self.accepts)

    def restore(self, st):
        self.temperature, self.trials, self.accepts = st

    def step(self, accepted: bool):
        self.temperature = max(self.temperature * self.cooling, self.floor)
        self.trials += 1
        if accepted:
            self.accepts += 1
        if self.trials >= self.window:
            rate = self.accepts / self.trials
            if rate < self.low:
                self.temperature = min(self.temperature * self.heat, self.ceiling)
            elif rate > self.high:
                self.temperature = max(self.temperature * self.cool, self.floor)
            self.trials = 0
            self.accepts = 0


def l2_norm(a, b) -> float:
    return s2.l2_norm(a, b)


def run_sa_island(evaluator: OptimizeS2Hip, particles: List[Mesh], target, box: BoundingBox, history: List[str],
                  island_id: int = 0, mode: int = 1, d1: float = 0.0, d2: float = 0.0, rotation=None,
                  blocked: Optional[Callable] = None, batch: int = 0, log=print) -> IslandResult:
    """run_sa_island (src/pipeline/optimize.rs:359-833) for voxel exact S2 on the
    HIP engine.  `batch` = 0 sizes the speculative runs from the recent
    acceptance rate; a positive value fixes their length (1: sequential)."""
    p = evaluator.params
    evaluator.require_hip()
    if evaluator.method != S2Method.VOXEL_EXACT:
        raise cfg.ConfigError("the batched island loop evaluates voxel exact S2 (mc_method: exact)")
    target = np.asarray(target, np.float64)
    key = stage_key(p.seed, ISLAND_STAGE + island_id)
    sched = _Schedule(p)
    verts = [np.ascontiguousarray(m.vertices, np.float64).copy() for m in particles]
    seed = B.Seed(evaluator.ctx, particles, box, evaluator.pitch, p.r_max)
    try:
        t_s2 = time.perf_counter()
        current_s2 = seed.exact_s2()
        s2_time = time.perf_counter() - t_s2
        current_loss = l2_norm(current_s2, target)
        report.push_history_s2(history, "Post-Pruning S2", current_s2)
        history.append(f"Post-Pruning Loss: {current_loss:.6f}")
        best_loss, best_s2, best_parts = current_loss, current_s2.copy(), [v.copy() for v in verts]
        log(f"[Info] Starting optimization with {len(verts)} particles.")
        log(f"[Info] Initial loss: {current_loss:.6f}")
        bounds_lo, bounds_hi = np.asarray(box.min, np.float64), np.asarray(box.max, np.float64)
        accepted_moves = evaluations = batches = 0
        collision_time = 0.0
        rate = 0.5
        it = 0
        stop = False
        while it < p.max_iterations and not stop:
            # proposals of the iterations up to the next acceptance, assuming none is accepted
            length = batch if batch > 0 else int(min(256, max(8, math.ceil(4.0 / max(rate, 1e-3)))))
            saved = sched.state()
            plan = []   # (iteration, index, new vertices or None, temperature, rng)
            j = it
            while j < p.max_iterations and len(plan) < length:
                rng = IterRng(key, j)
                T = sched.temperature
                t0 = time.perf_counter()
                idx, cand = _propose(rng, verts, p, T, sched.t0, box, bounds_lo, bounds_hi, mode, d1, d2, rotation,
                                     blocked)
                collision_time += time.perf_counter() - t0
                plan.append((j, idx, cand, T, rng))
                sched.step(False)
                j += 1
                if cand is not None and sched.temperature < 1e-9:
                    break   # the loop ends after this iteration unless it is accepted
            sched.restore(saved)
            cands = [(0, idx, cand) for (_, idx, cand, _, _) in plan if cand is not None]
            curves = []
            if cands:
                t0 = time.perf_counter()
                curves = B.batch_eval(evaluator.ctx, [seed], cands)
                s2_time += time.perf_counter() - t0
                batches += 1
            ci = 0
            n_eval = 0
            for (j, idx, cand, T, rng) in plan:
                it = j + 1
                if cand is None:
                    sched.step(False)
                    continue
                curve = curves[ci]
                ci += 1
                n_eval += 1
                loss = l2_norm(curve, target)
                delta = loss - current_loss
                accept = True if delta <= 0.0 else rng.bernoulli(min(max(math.exp(-delta / T), 0.0), 1.0))
                if accept:
                    t0 = time.perf_counter()
                    seed.accept(idx, cand)
                    s2_time += time.perf_counter() - t0
                    verts[idx] = cand
                    accepted_moves += 1
                    current_s2, current_loss = curve, loss
                    if current_loss < best_loss:
                        best_loss, best_s2, best_parts = current_loss, current_s2.copy(), [v.copy() for v in verts]
                        report.push_history_best(history, j, best_loss, best_s2)
                sched.step(accept)
                if sched.temperature < 1e-9:
                    stop = True
                    break
                if accept:
                    break
            evaluations += n_eval
            rate = 0.8 * rate + 0.2 * (1.0 / max(n_eval, 1) if n_eval else rate)
        log(f"[Info] Island {island_id}: {it} iterations, {accepted_moves} accepted, best loss {best_loss:.6f}")
        return IslandResult(best_loss, best_s2, best_parts, accepted_moves, it, evaluations, s2_time,
                            collision_time, batches)
    finally:
        seed.free()


def _propose(rng: IterRng, verts, p, T, t0, box, lo, hi, mode, d1, d2, rotation, blocked):
    """One iteration's move (src/pipeline/optimize.rs:460-640): the particle
    index and its new vertices, or None when a constraint blocks the move."""
    n = len(verts)
    idx = rng.index(n)
    v = verts[idx]
    scale = min(max(T / t0, 0.05), 1.0)
    roll = rng.uniform(0.0, 1.0)
    mt = p.max_translation
    if roll < 0.6:
        trans = np.array([rng.uniform(-mt, mt) * scale for _ in range(3)]) if mt > 0 else np.zeros(3)
        v = move_to_center(v, centroid(v) + trans)
        rot_limit = math.radians(p.max_rotation_deg) * scale
        if rot_limit > 1e-6:
            axis = sample_axis(rng, rotation)
            if axis is not None:
                v = rotate_about_center(v, axis, rng.uniform(-rot_limit, rot_limit))
    elif roll < 0.9:
        other = rng.index(n)
        if other != idx:
            c0, c1 = centroid(v), centroid(verts[other])
            d = c1 - c0
            dist = math.sqrt(float(d @ d))
            if dist > 1e-12:
                step = min(dist * 0.3, mt * scale * 2.0)
                v = move_to_center(v, c0 + d * (step / dist))
        axis = sample_axis(rng, rotation)
        if axis is not None:
            lim = math.radians(10.0) * scale
            v = rotate_about_center(v, axis, rng.uniform(-lim, lim))
    else:
        pos = np.array([rng.uniform(lo[k], hi[k]) for k in range(3)])
        v = move_to_center(v, pos)
        axis = sample_axis(rng, rotation)
        if axis is not None:
            v = rotate_about_center(v, axis, rng.uniform(0.0, 2.0 * math.pi))
    if mode == 3:
        v = wrap_centroid(v, box)
    if not boundary_ok(v, box, mode, d1, d2):
        return idx, None
    if blocked is not None and blocked(v, idx, verts):
        return idx, None
    return idx, np.ascontiguousarray(v)


# ------------------------------------------------------------------ the pipeline

class OptimizePipeline:
    """OptimizePipeline::run_in_pool (src/pipeline/optimize.rs:870-1062) for a
    single island on the HIP engine: load and split the input, drop particles
    outside the box, target S2 (manual_array or reference_stl), input S2, the
    island loop, the final S2 of the best packing, output STL and
    s2_history.txt."""

    def __init__(self, conf: cfg.OptimizationConfig, env=None, out=sys.stdout, blocked=None, batch: int = 0):
        self.c = conf
        self.env = os.environ if env is None else env
        self.out = out
        self.blocked = blocked
        self.batch = batch

    def log(self, msg):
        print(msg, file=self.out)

    def run(self) -> IslandResult:
        c, p = self.c, self.c.optimization
        timer = report.StageTimer("optimize", self.out)
        box = cfg.parse_box_dimensions(c.box_dimensions)
        mode = p.mode if p.mode is not None else 1
        d1 = p.min_boundary_dist or 0.0
        d2 = p.min_cross_boundary_depth or 0.0
        rotation = parse_rotation_mode(p.rotation_mode, p.rotation_axis_vector)
        self.log(f"[Info] Rotation mode: {p.rotation_mode or 'any'}")
        if (p.islands or 1) > 1:
            raise cfg.ConfigError("several islands with migration are not part of this engine; use islands: 1")
        mesh = load_stl(c.input_stl)
        gid, ng = granules(mesh)
        particles = []
        for g in range(ng):
            faces = mesh.faces[gid == g]
            used, inv = np.unique(faces.reshape(-1), return_inverse=True)
            particles.append(Mesh(mesh.vertices[used], inv.reshape(-1, 3).astype(np.uint32)))
        loaded = len(particles)
        lo, hi = np.asarray(box.min), np.asarray(box.max)
        particles = [m for m in particles
                     if np.all(m.vertices.min(axis=0) < hi) and np.all(m.vertices.max(axis=0) > lo)]
        if loaded > len(particles):
            self.log(f"[Info] Pre-filter removed {loaded - len(particles)} particles fully outside optimization bbox.")
        if not particles:
            raise cfg.ConfigError("No particles remain after bbox pre-filter for optimization")
        timer.stage("load")
        merged = merge_meshes(particles)
        ev = OptimizeS2Hip(p, box, merged, self.env)
        try:
            timer.stage("prepare_evaluator")
            self.log(f"[Info] {ev.description}")
            if c.target.type == "manual_array":
                if c.target.s2_array is None:
                    raise cfg.ConfigError("target.s2_array is required")
                target = np.asarray(c.target.s2_array, np.float64)
            elif c.target.type == "reference_stl":
                if not c.target.stl_path:
                    raise cfg.ConfigError("target.stl_path is required")
                ref = load_stl(c.target.stl_path)
                bb = cfg.parse_box_dimensions(c.target.stl_bounding_box) if c.target.stl_bounding_box else \
                    (mesh_bbox(ref) or BoundingBox.from_size((1.0, 1.0, 1.0)))
                target = ev.evaluate(ref, bb, p.r_max, max(p.mc_samples, 1000), fixed_eval_seed(p.seed, TARGET_STAGE))
            else:
                raise cfg.ConfigError(f"Unknown target type: {c.target.type}")
            timer.stage("target_s2")
            history = [ev.description]
            self.log(f"[Info] S2 config: method={ev.method}, r_max={p.r_max}, mc_samples={p.mc_samples}, "
                     f"voxel_pitch={p.voxel_pitch:.6f}")
            input_s2 = ev.evaluate(merged, box, p.r_max, max(p.mc_samples, 1000), fixed_eval_seed(p.seed, INPUT_STAGE))
            input_vf = s2.geometric_volume_fraction(merged, box)
            input_loss = l2_norm(input_s2, target)
            self.log(f"[Info] Input stage: VF {input_vf:.6f}, Loss {input_loss:.6f}, Method {ev.method}")
            report.push_history_s2(history, "Target S2", target)
            report.push_history_s2(history, "Input S2", input_s2)
            history.append(f"Input Loss: {input_loss:.6f}")
            timer.stage("input_s2")
            if p.prune_enabled:
                self.log("[Info] Selective pruning is not part of this engine; the input population is kept.")
            timer.stage("prune")
            history.append("Execution: workers=1, islands=1, active_island_limit=1")
            res = run_sa_island(ev, particles, target, box, history, 0, mode, d1, d2, rotation, self.blocked,
                                self.batch, self.log)
            timer.stage("anneal")
            best_mesh = merge_meshes([Mesh(v, m.faces) for v, m in zip(res.best_particles, particles)])
            best_s2 = ev.evaluate(best_mesh, box, p.r_max, max(p.mc_samples, 2000), fixed_eval_seed(p.seed, FINAL_STAGE))
            best_loss = l2_norm(best_s2, target)
            history.append(f"Selected Search Loss: {res.best_loss:.6f}")
            report.push_history_s2(history, "Final Best S2", best_s2)
            history.append(f"Final Best Loss: {best_loss:.6f}")
            timer.stage("final_s2")
            out = Path(c.output_path)
            out.parent.mkdir(parents=True, exist_ok=True)
            save_stl(out, best_mesh, "optimized_structure")
            timer.stage("write_stl")
            hist_path = out.parent / "s2_history.txt"
            report.write_history(hist_path, history)
            timer.stage("write_history")
            self.log("[Info] Optimization completed.")
            self.log(f"[Info] Best loss: {best_loss:.6f}")
            self.log(f"[Info] S2 history saved: {hist_path}")
            self.log(f"[Info] Timing summary: S2 {res.s2_time:.2f}s | collision/constraints {res.collision_time:.2f}s"
                     f" | {res.evaluations} candidates in {res.batches} batches")
            timer.total("total_in_pool")
            res.best_s2 = best_s2
            res.best_loss = best_loss
            return res
        finally:
            ev.close()
