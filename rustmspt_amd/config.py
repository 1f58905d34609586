"""YAML surface of the S2 pipelines and the backend policy, with `hip` added to
AccelerationMode.

Mirrors src/config/{acceleration,measurement,optimization}.rs (field names and
defaults), src/config/deserialize.rs (integers written as "40_000"),
src/compute/backend.rs (AccelerationMode, lowercase names) and
src/compute/policy.rs:138-218 (configured_mode, resolve_execution).
"""
from __future__ import annotations

import enum
import os
from dataclasses import dataclass, field, fields
from typing import Optional

import yaml


class ConfigError(ValueError):
    """RustMsptError::InvalidConfig."""


class GpuError(RuntimeError):
    """RustMsptError::Gpu (no CPU fallback allowed)."""


class AccelerationMode(str, enum.Enum):
    AUTO = "auto"
    CPU = "cpu"
    GPU = "gpu"
    HIP = "hip"


def _flex_int(v, name):
    """src/config/deserialize.rs:3-10 parse_usize_like: ints or strings with '_'."""
    if v is None:
        return None
    if isinstance(v, bool):
        raise ConfigError(f"invalid integer value for {name}: {v!r}")
    if isinstance(v, int):
        return v
    if isinstance(v, str):
        try:
            return int(v.replace("_", ""))
        except ValueError as e:
            raise ConfigError(f"invalid usize value '{v}': {e}") from None
    raise ConfigError(f"invalid integer value for {name}: {v!r}")


@dataclass
class AccelerationConfig:
    """src/config/acceleration.rs:5-57 (defaults: auto, wgpu, fallback on,
    250000 voxels/pixels, no budget, f32)."""
    mode: AccelerationMode = AccelerationMode.AUTO
    backend: str = "wgpu"
    cpu_fallback: bool = True
    gpu_min_voxels: int = 250_000
    gpu_min_pixels: int = 250_000
    gpu_memory_limit_mb: Optional[int] = None
    gpu_prefer_power: bool = False
    gpu_precision: str = "f32"

    @staticmethod
    def parse(d) -> "AccelerationConfig":
        d = dict(d or {})
        known = {f.name for f in fields(AccelerationConfig)}
        extra = set(d) - known
        if extra:
            raise ConfigError(f"unknown acceleration field(s): {sorted(extra)}")
        out = AccelerationConfig()
        if "mode" in d:
            try:
                out.mode = AccelerationMode(str(d["mode"]))
            except ValueError:
                raise ConfigError(f"unknown variant `{d['mode']}`, expected one of `auto`, `cpu`, `gpu`, `hip`")
        for k in ("backend", "gpu_precision"):
            if k in d:
                out.__dict__[k] = str(d[k])
        for k in ("cpu_fallback", "gpu_prefer_power"):
            if k in d:
                out.__dict__[k] = bool(d[k])
        for k in ("gpu_min_voxels", "gpu_min_pixels", "gpu_memory_limit_mb"):
            if k in d:
                out.__dict__[k] = _flex_int(d[k], k)
        return out


@dataclass
class MeasurementParams:
    """src/config/measurement.rs:6-20."""
    stl_path: str
    r_max: int
    voxel_pitch: float
    mc_method: str
    output_path: str
    bounding_box: Optional[list] = None
    stl_bounding_box: Optional[list] = None
    mc_samples: Optional[int] = None
    cpu_max: Optional[int] = None
    acceleration: AccelerationConfig = field(default_factory=AccelerationConfig)

    @staticmethod
    def parse(d) -> "MeasurementParams":
        d = dict(d)
        for req in ("stl_path", "r_max", "voxel_pitch", "mc_method", "output_path"):
            if req not in d:
                raise ConfigError(f"missing field `{req}`")
        return MeasurementParams(
            stl_path=str(d["stl_path"]), r_max=_flex_int(d["r_max"], "r_max"),
            voxel_pitch=float(d["voxel_pitch"]), mc_method=str(d["mc_method"]), output_path=str(d["output_path"]),
            bounding_box=d.get("bounding_box"), stl_bounding_box=d.get("stl_bounding_box"),
            mc_samples=_flex_int(d.get("mc_samples"), "mc_samples"), cpu_max=_flex_int(d.get("cpu_max"), "cpu_max"),
            acceleration=AccelerationConfig.parse(d.get("acceleration")))


def load_measurement_config(path) -> MeasurementParams:
    """src/config/mod.rs:53-56 load_yaml (JSON is valid YAML)."""
    with open(path) as fh:
        doc = yaml.safe_load(fh)
    if not isinstance(doc, dict) or "measurement" not in doc:
        raise ConfigError("missing field `measurement`")
    return MeasurementParams.parse(doc["measurement"])


@dataclass
class TargetConfig:
    """src/config/optimization.rs:5-11."""
    type: str
    s2_array: Optional[list] = None
    stl_path: Optional[str] = None
    stl_bounding_box: Optional[list] = None


@dataclass
class OptimizationParams:
    """src/config/optimization.rs:13-59: field names and optional fields as the
    reference's; integers accept the "5_000" spelling."""
    max_iterations: int
    initial_temperature: float
    cooling_rate: float
    r_max: int
    voxel_pitch: float
    mc_method: str
    mc_samples: int
    max_translation: float
    max_rotation_deg: float
    seed: Optional[int] = None
    adaptive_temp_window: Optional[int] = None
    target_acceptance_low: Optional[float] = None
    target_acceptance_high: Optional[float] = None
    adaptive_heat_factor: Optional[float] = None
    adaptive_cool_factor: Optional[float] = None
    adaptive_temp_ceiling_factor: Optional[float] = None
    min_neighbor_distance: Optional[float] = None
    mode: Optional[int] = None
    min_boundary_dist: Optional[float] = None
    min_cross_boundary_depth: Optional[float] = None
    prune_enabled: Optional[bool] = None
    prune_tolerance: Optional[float] = None
    prune_max_rounds: Optional[int] = None
    prune_eval_samples: Optional[int] = None
    cpu_max: Optional[int] = None
    orient_to_positive_volume: Optional[bool] = None
    rotation_mode: Optional[str] = None
    rotation_axis_vector: Optional[list] = None
    islands: Optional[int] = None
    migration_interval: Optional[int] = None
    acceleration: AccelerationConfig = field(default_factory=AccelerationConfig)

    _REQUIRED = ("max_iterations", "initial_temperature", "cooling_rate", "r_max", "voxel_pitch", "mc_method",
                 "mc_samples", "max_translation", "max_rotation_deg")
    _INTS = ("max_iterations", "r_max", "mc_samples", "adaptive_temp_window", "prune_max_rounds",
             "prune_eval_samples", "cpu_max", "islands", "migration_interval", "mode", "seed")
    _FLOATS = ("initial_temperature", "cooling_rate", "voxel_pitch", "max_translation", "max_rotation_deg",
               "target_acceptance_low", "target_acceptance_high", "adaptive_heat_factor", "adaptive_cool_factor",
               "adaptive_temp_ceiling_factor", "min_neighbor_distance", "min_boundary_dist",
               "min_cross_boundary_depth", "prune_tolerance")

    @staticmethod
    def parse(d) -> "OptimizationParams":
        d = dict(d)
        for req in OptimizationParams._REQUIRED:
            if req not in d:
                raise ConfigError(f"missing field `{req}`")
        known = {f.name for f in fields(OptimizationParams)}
        kw = {}
        for k, v in d.items():
            if k not in known:   # serde's default: unknown fields are ignored
                continue
            if k == "acceleration":
                kw[k] = AccelerationConfig.parse(v)
            elif v is None:
                kw[k] = None
            elif k in OptimizationParams._INTS:
                kw[k] = _flex_int(v, k)
            elif k in OptimizationParams._FLOATS:
                kw[k] = float(v)
            elif k in ("prune_enabled", "orient_to_positive_volume"):
                kw[k] = bool(v)
            elif k == "mc_method" or k == "rotation_mode":
                kw[k] = str(v)
            else:
                kw[k] = v
        if kw.get("mode") is not None and not 0 <= kw["mode"] <= 255:
            raise ConfigError("optimization.mode must fit u8")
        if kw.get("seed") is not None and kw["seed"] < 0:
            raise ConfigError("optimization.seed must be a u64")
        return OptimizationParams(**kw)


@dataclass
class OptimizationConfig:
    """src/config/optimization.rs:61-68: input.stl_path, output.path, target,
    box.dimensions, optimization."""
    input_stl: str
    output_path: str
    target: TargetConfig
    box_dimensions: list
    optimization: OptimizationParams


def load_optimization_config(path) -> OptimizationConfig:
    with open(path) as fh:
        doc = yaml.safe_load(fh)
    if not isinstance(doc, dict):
        raise ConfigError("optimization config must be a mapping")
    for k in ("input", "output", "target", "box", "optimization"):
        if k not in doc:
            raise ConfigError(f"missing field `{k}`")
    t = dict(doc["target"])
    if "type" not in t:
        raise ConfigError("missing field `type`")
    target = TargetConfig(str(t["type"]), t.get("s2_array"), t.get("stl_path"), t.get("stl_bounding_box"))
    return OptimizationConfig(str(doc["input"]["stl_path"]), str(doc["output"]["path"]), target,
                              list(doc["box"]["dimensions"]), OptimizationParams.parse(doc["optimization"]))


def parse_box_dimensions(v):
    """src/config/mod.rs:64-77: [sx, sy, sz] or [min.., max..]."""
    from .mesh import BoundingBox
    if len(v) == 3:
        return BoundingBox((0.0, 0.0, 0.0), tuple(float(x) for x in v))
    if len(v) == 6:
        return BoundingBox(tuple(float(x) for x in v[:3]), tuple(float(x) for x in v[3:]))
    raise ConfigError("box.dimensions must have length 3 or 6")


def configured_mode(acc: AccelerationConfig, env=None) -> AccelerationMode:
    """src/compute/policy.rs:138-155 (and optimize's resolve_mode,
    src/pipeline/optimize_execution.rs:42-55): RUSTMSPT_ACCELERATION overrides
    YAML; `hip` is accepted by both."""
    env = os.environ if env is None else env
    v = env.get("RUSTMSPT_ACCELERATION")
    if v is None:
        return acc.mode
    if v in ("cpu", "gpu", "auto", "hip"):
        return AccelerationMode(v)
    raise ConfigError(f"invalid RUSTMSPT_ACCELERATION: {v}")


@dataclass
class BackendSelection:
    backend: str                 # "cpu" or "hip"
    fallback_reason: Optional[str] = None


def hip_device_probe() -> Optional[str]:
    """None when a HIP device and the engine are usable, else the reason."""
    try:
        from .s2 import Context
        ctx = Context(0)
        ctx.close()
        return None
    except Exception as e:  # library missing or no device
        return f"HIP init failed: {e}"


def resolve_execution(acc: AccelerationConfig, requested: AccelerationMode, workload: int, threshold: int,
                      supports_gpu: bool, estimated_bytes: Optional[int], probe=hip_device_probe) -> BackendSelection:
    """src/compute/policy.rs:158-218 with a `hip` branch.  cpu (or auto below
    the threshold) never probes.  `hip` has no backend/precision restriction
    (the wgpu-only check applies to `gpu`); the budget gate and the cpu_fallback
    policy are the reference's.  `auto` resolves to hip when a device is present."""
    if requested == AccelerationMode.CPU or (requested == AccelerationMode.AUTO and workload < threshold):
        return BackendSelection("cpu")

    def failure(reason):
        if acc.cpu_fallback:
            return BackendSelection("cpu", reason)
        raise GpuError(reason)

    if not supports_gpu:
        return failure("selected method has no supported GPU execution path")
    if requested == AccelerationMode.GPU:
        if acc.backend != "wgpu" or acc.gpu_precision != "f32" or acc.gpu_prefer_power:
            raise ConfigError("GPU currently requires backend=wgpu, gpu_precision=f32 and gpu_prefer_power=false")
        return failure("the wgpu backend is not part of this engine; use acceleration.mode: hip")
    if acc.gpu_memory_limit_mb is not None:
        budget = acc.gpu_memory_limit_mb * 1024 * 1024
        if estimated_bytes is None:
            return failure("GPU task has no validated working-set estimate")
        if estimated_bytes > budget:
            return failure(f"GPU task requires {estimated_bytes} bytes, exceeding budget {budget}")
    reason = probe()
    if reason is not None:
        return failure(reason)
    return BackendSelection("hip")


def smooth_fft_length(n: int) -> int:
    """src/geometry/s2.rs:519-551: smallest 2,3,5-smooth length >= n."""
    target, best, p5 = max(n, 1), None, 1
    while True:
        p35 = p5
        while True:
            v = p35
            while v < target:
                v *= 2
            best = v if best is None or v < best else best
            if p35 >= target:
                break
            p35 *= 3
        if p5 >= target:
            break
        p5 *= 5
    return best


def hip_exact_bytes(dims, r_max: int, pitch: float, faces: int, fft: bool = False) -> int:
    """Device working set of one exact (or FFT32) evaluation: packed volume,
    triangle data, shell tables, count table and curve buffers (plus the pruned
    FFT buffers for FFT32)."""
    from .s2 import row_words, shell_count
    nx, ny, nz = dims
    packed = nx * ny * row_words(nz) * 8 + 512
    k = shell_count(dims, r_max, pitch)
    est = packed + faces * 136 + k * (12 + 4 + 8 + 8) + k * 8 // 2 + (r_max + 2) * 32
    if fft:
        f = [smooth_fft_length(max(2 * n - 1, 1)) for n in dims]
        h = f[2] // 2 + 1
        r = [min(n - 1, int(r_max / pitch + 0.5 / pitch) + 1) for n in dims]
        est += nx * f[1] * h * 8 + (2 * r[0] + 1) * f[1] * h * 8 + (2 * r[0] + 1) * (2 * r[1] + 1) * h * 8
        est += (2 * r[0] + 1) * (2 * r[1] + 1) * (2 * r[2] + 1) * 4 + k * 4
    return est
