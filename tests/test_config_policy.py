"""YAML surface and backend policy (CPU only): `hip` joins auto/cpu/gpu in both
the YAML enum and the environment override; the decision matrix follows
src/compute/policy.rs:158-218 with an injected probe (the reference's own
forcing pattern, src/pipeline/optimize_execution.rs:58-64)."""
import json

import pytest

from rustmspt_amd import config as cfg


def acc(**kw):
    return cfg.AccelerationConfig.parse(kw)


def test_defaults_match_reference():
    a = cfg.AccelerationConfig.parse({})
    assert (a.mode, a.backend, a.cpu_fallback, a.gpu_min_voxels, a.gpu_memory_limit_mb, a.gpu_precision) == \
        (cfg.AccelerationMode.AUTO, "wgpu", True, 250000, None, "f32")
    with pytest.raises(cfg.ConfigError):
        cfg.AccelerationConfig.parse({"mode": "cuda"})
    assert cfg.AccelerationConfig.parse({"mode": "hip"}).mode == cfg.AccelerationMode.HIP


def test_measurement_yaml_and_flexible_ints(tmp_path):
    p = tmp_path / "m.yaml"
    p.write_text("""measurement:
  stl_path: "a.stl"
  stl_bounding_box: [0.0, 0.0, 0.0, 50.0, 50.0, 50.0]
  r_max: 20
  voxel_pitch: 1.0
  mc_method: 'both'
  mc_samples: 40_000
  cpu_max: -1
  output_path: "out.txt"
  acceleration:
    mode: hip
    gpu_memory_limit_mb: "1_024"
""")
    m = cfg.load_measurement_config(p)
    assert m.mc_samples == 40000 and m.cpu_max == -1 and m.acceleration.mode == cfg.AccelerationMode.HIP
    assert m.acceleration.gpu_memory_limit_mb == 1024
    j = tmp_path / "m.json"
    j.write_text(json.dumps({"measurement": {"stl_path": "a", "output_path": "b", "bounding_box": [4, 4, 4],
                                             "r_max": 2, "voxel_pitch": 0.0, "mc_method": "exact"}}))
    assert cfg.load_measurement_config(j).bounding_box == [4, 4, 4]
    assert cfg.parse_box_dimensions([1, 2, 3]).max == (1.0, 2.0, 3.0)
    with pytest.raises(cfg.ConfigError):
        cfg.parse_box_dimensions([1, 2])


def test_env_override_accepts_hip_and_rejects_unknown():
    a = acc(mode="cpu")
    assert cfg.configured_mode(a, {}) == cfg.AccelerationMode.CPU
    for v in ("cpu", "gpu", "auto", "hip"):
        assert cfg.configured_mode(a, {"RUSTMSPT_ACCELERATION": v}).value == v
    with pytest.raises(cfg.ConfigError):
        cfg.configured_mode(a, {"RUSTMSPT_ACCELERATION": "invalid"})


def never_probe():
    raise AssertionError("probed")


def test_decision_matrix():
    M = cfg.AccelerationMode
    ok = lambda: None  # noqa: E731
    down = lambda: "HIP init failed: no device"  # noqa: E731
    # cpu and auto-below-threshold never probe
    assert cfg.resolve_execution(acc(), M.CPU, 10 ** 9, 1, True, 1, never_probe).backend == "cpu"
    assert cfg.resolve_execution(acc(), M.AUTO, 10, 100, True, 1, never_probe).backend == "cpu"
    # hip selected, auto above threshold resolves to hip
    assert cfg.resolve_execution(acc(), M.HIP, 10, 100, True, 1, ok).backend == "hip"
    assert cfg.resolve_execution(acc(), M.AUTO, 1000, 100, True, 1, ok).backend == "hip"
    # unavailable device: fallback when allowed, error otherwise
    sel = cfg.resolve_execution(acc(), M.HIP, 10, 0, True, 1, down)
    assert sel.backend == "cpu" and "no device" in sel.fallback_reason
    with pytest.raises(cfg.GpuError):
        cfg.resolve_execution(acc(cpu_fallback=False), M.HIP, 10, 0, True, 1, down)
    # budget gate before the probe: 0 MiB cannot hold the working set
    sel = cfg.resolve_execution(acc(gpu_memory_limit_mb=0), M.HIP, 10, 0, True, 4096, never_probe)
    assert sel.backend == "cpu" and "exceeding budget" in sel.fallback_reason
    with pytest.raises(cfg.GpuError):
        cfg.resolve_execution(acc(gpu_memory_limit_mb=0, cpu_fallback=False), M.HIP, 10, 0, True, 4096, never_probe)
    assert cfg.resolve_execution(acc(gpu_memory_limit_mb=1), M.HIP, 10, 0, True, None, never_probe).backend == "cpu"
    # the wgpu backend is not part of this engine; its config checks still apply
    with pytest.raises(cfg.ConfigError):
        cfg.resolve_execution(acc(gpu_precision="f64"), M.GPU, 10, 0, True, 1, ok)
    assert cfg.resolve_execution(acc(), M.GPU, 10, 0, True, 1, ok).backend == "cpu"
    # a method without a device path
    assert cfg.resolve_execution(acc(), M.HIP, 10, 0, False, 1, ok).backend == "cpu"


def test_smooth_length_matches_reference_cases():
    assert [cfg.smooth_fft_length(2 * n - 1) for n in (7, 11, 13)] == [15, 24, 25]
    assert cfg.smooth_fft_length(79) == 80 and cfg.smooth_fft_length(511) == 512 and cfg.smooth_fft_length(1535) == 1536
