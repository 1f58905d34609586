"""optimize on the HIP engine: the YAML surface, the S2 backend gates, and a
seeded single-island run whose batched candidate evaluation is checked against
the sequential loop and against the oracle (after the reference's
tests/optimize_execution_tests.rs:48-89, 224-240)."""
import json
import math

import numpy as np
import pytest

from rustmspt_amd import config as cfg
from rustmspt_amd import mesh as M
from rustmspt_amd import optimize as O
from rustmspt_amd import synth

REFERENCE_YAML = """
input: {stl_path: "in.stl"}
target: {type: 'manual_array', s2_array: [0.063074, 0.051980, 0.041610], stl_path: "p.stl", stl_bounding_box: []}
box: {dimensions: [50.0, 50.0, 50.0]}
optimization:
  max_iterations: 5_000
  initial_temperature: 0.1
  cooling_rate: 0.99
  adaptive_temp_window: 50
  target_acceptance_low: 0.20
  target_acceptance_high: 0.45
  adaptive_heat_factor: 1.08
  adaptive_cool_factor: 0.94
  adaptive_temp_ceiling_factor: 5.0
  r_max: 10
  voxel_pitch: 1.0
  mc_method: 'monte_carlo'
  mc_samples: 8_000
  max_translation: 25.0
  max_rotation_deg: 30.0
  rotation_mode: 'none'
  rotation_axis_vector: [0.0, 0.0, 1.0]
  min_neighbor_distance: 1.0
  mode: 2
  min_boundary_dist: 1.5
  min_cross_boundary_depth: 3.0
  prune_enabled: true
  prune_tolerance: 0.01
  prune_max_rounds: 120
  prune_eval_samples: 1000
  cpu_max: -1
  orient_to_positive_volume: false
  acceleration:
    mode: auto
output: {path: "out.stl"}
"""


def test_reference_config_parses(tmp_path):
    # the layout of the reference's data/input/optimize_config.yaml
    (tmp_path / "c.yaml").write_text(REFERENCE_YAML)
    c = cfg.load_optimization_config(tmp_path / "c.yaml")
    p = c.optimization
    assert (p.max_iterations, p.mc_samples, p.r_max, p.mode) == (5000, 8000, 10, 2)
    assert p.seed is None and p.islands is None and p.acceleration.mode == cfg.AccelerationMode.AUTO
    assert c.target.type == "manual_array" and c.target.s2_array[0] == 0.063074
    assert c.box_dimensions == [50.0, 50.0, 50.0] and c.output_path == "out.stl"
    with pytest.raises(cfg.ConfigError):
        cfg.OptimizationParams.parse({"max_iterations": 1})
    q = cfg.OptimizationParams.parse(dict(params(), seed=7, acceleration={"mode": "hip"}, unknown_field=3))
    assert q.seed == 7 and q.acceleration.mode == cfg.AccelerationMode.HIP


def params(method="exact", pitch=1.0, **acc):
    return {"max_iterations": 8, "initial_temperature": 0.1, "cooling_rate": 0.99, "r_max": 0,
            "voxel_pitch": pitch, "mc_method": method, "mc_samples": 200, "max_translation": 0.0,
            "max_rotation_deg": 0.0, "prune_enabled": False, "acceleration": dict({"mode": "cpu"}, **acc)}


def test_backend_gates():
    """backend_gates_preserve_methods_without_gpu_probe
    (src/pipeline/optimize_execution.rs:388-447) with the hip arm: cpu and small
    auto never probe; gpu keeps voxel methods on the CPU (and refuses it when
    fallback is off); hip runs every method on the device."""
    def no_probe():
        raise AssertionError("probed")

    hip = lambda: cfg.BackendSelection("hip")
    for method in ("exact", "monte_carlo"):
        for pitch in (-1.0, 0.0, 1.0):
            p = cfg.OptimizationParams.parse(params(method, pitch, cpu_fallback=False))
            assert O.select_s2_backend(p, cfg.AccelerationMode.CPU, 10 ** 12, 0, no_probe).backend == "cpu"
            assert O.select_s2_backend(p, cfg.AccelerationMode.AUTO, 1, 0, no_probe).backend == "cpu"
            voxel = O.S2Method.resolve(p) != O.S2Method.MESH_MC
            if voxel:
                with pytest.raises(cfg.GpuError):
                    O.select_s2_backend(p, cfg.AccelerationMode.GPU, 10 ** 12, 0, no_probe)
            assert O.select_s2_backend(p, cfg.AccelerationMode.HIP, 10 ** 9, 100, hip).backend == "hip"
            assert O.select_s2_backend(p, cfg.AccelerationMode.AUTO, 10 ** 9, 100, hip).backend == "hip"
            p.acceleration.cpu_fallback = True
            if voxel:
                sel = O.select_s2_backend(p, cfg.AccelerationMode.GPU, 10 ** 12, 0, no_probe)
                assert sel.backend == "cpu" and "no compatible optimizer GPU evaluator" in sel.fallback_reason
    # a budget the evaluation does not fit: fallback reason, or an error without fallback
    p = cfg.OptimizationParams.parse(params("exact", 1.0, gpu_memory_limit_mb=0))
    sel = O.select_s2_backend(p, cfg.AccelerationMode.HIP, 10 ** 6, 10, no_probe)
    assert sel.backend == "cpu" and "exceeding budget" in sel.fallback_reason
    p.acceleration.cpu_fallback = False
    with pytest.raises(cfg.GpuError):
        O.select_s2_backend(p, cfg.AccelerationMode.HIP, 10 ** 6, 10, no_probe)


def test_moves_and_constraints():
    rng = O.IterRng(O.stage_key(5, O.ISLAND_STAGE), 3)
    again = O.IterRng(O.stage_key(5, O.ISLAND_STAGE), 3)
    assert [rng.uniform(0, 1) for _ in range(4)] == [again.uniform(0, 1) for _ in range(4)]
    # keys one bit apart (above 2^63) give different streams
    k = O.stage_key(42, O.ISLAND_STAGE)
    assert k >= 1 << 63 and O.IterRng(k, 0).uniform(0, 1) != O.IterRng(k ^ 1, 0).uniform(0, 1)
    assert O.fixed_eval_seed(None, 2) is None and O.fixed_eval_seed(1, 2) != O.fixed_eval_seed(1, 3)
    v = M.icosphere_mesh(np.array([5.0, 5.0, 5.0]), 1.0, 1).vertices
    r = O.rotate_about_center(v, [0.0, 0.0, 2.0], math.pi / 2)
    assert np.allclose(O.centroid(r), O.centroid(v)) and np.allclose(r[:, 2], v[:, 2])
    assert np.allclose(r[:, 0] - 5.0, -(v[:, 1] - 5.0))
    box = M.BoundingBox((0.0, 0.0, 0.0), (10.0, 10.0, 10.0))
    assert O.boundary_ok(v, box, 1, 0.0, 0.0) and not O.boundary_ok(v, box, 1, 4.5, 0.0)
    out = v + np.array([5.5, 0.0, 0.0])   # x in [9.5, 11.5]: 0.5 inside x = 10, 1.5 beyond
    assert not O.boundary_ok(out, box, 1, 0.0, 0.0)
    assert O.boundary_ok(out, box, 2, 0.0, 0.4) and not O.boundary_ok(out, box, 2, 0.0, 0.6)
    w = O.wrap_centroid(v + np.array([12.0, 0.0, 0.0]), box)
    assert np.allclose(O.centroid(w), [7.0, 5.0, 5.0])


# ------------------------------------------------------------------ on the GPU

def write_case(tmp, seed, iters=60, batch_target=None):
    rng = synth.rng_for(31, 1)
    parts = synth.place_particles(12, 40, rng, synth.ellipsoid_maker(3.5, 5.0), gap=0.5)
    M.save_stl(tmp / "input.stl", synth.packing_mesh(parts))
    # the target: the same particles shifted by up to 2.5 voxels per axis, so
    # moves that lower the loss exist and each seed finds its own best packing
    trng = synth.rng_for(32, 1)
    tparts = [p.moved(trng.uniform(-2.5, 2.5, 3), np.eye(3)) for p in parts]
    M.save_stl(tmp / "target.stl", synth.packing_mesh(tparts))
    conf = {"input": {"stl_path": str(tmp / "input.stl")}, "output": {"path": str(tmp / "out" / "output.stl")},
            "box": {"dimensions": [0, 0, 0, 40, 40, 40]},
            "target": {"type": "reference_stl", "stl_path": str(tmp / "target.stl"),
                       "stl_bounding_box": [0, 0, 0, 40, 40, 40]},
            "optimization": {"seed": seed, "max_iterations": iters, "initial_temperature": 1e-4,
                             "cooling_rate": 0.97, "r_max": 8, "voxel_pitch": 1.0, "mc_method": "exact",
                             "mc_samples": 1500, "max_translation": 3.0, "max_rotation_deg": 25.0,
                             "rotation_mode": "any", "prune_enabled": False, "islands": 1,
                             "acceleration": {"mode": "hip"}}}
    (tmp / "c.json").write_text(json.dumps(conf))
    return cfg.load_optimization_config(tmp / "c.json")


@pytest.mark.gpu
def test_seeded_island_final_curve_equals_oracle_and_is_reproducible(tmp_path):
    from oracle import oracle as o
    runs = {}
    for name, seed, batch in (("a", 42, 0), ("again", 42, 0), ("sequential", 42, 1), ("b32", 42, 32),
                              ("other", 43, 0)):
        d = tmp_path / name
        d.mkdir()
        conf = write_case(d, seed)
        res = O.OptimizePipeline(conf, env={}, batch=batch).run()
        hist = (d / "out" / "s2_history.txt").read_text()
        runs[name] = (res, hist, (d / "out" / "output.stl").read_bytes())
    res, hist, stl = runs["a"]
    assert "effective=hip, method=voxel_exact" in hist
    assert res.accepted > 0 and res.evaluations > res.accepted
    # the same seed gives the same run, whatever the batch length (1 = the sequential loop)
    for name in ("again", "sequential", "b32"):
        assert runs[name][1] == hist and runs[name][2] == stl, name
    assert runs["sequential"][0].batches == runs["sequential"][0].evaluations
    assert runs["a"][0].batches < runs["a"][0].evaluations
    assert runs["other"][2] != stl
    # Final Best S2: the exact curve of the best packing (f64 vertices), bit for bit,
    # and of the saved STL (f32 vertices) to 1e-6, as the reference's test checks
    final = [float(x) for x in next(l for l in hist.splitlines()
                                    if l.startswith("Final Best S2: ")).split(": ")[1].split()]
    best = M.merge_meshes([M.Mesh(v, f) for v, f in zip(res.best_particles, _faces(tmp_path / "a"))])
    occ = o.voxelize(best.vertices, best.faces, [0, 0, 0], 1.0, [40, 40, 40])
    want = o.exact_s2(occ, 8)
    assert np.array_equal(res.best_s2, want)
    assert np.max(np.abs(np.array(final) - want)) <= 5e-7 + 1e-12   # printed with six decimals
    saved = M.load_stl(tmp_path / "a" / "out" / "output.stl")
    occ2 = o.voxelize(saved.vertices, saved.faces, [0, 0, 0], 1.0, [40, 40, 40])
    assert np.max(np.abs(o.exact_s2(occ2, 8) - np.array(final))) < 1e-6
    loss = float(next(l for l in hist.splitlines() if l.startswith("Final Best Loss: ")).split(": ")[1])
    assert loss <= float(next(l for l in hist.splitlines() if l.startswith("Input Loss: ")).split(": ")[1])


def _faces(d):
    mesh = M.load_stl(d / "input.stl")
    gid, ng = M.granules(mesh)
    out = []
    for g in range(ng):
        f = mesh.faces[gid == g]
        _, inv = np.unique(f.reshape(-1), return_inverse=True)
        out.append(inv.reshape(-1, 3).astype(np.uint32))
    return out
