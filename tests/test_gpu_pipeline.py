"""measure pipeline end to end on the GPU: the reference's integration contracts
(tests/measure_execution_tests.rs) with acceleration.mode hip."""
import json

import numpy as np
import pytest

from oracle import oracle as o
from rustmspt_amd import config as cfg
from rustmspt_amd import mesh as M
from rustmspt_amd import pipeline, synth

pytestmark = pytest.mark.gpu


def run(tmp_path, m, env=None, **over):
    doc = {"measurement": {"stl_path": str(tmp_path / "in.stl"), "output_path": str(tmp_path / "out.txt"),
                           "r_max": 2, "voxel_pitch": 0.0, "mc_method": "exact", "mc_samples": 200,
                           "bounding_box": [4, 4, 4], "acceleration": {"mode": "hip"}}}
    doc["measurement"].update(over)
    M.save_stl(tmp_path / "in.stl", m)
    (tmp_path / "c.json").write_text(json.dumps(doc))
    params = cfg.load_measurement_config(tmp_path / "c.json")
    pipeline.MeasurePipeline(params, env=env or {}).run()
    return (tmp_path / "out.txt").read_text()


def test_single_voxel_cube_report(tmp_path):
    text = run(tmp_path, M.box_mesh([0.2, 0.2, 0.2], [1.4, 1.4, 1.4]))
    assert "exact=hip/" in text
    for r, v in enumerate([1 / 64, 0.0, 0.0]):
        assert f"{r}: {v:.6f}" in text
    assert "Method: exact" in text and text.startswith("Volume Fraction: ")


def test_policy_paths(tmp_path):
    cube = M.box_mesh([0.2, 0.2, 0.2], [1.4, 1.4, 1.4])
    with pytest.raises(pipeline.CpuPathUnavailable):
        run(tmp_path, cube, env={"RUSTMSPT_ACCELERATION": "cpu"})
    with pytest.raises(cfg.ConfigError):
        run(tmp_path, cube, env={"RUSTMSPT_ACCELERATION": "invalid"})
    with pytest.raises(cfg.GpuError):
        run(tmp_path, cube, acceleration={"mode": "hip", "cpu_fallback": False, "gpu_memory_limit_mb": 0})


def test_solid_box_all_ones(tmp_path):
    text = run(tmp_path, M.box_mesh([0, 0, 0], [32, 32, 32]), r_max=16, voxel_pitch=1.0, bounding_box=[32, 32, 32],
               acceleration={"mode": "hip", "cpu_fallback": False, "gpu_memory_limit_mb": 64})
    for r in range(17):
        assert f"{r}: 1.000000" in text


def test_both_methods_match_oracle(tmp_path):
    mesh, bbox, _ = synth.config_plain(seed=5, n=8, grid=40)
    text = run(tmp_path, mesh, r_max=8, voxel_pitch=1.0, bounding_box=[40, 40, 40], mc_method="both",
               mc_samples=100000)
    occ = o.voxelize(mesh.vertices, mesh.faces, [0, 0, 0], 1.0, [40, 40, 40])
    want = o.exact_s2(occ, 8)
    block = text.split("S2 Values [exact]:\n")[1].split("S2(0)-VF")[0]
    for r, v in enumerate(want):
        assert f"{r}: {v:.6f}" in block
    assert "L2 error [exact vs monte_carlo]:" in text and "monte_carlo=hip/" in text


def test_device_and_host_ingest_reports_match(tmp_path):
    mesh, bbox, _ = synth.config_plain(seed=9, n=10, grid=36)
    kw = dict(r_max=6, voxel_pitch=1.0, bounding_box=[36, 36, 36], mc_method="both", mc_samples=20000)
    dev = run(tmp_path, mesh, env={"RUSTMSPT_S2_SEED": "4"}, **kw)
    host = run(tmp_path, mesh, env={"RUSTMSPT_S2_SEED": "4", "S2HIP_HOST_INGEST": "1"}, **kw)
    assert dev == host and "exact=hip/" in dev


def test_monte_carlo_with_pitch_zero_is_mesh_mc(tmp_path):
    # calculate_s2 routes monte_carlo with voxel_pitch <= 0 to the continuous mesh MC
    # (src/geometry/s2.rs:1351-1356); S2(0) is the geometric volume fraction
    mesh = M.merge_meshes([M.icosphere_mesh(np.array([1.2, 1.2, 1.0]), 0.6, 3),
                           M.icosphere_mesh(np.array([2.6, 2.8, 3.0]), 0.8, 3)])
    text = run(tmp_path, mesh, env={"RUSTMSPT_S2_SEED": "21"}, r_max=3, voxel_pitch=0.0, mc_method="monte_carlo",
               mc_samples=5000)
    from rustmspt_amd import s2
    ctx = s2.Context(0)
    try:
        back = M.load_stl(tmp_path / "in.stl")
        box = M.BoundingBox((0.0, 0.0, 0.0), (4.0, 4.0, 4.0))
        want = s2.mesh_mc_s2(ctx, back, box, 3, 5000, 21)
    finally:
        ctx.close()
    assert f"Volume Fraction: {want[0]:.6f}" in text
    for r, v in enumerate(want):
        assert f"{r}: {v:.6f}" in text


def test_boxes_crossing_the_region_report_clipped_volume_fraction(tmp_path):
    # two outward-wound boxes cut by the 4^3 measurement box: in-box volumes
    # 3.2 x 2.1 x 1.3 = 8.736 and 1.4 x 0.9 x 1.7 = 2.142, so the geometric
    # volume fraction is 10.878 / 64 = 0.16996875 (volume_fraction_in_bbox,
    # src/geometry/volume.rs:437-469); the voxel centres they cover at pitch 1
    # are 3 x 2 x 1 and 1 x 1 x 2, so S2(0) = 8 / 64 = 0.125 and the
    # S2(0)-VF difference is 0.04496875
    def outward(lo, hi):
        m = M.box_mesh(lo, hi)
        return M.Mesh(m.vertices, m.faces[:, [0, 2, 1]].copy())
    mesh = M.merge_meshes([outward([-1.0, 0.6, 0.9], [3.2, 2.7, 2.2]), outward([2.6, 3.1, -0.5], [5.0, 4.7, 1.7])])
    text = run(tmp_path, mesh, r_max=1, voxel_pitch=1.0, bounding_box=[4, 4, 4])
    assert text.startswith("Volume Fraction: 0.169969\n"), text
    assert "S2(0)-VF diff [exact]: 0.044969\n" in text
    assert "0: 0.125000\n" in text
