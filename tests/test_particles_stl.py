"""The reference's own sample input, data/input/particles.stl (15 particles,
5,600 faces of real, non-synthetic geometry), kept as tests/golden/particles.stl:
the host loader's weld and the granule split against the oracle, the clipped
volume's additivity, and (marker gpu) device ingest, voxelization, exact S2 and
the measure pipeline against the oracle."""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as o
from rustmspt_amd import config as cfg
from rustmspt_amd import mesh as M
from rustmspt_amd import s2

STL = Path(__file__).resolve().parent / "golden" / "particles.stl"


def corners():
    data = STL.read_bytes()
    n = int(np.frombuffer(data[80:84], "<u4")[0])
    rec = np.frombuffer(data[84:84 + 50 * n], dtype=[("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")])
    return rec["v"].reshape(-1, 3).astype(np.float64)


def test_fixture_weld_and_granules_match_oracle():
    m = M.load_stl(STL)
    v, t = o.weld(corners())
    assert (len(v), len(t) // 3) == (2830, 5600)
    assert np.array_equal(m.vertices, v) and np.array_equal(m.faces.reshape(-1), t)
    gid, ng = M.granules(m)
    go, no = o.granules(t.reshape(-1, 3), len(v))
    assert ng == no == 15 and np.array_equal(gid, go)
    # every particle is closed and wound outward (positive signed volume)
    for g in range(ng):
        f = m.faces[gid == g]
        a, b, c = (m.vertices[f[:, k]] for k in range(3))
        assert np.einsum("ij,ij->i", a, np.cross(b, c)).sum() > 0


def test_fixture_clipped_volume_is_additive():
    """Splitting the box by a plane through the particles splits the clipped
    volume: left + right equals the whole (the capped cuts close each piece)."""
    m = M.load_stl(STL)
    bb = M.mesh_bbox(m)
    lo, hi = np.array(bb.min), np.array(bb.max)
    whole = s2.mesh_volume_in_bbox(m, bb)
    assert whole > 4000.0
    for axis in range(3):
        for frac in (0.31, 0.5, 0.77):
            cut = lo[axis] + frac * (hi[axis] - lo[axis])
            h1, l2 = hi.copy(), lo.copy()
            h1[axis] = l2[axis] = cut
            parts = s2.mesh_volume_in_bbox(m, M.BoundingBox(tuple(lo), tuple(h1))) + \
                s2.mesh_volume_in_bbox(m, M.BoundingBox(tuple(l2), tuple(hi)))
            assert abs(parts - whole) <= 1e-9 * whole, (axis, frac)
    vf = s2.geometric_volume_fraction(m, bb)
    assert vf == pytest.approx(whole / float(np.prod(hi - lo)), rel=1e-12)


# ------------------------------------------------------------------ on the GPU

@pytest.mark.gpu
def test_fixture_device_ingest_matches_oracle(ctx):
    from rustmspt_amd.devmesh import DeviceMesh
    dm = DeviceMesh.from_stl_bytes(ctx, [STL.read_bytes()])
    mesh, gid = dm.read()
    v, t = o.weld(corners())
    assert np.array_equal(mesh.vertices, v) and np.array_equal(mesh.faces.reshape(-1), t)
    go, ng = o.granules(t.reshape(-1, 3), len(v))
    assert dm.ngranules == ng == 15 and np.array_equal(gid, go)


@pytest.mark.gpu
@pytest.mark.parametrize("pitch", [1.0, 0.5])
def test_fixture_voxels_and_exact_s2_match_oracle(ctx, pitch):
    m = M.load_stl(STL)
    bb = M.mesh_bbox(m)
    dims = s2.occupancy_dims(bb, pitch)
    assert list(dims) == list(o.occupancy_dims(bb.min, bb.max, pitch))
    vol = s2.voxelize(ctx, m, bb, pitch)
    occ = vol.to_bytes()
    want = o.voxelize(m.vertices, m.faces, list(bb.min), pitch, dims)
    assert np.array_equal(occ, want) and want.sum() > 1000
    r_max = 20
    assert np.array_equal(vol.exact_s2(r_max, pitch), o.exact_s2(want, r_max, pitch))
    vol.free()


@pytest.mark.gpu
def test_fixture_measure_report(tmp_path):
    """The reference's measure settings (data/input/measure_config.yaml: r_max 20,
    pitch 1, both methods, 40,000 samples) on the sample particles, with a 50^3
    box that cuts through several of them."""
    from rustmspt_amd import pipeline
    box = [110.0, 615.0, 1210.0, 160.0, 665.0, 1260.0]
    doc = {"measurement": {"stl_path": str(STL), "stl_bounding_box": box, "r_max": 20, "voxel_pitch": 1.0,
                           "mc_method": "both", "mc_samples": 40000, "output_path": str(tmp_path / "s2.txt"),
                           "acceleration": {"mode": "hip"}}}
    (tmp_path / "c.json").write_text(json.dumps(doc))
    pipeline.MeasurePipeline(cfg.load_measurement_config(tmp_path / "c.json"), env={"RUSTMSPT_S2_SEED": "3"}).run()
    text = (tmp_path / "s2.txt").read_text()
    m = M.load_stl(STL)
    bb = M.BoundingBox(tuple(box[:3]), tuple(box[3:]))
    occ = o.voxelize(m.vertices, m.faces, box[:3], 1.0, [50, 50, 50])
    want = o.exact_s2(occ, 20)
    block = text.split("S2 Values [exact]:\n")[1].split("S2(0)-VF")[0]
    for r, v in enumerate(want):
        assert f"{r}: {v:.6f}" in block
    vf = s2.geometric_volume_fraction(m, bb)
    assert text.startswith(f"Volume Fraction: {vf:.6f}")
    # particles cross the box: the clipped in-box volume is below the whole mesh's
    assert 0.0 < vf * 50 ** 3 < s2.mesh_volume_in_bbox(m, M.mesh_bbox(m))
