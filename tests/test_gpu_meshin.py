"""Device-resident meshes: STL weld and granule split on the device equal the
oracle's restatement of the reference's weld rule and granule order (and the
host loader), and voxelizing the resident mesh gives the same bits as the host
preparation path.  Requires the GPU (marker gpu)."""
import numpy as np
import pytest

from oracle import oracle as o
from rustmspt_amd import mesh as M
from rustmspt_amd import s2, synth
from rustmspt_amd.devmesh import DeviceMesh

pytestmark = pytest.mark.gpu


def binary_stl(mesh: M.Mesh, path):
    M.save_stl(path, mesh)
    return path.read_bytes()


def ascii_stl(corners: np.ndarray) -> bytes:
    lines = ["solid t"]
    for f in range(len(corners) // 3):
        lines += ["facet normal 0 0 0", "outer loop"]
        lines += [f"vertex {x!r} {y!r} {z!r}" for x, y, z in corners[3 * f:3 * f + 3].tolist()]
        lines += ["endloop", "endfacet"]
    lines.append("endsolid t")
    return ("\n".join(lines) + "\n").encode()


def corners_of(data: bytes) -> np.ndarray:
    """f64 corners a binary STL holds (f32 widened)."""
    n = int(np.frombuffer(data[80:84], "<u4")[0])
    rec = np.frombuffer(data[84:84 + 50 * n], dtype=[("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")])
    return rec["v"].reshape(-1, 3).astype(np.float64)


def check_against_oracle(dm: DeviceMesh, corners, part=None):
    mesh, gid = dm.read()
    v, t = o.weld(corners, part)
    assert np.array_equal(mesh.vertices, v) and np.array_equal(mesh.faces.reshape(-1), t)
    go, ng = o.granules(t.reshape(-1, 3), len(v))
    assert dm.ngranules == ng and np.array_equal(gid, go)
    bb = M.mesh_bbox(mesh)
    assert dm.bbox.min == bb.min and dm.bbox.max == bb.max
    return mesh, gid


def test_binary_weld_and_granules(ctx, tmp_path):
    mesh, bbox, _ = synth.config_plain(seed=3, n=40, grid=48)
    rng = np.random.default_rng(1)
    perm = rng.permutation(len(mesh.faces))  # interleave the particles' faces
    mesh = M.Mesh(mesh.vertices, mesh.faces[perm])
    data = binary_stl(mesh, tmp_path / "p.stl")
    dm = DeviceMesh.from_stl_bytes(ctx, [data])
    got, gid = check_against_oracle(dm, corners_of(data))
    host = M.load_stl(tmp_path / "p.stl")
    hg, hn = M.granules(host)
    assert np.array_equal(got.vertices, host.vertices) and np.array_equal(got.faces, host.faces)
    assert np.array_equal(gid, hg) and dm.ngranules == hn == 40
    dm.free()


def test_weld_key_boundaries_ascii(ctx):
    # coordinates on and around the 1e-6 rounding boundaries, negative values,
    # exact duplicates and near-duplicates that weld
    rng = np.random.default_rng(7)
    base = rng.integers(-50, 50, (600, 3)).astype(np.float64)
    jitter = rng.choice([0.0, 0.49e-6, -0.49e-6, 0.5e-6, -0.5e-6, 0.51e-6, 1e-7, 3e-6], (600, 3))
    pts = base * 1e-3 + jitter
    idx = rng.integers(0, 600, 3 * 900)
    corners = pts[idx]
    dm = DeviceMesh.from_stl_bytes(ctx, [ascii_stl(corners)])
    check_against_oracle(dm, corners)
    dm.free()


def test_files_never_weld_across(ctx, tmp_path):
    a = M.icosphere_mesh(np.array([0.0, 0.0, 0.0]), 1.0, 2)
    b = M.icosphere_mesh(np.array([0.0, 0.0, 0.0]), 1.0, 1)     # shares vertex positions with a
    c = M.box_mesh([3, 3, 3], [4, 5, 6])
    files = [binary_stl(a, tmp_path / "a.stl"), binary_stl(b, tmp_path / "b.stl")]
    cc = c.vertices[c.faces].reshape(-1, 3)
    files.append(ascii_stl(cc))
    (tmp_path / "c.stl").write_bytes(files[2])
    corners = np.concatenate([corners_of(files[0]), corners_of(files[1]), cc])
    part = np.concatenate([np.full(3 * len(a.faces), 0), np.full(3 * len(b.faces), 1), np.full(len(cc), 2)])
    dm = DeviceMesh.from_stl_bytes(ctx, files)
    got, gid = check_against_oracle(dm, corners, part)
    assert dm.ngranules == 3
    # the folder loader reads the same files in name order
    dmf = DeviceMesh.load_stl(ctx, tmp_path)
    f_mesh, f_gid = dmf.read()
    host = M.load_stl_or_merge_folder(tmp_path)
    assert np.array_equal(f_mesh.vertices, host.vertices) and np.array_equal(f_mesh.faces, host.faces)
    assert np.array_equal(f_gid, M.granules(host)[0])
    dm.free()
    dmf.free()


def test_stl_errors(ctx, tmp_path):
    with pytest.raises(Exception, match="Binary STL size mismatch"):
        DeviceMesh.from_stl_bytes(ctx, [b"\0" * 80 + np.array([5], "<u4").tobytes() + b"\0" * 100])
    with pytest.raises(Exception, match="cannot read"):
        DeviceMesh.load_stl(ctx, tmp_path / "missing.stl")


@pytest.mark.parametrize("case", ["packing", "overlap", "crossing", "fine_pitch", "outside"])
def test_resident_voxelize_equals_host(ctx, case):
    if case == "packing":
        mesh, bbox, _ = synth.config_plain(seed=11, n=30, grid=40)
        pitch = 1.0
    elif case == "overlap":
        mesh = M.merge_meshes([M.icosphere_mesh(np.array([8.0, 8.0, 8.0]), 5.0, 3),
                               M.icosphere_mesh(np.array([11.0, 9.0, 8.5]), 4.0, 3)])
        bbox, pitch = M.BoundingBox((0.0, 0.0, 0.0), (20.0, 19.0, 18.0)), 0.5
    elif case == "crossing":
        mesh = M.merge_meshes([M.icosphere_mesh(np.array([1.0, 9.0, 9.0]), 4.0, 3),
                               M.box_mesh([10.0, -2.0, 5.0], [17.5, 6.25, 13.0])])
        bbox, pitch = M.BoundingBox((0.0, 0.0, 0.0), (16.0, 16.0, 16.0)), 0.75
    elif case == "fine_pitch":
        mesh = M.icosphere_mesh(np.array([0.3, 0.2, 0.1]), 0.25, 3)
        bbox, pitch = M.BoundingBox((0.0, 0.0, 0.0), (0.6, 0.45, 0.4)), 0.004
    else:
        mesh = M.merge_meshes([M.icosphere_mesh(np.array([5.0, 5.0, 5.0]), 3.0, 2),
                               M.icosphere_mesh(np.array([50.0, 5.0, 5.0]), 3.0, 2)])
        bbox, pitch = M.BoundingBox((0.0, 0.0, 0.0), (10.0, 10.0, 10.0)), 1.0
    host = s2.voxelize(ctx, mesh, bbox, pitch)
    dm = DeviceMesh.from_mesh(ctx, mesh)
    dev = dm.voxelize(bbox, pitch)
    assert dev.dims == host.dims
    assert np.array_equal(dev.read_packed(), host.read_packed())
    assert dev.occupied == host.occupied > 0
    host.free()
    dev.free()
    dm.free()


def test_large_packing_ingest_equals_host(ctx, tmp_path):
    # the 500-particle single-measure packing: STL bytes -> weld -> granules ->
    # voxels entirely on the device, equal to the host loader and voxelizer
    mesh, bbox, _ = synth.config_large(seed=2, n=500, grid=128)
    p = tmp_path / "large.stl"
    M.save_stl(p, mesh)
    dm = DeviceMesh.load_stl(ctx, p)
    host_mesh = M.load_stl(p)
    got, gid = dm.read()
    assert np.array_equal(got.vertices, host_mesh.vertices) and np.array_equal(got.faces, host_mesh.faces)
    hg, hn = M.granules(host_mesh)
    assert np.array_equal(gid, hg) and dm.ngranules == hn
    host = s2.voxelize(ctx, host_mesh, bbox, 1.0)
    dev = dm.voxelize(bbox, 1.0)
    assert np.array_equal(dev.read_packed(), host.read_packed())
    host.free()
    dev.free()
    dm.free()
