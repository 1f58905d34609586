"""Continuous mesh Monte Carlo on the GPU: counts equal the oracle's restatement
of the estimator on the same Philox stream (whole-mesh parity by a full
triangle scan), and the estimates match the closed-form two-point probability of
a sphere.  Requires the GPU (marker gpu)."""
import math

import numpy as np
import pytest

from oracle import oracle as o
from rustmspt_amd import s2, synth
from rustmspt_amd.mesh import BoundingBox, icosphere_mesh, merge_meshes

pytestmark = pytest.mark.gpu


def spheres(centres, radii, level=2):
    return merge_meshes([icosphere_mesh(np.asarray(c, np.float64), r, level) for c, r in zip(centres, radii)])


@pytest.mark.parametrize("case", ["separate", "overlapping", "crossing_box"])
def test_counts_equal_oracle_stream(ctx, case):
    if case == "separate":
        mesh = spheres([(6, 6, 6), (15, 9, 14), (8, 17, 16)], [4.0, 3.0, 5.0])
    elif case == "overlapping":  # whole-mesh parity: the overlap reads as outside
        mesh = spheres([(10, 10, 10), (13, 11, 10)], [5.0, 4.0])
    else:
        mesh = spheres([(2, 10, 10), (14, 12, 11)], [4.0, 5.0])
    bbox = BoundingBox(np.zeros(3), np.full(3, 22.0))
    got = s2.mesh_mc_counts(ctx, mesh, bbox, 7, 3000, seed=11)
    want = o.mesh_mc_counts(mesh.vertices, mesh.faces, bbox.min, bbox.max, 7, 3000, 11)
    assert np.array_equal(got, want)
    assert np.all(got[1:, 0] <= got[1:, 1]) and np.all(got[1:, 1] <= 3000)


def test_sample_floor_and_seed(ctx):
    mesh = spheres([(5, 5, 5)], [3.0])
    bbox = BoundingBox(np.zeros(3), np.full(3, 10.0))
    a = s2.mesh_mc_counts(ctx, mesh, bbox, 3, 10, seed=1)   # raised to 200 samples per radius
    assert np.array_equal(a, o.mesh_mc_counts(mesh.vertices, mesh.faces, bbox.min, bbox.max, 3, 10, 1))
    assert np.all(a[1:, 1] <= 200) and a[1:, 1].sum() > 0
    b = s2.mesh_mc_counts(ctx, mesh, bbox, 3, 10, seed=2)
    assert not np.array_equal(a, b)


def test_sphere_two_point_probability(ctx):
    # a sphere well inside a cube: p inside the sphere keeps q in the box for r <= 8,
    # so hits / N = (V / B) f(r) with f the sphere's overlap fraction, and
    # valid / N = E_d[prod_i (1 - r |d_i| / L)] over uniform directions
    R, L, r_max, n = 6.0, 30.0, 8, 400_000
    mesh = icosphere_mesh(np.array([15.0, 15.0, 15.0]), R, 4)
    bbox = BoundingBox(np.zeros(3), np.full(3, L))
    hv = s2.mesh_mc_counts(ctx, mesh, bbox, r_max, n, seed=3).astype(np.float64)
    vol = s2.geometric_volume_fraction(mesh, bbox) * L ** 3
    rng = np.random.default_rng(0)
    d = rng.normal(size=(4_000_000, 3))
    d = np.abs(d / np.linalg.norm(d, axis=1, keepdims=True))
    for r in range(1, r_max + 1):
        f = 1 - 3 * r / (4 * R) + r ** 3 / (16 * R ** 3)
        p_hit = vol / L ** 3 * f
        sig = math.sqrt(p_hit * (1 - p_hit) / n)
        # faceting of the level-4 icosphere moves the overlap fraction by well under 1 %
        assert abs(hv[r, 0] / n - p_hit) <= 4 * sig + 0.01 * p_hit, (r, hv[r, 0] / n, p_hit)
        p_valid = float(np.mean(np.prod(1 - r * d / L, axis=1)))
        sv = math.sqrt(p_valid * (1 - p_valid) / n)
        assert abs(hv[r, 1] / n - p_valid) <= 4 * sv + 2e-3, (r, hv[r, 1] / n, p_valid)


def test_calculate_s2_routes_pitch_zero_to_mesh_mc(ctx):
    mesh = spheres([(7, 7, 7), (16, 15, 14)], [4.0, 5.0], level=3)
    bbox = BoundingBox(np.zeros(3), np.full(3, 24.0))
    curve = s2.calculate_s2(ctx, mesh, bbox, 6, 0.0, "monte_carlo", 5000, seed=9)
    hv = o.mesh_mc_counts(mesh.vertices, mesh.faces, bbox.min, bbox.max, 6, 5000, 9)
    assert curve[0] == s2.geometric_volume_fraction(mesh, bbox)
    for r in range(1, 7):
        assert curve[r] == hv[r, 0] / hv[r, 1]
    # a granule crossing the box: S2(0) is its clipped in-box volume fraction
    # (volume_fraction_in_bbox clips it, src/geometry/volume.rs:424-430)
    crossing = spheres([(1, 10, 10)], [4.0])
    c = s2.mesh_mc_s2(ctx, crossing, bbox, 4, 1000, 1)
    whole = s2.mesh_volume_in_bbox(crossing, BoundingBox(np.full(3, -50.0), np.full(3, 50.0)))
    assert c[0] == s2.geometric_volume_fraction(crossing, bbox)
    assert 0.5 * whole / 24 ** 3 < c[0] < 0.75 * whole / 24 ** 3


def test_packing_against_voxel_exact(ctx):
    # a particle packing: the continuous estimate and the fine-voxel exact S2
    # agree to within sampling noise plus the voxel discretization
    # (radii are physical: shell r lies at r / pitch voxels)
    mesh, bbox, _ = synth.config_plain(seed=4, n=12, grid=32)
    n = 200_000
    mc = s2.mesh_mc_s2(ctx, mesh, bbox, 8, n, seed=5)
    vol = s2.voxelize(ctx, mesh, bbox, 0.25)
    ex = vol.exact_s2(8, 0.25)
    vol.free()
    assert abs(mc[0] - ex[0]) < 0.05 * ex[0]
    for r in range(1, 9):
        sig = math.sqrt(max(ex[r] * (1 - ex[r]), 1e-12) / n)
        assert abs(mc[r] - ex[r]) <= 4 * sig + 0.1 * ex[r], (r, mc[r], ex[r])


def test_resident_mesh_counts_equal_host_entry_and_oracle(ctx):
    from rustmspt_amd.devmesh import DeviceMesh
    mesh = spheres([(10, 10, 10), (13, 11, 10), (4, 16, 15)], [5.0, 4.0, 3.0])   # overlap and a separate one
    bbox = BoundingBox(np.zeros(3), np.full(3, 22.0))
    dm = DeviceMesh.from_mesh(ctx, mesh)
    got = dm.mesh_mc_counts(bbox, 6, 4000, seed=13)
    assert np.array_equal(got, s2.mesh_mc_counts(ctx, mesh, bbox, 6, 4000, seed=13))
    assert np.array_equal(got, o.mesh_mc_counts(mesh.vertices, mesh.faces, bbox.min, bbox.max, 6, 4000, 13))
    dm.free()


def test_rigid_moves_through_set_vertices(ctx):
    # an optimize-style sequence: move one particle of a resident packing at a
    # time and re-evaluate; each result equals the host-array entry on the moved mesh
    from rustmspt_amd.devmesh import DeviceMesh
    parts = [icosphere_mesh(np.asarray(c, np.float64), r, 2) for c, r in
             [((5, 5, 5), 3.0), ((13, 6, 9), 2.5), ((8, 14, 12), 3.5), ((15, 15, 4), 2.0)]]
    mesh = merge_meshes(parts)
    starts = np.cumsum([0] + [len(p.vertices) for p in parts])
    bbox = BoundingBox(np.zeros(3), np.full(3, 20.0))
    dm = DeviceMesh.from_mesh(ctx, mesh)
    rng = np.random.default_rng(2)
    verts = mesh.vertices.copy()
    for step in range(4):
        i = int(rng.integers(0, len(parts)))
        shift = rng.uniform(-2.0, 2.0, 3)
        sl = slice(starts[i], starts[i + 1])
        verts[sl] = verts[sl] + shift
        dm.set_vertices(int(starts[i]), verts[sl])
        moved = type(mesh)(verts.copy(), mesh.faces)
        assert np.array_equal(dm.mesh_mc_counts(bbox, 5, 3000, seed=step),
                              s2.mesh_mc_counts(ctx, moved, bbox, 5, 3000, seed=step))
        assert dm.bbox.min == tuple(np.fmin.reduce(verts, axis=0).tolist())
    with pytest.raises(Exception, match="vertex range outside"):
        dm.set_vertices(len(verts) - 2, np.zeros((3, 3)))
    dm.free()
