"""Device voxelization vs the oracle's f64 predicate: bit-identical occupancy.
Requires the GPU (marker gpu)."""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as o
from rustmspt_amd import mesh as M
from rustmspt_amd import s2, synth

pytestmark = pytest.mark.gpu
G = Path(__file__).resolve().parent / "golden"


def device_occ(ctx, mesh, bmin, pitch, dims):
    bbox = M.BoundingBox(tuple(bmin), tuple(bmin[a] + dims[a] * pitch for a in range(3)))
    vol = s2.voxelize(ctx, mesh, bbox, pitch)
    assert vol.dims == list(dims)
    occ = vol.to_bytes()
    assert vol.occupied == int(occ.sum())
    vol.free()
    return occ


def check(ctx, mesh, bmin, pitch, dims):
    got = device_occ(ctx, mesh, bmin, pitch, dims)
    want = o.voxelize(mesh.vertices, mesh.faces, bmin, pitch, dims)
    assert np.array_equal(got, want), int(np.count_nonzero(got != want))
    return got


def test_known_boxes(ctx):
    known = json.loads((G / "known_answers.json").read_text())
    ka = known["box_occupancy"]
    occ = check(ctx, M.box_mesh(*ka["box"]), [0.0, 0.0, 0.0], 1.0, ka["domain"])
    x, y, z = np.meshgrid(*[np.arange(n) for n in ka["domain"]], indexing="ij")
    assert np.array_equal(occ, ((x < 4) & (y < 2) & (z < 2)).astype(np.uint8))
    tb = known["translated_box"]
    for origin in tb["origins"]:
        b0 = [origin + v for v in tb["box"][0]]
        b1 = [origin + v for v in tb["box"][1]]
        occ = check(ctx, M.box_mesh(b0, b1), [origin] * 3, 1.0, tb["domain"])
        x, y, z = np.meshgrid(*[np.arange(n) for n in tb["domain"]], indexing="ij")
        assert np.array_equal(occ, ((x < 2) & (y == 0) & (z == 0)).astype(np.uint8)), origin


def test_single_voxel_cube(ctx):
    occ = check(ctx, M.box_mesh([0.2, 0.2, 0.2], [1.4, 1.4, 1.4]), [0.0, 0.0, 0.0], 1.0, [4, 4, 4])
    assert occ.sum() == 1 and occ[0, 0, 0] == 1


def test_adversarial_meshes(ctx):
    # faces exactly on voxel centres, ulp-offset faces, shared faces, coincident
    # duplicates, a 3e-7 slab, tiny features (patterned on src/gpu/certify.rs:220-277)
    cases = [
        M.box_mesh([0.5, 0.5, 0.5], [3.5, 2.5, 4.5]),
        M.box_mesh([np.nextafter(0.5, 1), 0.5, np.nextafter(0.5, 0)], [3.5, np.nextafter(2.5, 3), 4.5]),
        M.merge_meshes([M.box_mesh([1, 1, 1], [3, 3, 3]), M.box_mesh([3, 1, 1], [5, 3, 3])]),
        M.merge_meshes([M.box_mesh([1, 1, 1], [4, 4, 4]), M.box_mesh([1, 1, 1], [4, 4, 4])]),
        M.box_mesh([0.0, 0.0, 2.5 - 1.5e-7], [6.0, 6.0, 2.5 + 1.5e-7]),
        M.box_mesh([2.49, 2.49, 2.49], [2.51, 2.51, 2.51]),
        M.box_mesh([-1.0, -1.0, -1.0], [7.0, 7.0, 7.0]),
    ]
    dup = M.box_mesh([1.0, 1.5, 0.5], [5.5, 4.5, 5.5])
    cases.append(M.Mesh(dup.vertices, np.concatenate([dup.faces, dup.faces[:4]])))
    for m in cases:
        check(ctx, m, [0.0, 0.0, 0.0], 1.0, [6, 6, 6])
        check(ctx, m, [0.25, -0.5, 0.0], 0.5, [12, 13, 11])


def test_icosphere_packings(ctx):
    for seed, grid, n, overlap in [(1, 32, 6, False), (2, 48, 12, True), (3, 40, 8, False)]:
        rng = synth.rng_for(seed, 9)
        parts = synth.place_particles(n, grid, rng, synth.mixed_maker(3.0, 7.0), overlap=overlap)
        mesh = synth.packing_mesh(parts)
        check(ctx, mesh, [0.0, 0.0, 0.0], 1.0, [grid] * 3)
    # particles crossing the domain boundary and a fractional pitch
    parts = [synth.Particle(np.array(c, float), np.array([5.0, 3.0, 4.0]), synth.random_rotation(synth.rng_for(4, 1)))
             for c in ([0.5, 10, 10], [20, 19.7, 1.0], [10, 10, 10])]
    check(ctx, synth.packing_mesh(parts), [0.0, 0.0, 0.0], 0.75, [27, 27, 27])


def test_far_origin_icosphere(ctx):
    for origin in (1e6, 1e9):
        m = M.icosphere_mesh([origin + 8.3, origin + 7.9, origin + 8.6], 5.5, 3)
        check(ctx, m, [origin] * 3, 1.0, [17, 17, 17])


def test_plain_config_subset(ctx):
    mesh, bbox, parts = synth.config_plain(seed=1, n=20, grid=128)
    got = device_occ(ctx, mesh, [0.0, 0.0, 0.0], 1.0, [128, 128, 128])
    want = o.voxelize(mesh.vertices, mesh.faces, [0.0, 0.0, 0.0], 1.0, [128, 128, 128])
    assert np.array_equal(got, want)
