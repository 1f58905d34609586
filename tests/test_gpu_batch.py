"""Batched candidate evaluation (exact delta counts) vs full evaluation of each
moved packing: curves bit-identical.  Requires the GPU (marker gpu)."""
import numpy as np
import pytest

from rustmspt_amd import batch, s2, synth

pytestmark = pytest.mark.gpu


def make_seed(ctx, grid, n, seed, r_max, overlap=False):
    rng = synth.rng_for(seed, 31)
    parts = synth.place_particles(n, grid, rng, synth.ellipsoid_maker(0.06 * grid, 0.10 * grid), overlap=overlap,
                                  gap=0.5)
    bbox = s2.BoundingBox.from_size((grid, grid, grid))
    return batch.Seed(ctx, [p.mesh() for p in parts], bbox, 1.0, r_max), parts, rng


def random_move(rng, p, reach=2.0):
    shift = rng.uniform(-reach, reach, 3)
    rot = synth.random_rotation(rng)
    return p.moved(shift, rot).mesh().vertices


def full_eval(ctx, sd, index, newv):
    mesh = sd.moved_mesh(index, newv)
    vol = s2.voxelize(ctx, mesh, sd.bbox, sd.pitch)
    out = vol.exact_s2(sd.r_max, sd.pitch)
    vol.free()
    return out


def test_seed_curve_equals_full(ctx):
    sd, parts, _ = make_seed(ctx, 40, 10, 1, 10)
    mesh = sd.moved_mesh(0, sd.meshes[0].vertices)
    vol = s2.voxelize(ctx, mesh, sd.bbox, 1.0)
    assert np.array_equal(sd.exact_s2(), vol.exact_s2(10))


@pytest.mark.parametrize("grid,n,r_max,overlap", [(32, 8, 8, False), (48, 14, 12, True), (40, 10, 10, False)])
def test_batch_equals_full_evaluation(ctx, grid, n, r_max, overlap):
    sd, parts, rng = make_seed(ctx, grid, n, grid + n, r_max, overlap)
    moves = []
    for k in range(12):
        i = int(rng.integers(0, len(parts)))
        moves.append((0, i, random_move(rng, parts[i], reach=2.0 if k % 3 else 6.0)))
    curves = batch.batch_eval(ctx, [sd], moves)
    for (si, i, newv), got in zip(moves, curves):
        assert np.array_equal(got, full_eval(ctx, sd, i, newv)), i


def test_mixed_grid_sizes_in_one_batch(ctx):
    seeds, parts_all, rngs = [], [], []
    for g, n in [(24, 5), (36, 9), (52, 14)]:
        sd, parts, rng = make_seed(ctx, g, n, 7 * g, g // 4)
        seeds.append(sd)
        parts_all.append(parts)
        rngs.append(rng)
    moves = []
    for k in range(18):
        si = k % 3
        i = int(rngs[si].integers(0, len(parts_all[si])))
        moves.append((si, i, random_move(rngs[si], parts_all[si][i])))
    curves = batch.batch_eval(ctx, seeds, moves)
    for (si, i, newv), got in zip(moves, curves):
        assert np.array_equal(got, full_eval(ctx, seeds[si], i, newv))


def test_move_out_of_grid_and_identity(ctx):
    sd, parts, rng = make_seed(ctx, 32, 6, 3, 8)
    same = batch.batch_eval(ctx, [sd], [(0, 0, sd.meshes[0].vertices)])[0]
    assert np.array_equal(same, sd.exact_s2())
    far = parts[1].moved(np.array([100.0, 0.0, 0.0]), np.eye(3)).mesh().vertices
    got = batch.batch_eval(ctx, [sd], [(0, 1, far)])[0]
    assert np.array_equal(got, full_eval(ctx, sd, 1, far))
