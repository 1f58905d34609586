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


def test_graph_replays_with_eager_work_between(ctx):
    # three batches of different sizes (the workspace grows, the graph is
    # recaptured, then replayed), with unrelated eager work in between
    sd, parts, rng = make_seed(ctx, 44, 12, 91, 11, overlap=True)
    other = s2.Volume.from_bytes(ctx, (np.random.default_rng(5).random((30, 31, 33)) < 0.3).astype(np.uint8))
    ref_other = other.exact_s2(9)
    for count in (5, 21, 9):
        moves = []
        for _ in range(count):
            i = int(rng.integers(0, len(parts)))
            moves.append((0, i, random_move(rng, parts[i])))
        curves = batch.batch_eval(ctx, [sd], moves)
        assert np.array_equal(other.exact_s2(9), ref_other)
        for (si, i, newv), got in zip(moves, curves):
            assert np.array_equal(got, full_eval(ctx, sd, i, newv)), (count, i)
    other.free()


def test_graph_and_eager_launches_agree(ctx, monkeypatch):
    sd, parts, rng = make_seed(ctx, 36, 9, 17, 9)
    moves = [(0, int(i), random_move(rng, parts[int(i)])) for i in rng.integers(0, len(parts), 17)]
    graph = batch.batch_eval(ctx, [sd], moves)
    monkeypatch.setenv("S2HIP_BATCH_GRAPH", "0")
    eager = batch.batch_eval(ctx, [sd], moves)
    for a, b in zip(graph, eager):
        assert np.array_equal(a, b)


def test_accept_updates_seed_exactly(ctx):
    sd, parts, rng = make_seed(ctx, 40, 10, 23, 10, overlap=True)
    for step in range(3):
        i = int(rng.integers(0, len(parts)))
        newv = random_move(rng, parts[i])
        want = full_eval(ctx, sd, i, newv)
        moved = sd.moved_mesh(i, newv)
        got = sd.accept(i, newv)
        assert np.array_equal(got, want), step
        # the seed's volume is the moved packing's voxelization, its counts match a recount
        ref = s2.voxelize(ctx, moved, sd.bbox, 1.0)
        assert np.array_equal(sd.volume().read_packed(), ref.read_packed())
        ref.free()
        assert np.array_equal(sd.exact_s2(), want)
        # a candidate against the updated seed still matches a full evaluation
        j = int(rng.integers(0, len(parts)))
        cand = sd.meshes[j].vertices + rng.uniform(-1.5, 1.5, 3)
        assert np.array_equal(batch.batch_eval(ctx, [sd], [(0, j, cand)])[0], full_eval(ctx, sd, j, cand))


def test_particle_with_two_components(ctx):
    rng = synth.rng_for(77, 3)
    parts = synth.place_particles(6, 36, rng, synth.ellipsoid_maker(2.5, 3.5), gap=1.0)
    meshes = [p.mesh() for p in parts]
    # particle 0 = two disjoint ellipsoids (two granules in one particle)
    from rustmspt_amd.mesh import merge_meshes
    pair = merge_meshes([meshes[0], meshes[1]])
    meshes = [pair] + meshes[2:]
    bbox = s2.BoundingBox.from_size((36, 36, 36))
    sd = batch.Seed(ctx, meshes, bbox, 1.0, 9)
    moves = [(0, 0, pair.vertices + np.array([1.25, -0.5, 0.75])), (0, 1, meshes[1].vertices - 1.0)]
    for (si, i, newv), got in zip(moves, batch.batch_eval(ctx, [sd], moves)):
        assert np.array_equal(got, full_eval(ctx, sd, i, newv))


@pytest.mark.parametrize("mode", ["fft", "direct"])
def test_region_correlation_and_direct_deltas_agree(ctx, monkeypatch, mode):
    """Delta counts by FP64 correlation over frames around the changed boxes
    (forced with S2HIP_BATCH_DELTA=fft) and by direct counting give the same
    curves, equal to full evaluations: mixed grids, anisotropic grids,
    particles crossing the grid faces, moves far out of the grid, overlaps."""
    seeds, parts_all, rngs = [], [], []
    for g, n, r in [(28, 6, 7), (44, 12, 11), (60, 16, 15)]:
        sd, parts, rng = make_seed(ctx, g, n, 5 * g + 1, r, overlap=True)
        seeds.append(sd)
        parts_all.append(parts)
        rngs.append(rng)
    rng = synth.rng_for(12, 7)
    aparts = synth.place_particles(7, 30, rng, synth.ellipsoid_maker(2.5, 4.5), overlap=True)
    # an anisotropic grid with particles cut by its faces
    aparts += [synth.Particle(np.array(c, float), np.array([4.0, 2.5, 3.0]), synth.random_rotation(rng))
               for c in ([0.5, 10, 10], [29.6, 20, 5], [12, 0.2, 23.5])]
    seeds.append(batch.Seed(ctx, [p.mesh() for p in aparts], s2.BoundingBox((0, 0, 0), (30, 22, 26)), 1.0, 8))
    parts_all.append(aparts)
    rngs.append(rng)
    moves = []
    for k in range(28):
        si = k % 4
        i = int(rngs[si].integers(0, len(parts_all[si])))
        reach = 9.0 if k % 7 == 0 else 2.5
        moves.append((si, i, random_move(rngs[si], parts_all[si][i], reach)))
    moves.append((0, 1, parts_all[0][1].moved(np.array([80.0, 0, 0]), np.eye(3)).mesh().vertices))
    monkeypatch.setenv("S2HIP_BATCH_DELTA", mode)
    got = batch.batch_eval(ctx, seeds, moves)
    monkeypatch.setenv("S2HIP_BATCH_DELTA", "direct" if mode == "fft" else "fft")
    other = batch.batch_eval(ctx, seeds, moves)
    for k, ((si, i, newv), a, b) in enumerate(zip(moves, got, other)):
        assert np.array_equal(a, b), k
        if k % 3 == 0:
            assert np.array_equal(a, full_eval(ctx, seeds[si], i, newv)), k


def test_accept_with_region_correlation_deltas(ctx, monkeypatch):
    monkeypatch.setenv("S2HIP_BATCH_DELTA", "fft")
    sd, parts, rng = make_seed(ctx, 48, 12, 29, 12, overlap=True)
    for step in range(3):
        i = int(rng.integers(0, len(parts)))
        newv = random_move(rng, parts[i])
        want = full_eval(ctx, sd, i, newv)
        assert np.array_equal(sd.accept(i, newv), want), step
        assert np.array_equal(sd.exact_s2(), want)
