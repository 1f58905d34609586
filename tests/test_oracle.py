"""Pins the CPU oracle (oracle/s2_oracle.c) against golden vectors, closed forms
and brute force.  CPU only."""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as o

G = Path(__file__).resolve().parent / "golden"


def load(name):
    return json.loads((G / name).read_text())


@pytest.fixture(scope="module")
def tiny():
    return load("tiny_pair_counts.json")


@pytest.fixture(scope="module")
def known():
    return load("known_answers.json")


def test_packed_and_bool_counts_match_brute_force(tiny):
    # src/geometry/s2.rs:1742-1772 and 1835-1860 grids; every in-domain offset
    for case in tiny:
        dims = case["dims"]
        occ = np.array(case["occ"], np.uint8).reshape(dims)
        words = o.pack(occ)
        for dx, dy, dz, h, v in case["counts"]:
            if v == 0:
                continue
            assert o.pair_counts_bits(words, dims, (dx, dy, dz)) == (h, v), (dims, dx, dy, dz)
            assert o.pair_counts_bool(occ, (dx, dy, dz)) == (h, v), (dims, dx, dy, dz)


def test_counts_symmetric_and_closed_forms():
    rng = np.random.default_rng(7)
    occ = (rng.random((9, 8, 70)) < 0.35).astype(np.uint8)
    words = o.pack(occ)
    for d in [(1, 2, 3), (0, 0, 64), (-3, 1, 65), (8, -7, 0), (2, 0, -69)]:
        assert o.pair_counts_bits(words, occ.shape, d) == o.pair_counts_bits(words, occ.shape, tuple(-x for x in d))
    # solid: hits == valid; empty: hits == 0
    solid = np.ones((5, 6, 130), np.uint8)
    ws = o.pack(solid)
    for d in [(0, 0, 1), (4, -5, 129), (1, 1, 64)]:
        h, v = o.pair_counts_bits(ws, solid.shape, d)
        assert h == v == (5 - abs(d[0])) * (6 - abs(d[1])) * (130 - abs(d[2]))
    assert o.pair_counts_bits(o.pack(np.zeros((3, 3, 9), np.uint8)), (3, 3, 9), (1, 1, 1))[0] == 0


def test_shell_sizes_match_lattice_count():
    sizes = load("shell_sizes_pitch1.json")
    for r, n in enumerate(sizes, start=1):
        assert len(o.shell_offsets(float(r), 0.5)) == n, r


def test_shell_order_is_dx_dy_dz_ascending():
    # src/geometry/s2.rs:2056-2071: the lazy iterator equals the cube scan order
    for r, w in [(1.0, 0.5), (3.5, 0.5), (8.0, 0.5), (2.0, 1.0), (17.0, 0.5), (0.5, 2.5)]:
        offs = o.shell_offsets(r, w)
        keys = [tuple(x) for x in offs.tolist()]
        assert keys == sorted(keys)
        lo2 = max(r - w, 0.0) ** 2
        hi2 = (r + w) ** 2
        lim = int(np.ceil(r + w + 1))
        ax = np.arange(-lim, lim + 1)
        cube = [(a, b, c) for a in ax for b in ax for c in ax
                if (a, b, c) != (0, 0, 0) and lo2 <= a * a + b * b + c * c < hi2]
        assert keys == [tuple(int(v) for v in k) for k in cube]
    assert o.shell_offsets(0.0, 0.5).tolist() == [[0, 0, 0]]
    assert o.shell_offsets(1e-12, 0.5).tolist() == [[0, 0, 0]]


def test_shell_work_matches_enumeration(known):
    for dims, r_max, k in known["shell_work"]:
        assert o.exact_shell_work(dims, r_max, 1.0)[0] == k
    # src/geometry/s2.rs:1899-1917 cases: K and W by direct enumeration
    for dims, r_max, pitch in [([6, 4, 3], 6, 1.0), ([9, 2, 7], 5, 0.5), ([5, 5, 5], 4, 2.0), ([1, 1, 30], 12, 0.7)]:
        offs, rs = o.domain_offsets(dims, r_max, pitch)
        w = sum(int(np.prod([dims[a] - abs(int(d[a])) for a in range(3)])) for d in offs)
        k, wk, _ = o.exact_shell_work(dims, r_max, pitch)
        assert k == len(offs) and wk == w


def test_planner_at_benchmark_grid():
    # 256^3, r_max=64, 8 workers -> direct ~518 s modeled, FFT over budget
    p = o.plan_exact_cpu([256, 256, 256], 64, 1.0, 8)
    assert p["method"] == "direct" and p["offsets"] == 1123772
    assert abs(p["direct_seconds"] - 517.7) < 1.0 and p["fft_bytes"] > 768 * 1024 * 1024
    assert o.plan_exact_cpu([128, 128, 128], 32, 1.0, 8)["method"] == "fft"
    # src/geometry/s2.rs:1951-1954
    assert o.plan_exact_cpu([64, 64, 64], 60, 1.0, 1, 2 ** 62)["method"] == "fft"
    assert o.plan_exact_cpu([200, 200, 200], 1, 1.0, 1, 2 ** 62)["method"] == "direct"


def test_padded_dims_and_smooth_lengths(known):
    for dims, want in known["padded_fft_dims"]:
        assert o.padded_fft_dims(dims) == want

    def smooth(v):
        for p in (2, 3, 5):
            while v % p == 0:
                v //= p
        return v == 1

    for n in range(0, 1200):  # src/geometry/s2.rs:1877-1895 brute-force minimality
        need = max(2 * n - 1, 1)
        want = next(v for v in range(need, 4 * need + 8) if smooth(v))
        assert o.smooth_fft_length(need) == want


def test_fill_missing_is_natural_cubic_spline():
    for case in load("spline_cases.json"):
        got = o.fill_missing(case["values"], case["support"], case["vf"])
        np.testing.assert_allclose(got, case["expect"], rtol=0, atol=1e-12)


def test_curve_is_ordered_sum_of_ratios(tiny):
    case = tiny[5]  # [1,2,9] i%7<3
    dims = [6, 4, 3]
    occ = (np.arange(72) % 7 < 3).astype(np.uint8).reshape(dims)
    curve = o.exact_s2(occ, 6)
    offs, rs = o.domain_offsets(dims, 6, 1.0)
    vf = occ.sum() / occ.size
    assert curve[0] == vf
    for r in range(1, 7):
        s = 0.0
        for d in offs[rs[r]:rs[r + 1]]:
            h, v = o.brute_pair_counts(occ, d)
            s = s + h / v
        want = s / (rs[r + 1] - rs[r]) if rs[r + 1] > rs[r] else None
        if want is not None:
            assert curve[r] == want  # bit-identical ordered fold
    assert case["dims"] == [1, 2, 9]


def test_fft_rounds_to_brute_force_counts(tiny):
    # src/geometry/s2.rs:1835-1860: the f64 correlation rounds to the integer counts
    for case in tiny[5:12]:
        dims = case["dims"]
        occ = np.array(case["occ"], np.uint8).reshape(dims)
        corr = o.fft_autocorr(occ)
        offs = np.array([[r[0], r[1], r[2]] for r in case["counts"]], np.int64)
        want = np.array([r[3] for r in case["counts"]], np.uint64)
        assert np.array_equal(o.fft_hits(corr, offs), want), dims
        idx = tuple(np.mod(offs[:, a], corr.shape[a]) for a in range(3))
        assert np.max(np.abs(corr[idx] - want.astype(float))) < 1e-6


def test_fft_matches_numpy_fft():
    rng = np.random.default_rng(3)
    occ = (rng.random((10, 7, 12)) < 0.4).astype(np.uint8)
    pad = o.padded_fft_dims(occ.shape)
    corr = o.fft_autocorr(occ, pad)
    f = np.fft.fftn(occ.astype(float), s=pad, axes=(0, 1, 2))
    ref = np.fft.ifftn(f * np.conj(f)).real
    np.testing.assert_allclose(corr, ref, atol=1e-9)


def test_philox_known_answers(known):
    for kat in known["philox4x32_10"]:
        assert o.philox4x32_10(kat["ctr"], kat["key"]) == kat["out"]


def test_voxelize_known_boxes(known):
    ka = known["box_occupancy"]
    (b0, b1) = ka["box"]
    verts, faces = _box(b0, b1)
    occ = o.voxelize(verts, faces, [0, 0, 0], 1.0, ka["domain"])
    x, y, z = np.meshgrid(*[np.arange(n) for n in ka["domain"]], indexing="ij")
    assert np.array_equal(occ, ((x < 4) & (y < 2) & (z < 2)).astype(np.uint8))
    tb = known["translated_box"]
    for origin in tb["origins"]:
        b0 = [origin + v for v in tb["box"][0]]
        b1 = [origin + v for v in tb["box"][1]]
        verts, faces = _box(b0, b1)
        occ = o.voxelize(verts, faces, [origin] * 3, 1.0, tb["domain"])
        x, y, z = np.meshgrid(*[np.arange(n) for n in tb["domain"]], indexing="ij")
        assert np.array_equal(occ, ((x < 2) & (y == 0) & (z == 0)).astype(np.uint8)), origin


def test_known_single_voxel_curve(known):
    ka = known["single_voxel_cube"]
    verts, faces = _box(*ka["box"])
    occ = o.voxelize(verts, faces, [0, 0, 0], ka["pitch"], ka["domain"])
    assert occ.sum() == 1 and occ[0, 0, 0] == 1
    assert o.exact_s2(occ, ka["r_max"]).tolist() == ka["s2"]
    kb = known["two_by_two_cube"]
    verts, faces = _box(*kb["box"])
    occ = o.voxelize(verts, faces, [0, 0, 0], kb["pitch"], kb["domain"])
    assert occ.sum() / occ.size == kb["s2_0"]
    ks = known["solid"]
    curve = o.exact_s2(np.ones(ks["domain"], np.uint8), ks["r_max"])
    assert np.all(curve == ks["value"])


def test_voxelized_icosphere_matches_analytic_sphere():
    from rustmspt_amd.mesh import icosphere_mesh
    c, r = np.array([10.3, 9.7, 11.1]), 6.2
    m = icosphere_mesh(c, r, 3)
    occ = o.voxelize(m.vertices, m.faces, [0, 0, 0], 1.0, [22, 22, 22])
    g = np.stack(np.meshgrid(*[np.arange(22) + 0.5] * 3, indexing="ij"), -1)
    dist = np.linalg.norm(g - c, axis=-1)
    # level-3 faceting stays within ~0.2% of the radius: decided cells away from the surface must agree
    inner, outer = dist < r * 0.98, dist > r * 1.0001
    assert np.all(occ[inner] == 1) and np.all(occ[outer] == 0)


def test_granules_follow_bfs_order():
    from rustmspt_amd.mesh import icosphere_mesh, merge_meshes
    m = merge_meshes([icosphere_mesh([0, 0, 0], 1, 1), icosphere_mesh([5, 0, 0], 1, 0),
                      icosphere_mesh([9, 0, 0], 2, 1)])
    comp, n = o.granules(m.faces, len(m.vertices))
    assert n == 3
    assert comp.tolist() == [0] * 80 + [1] * 20 + [2] * 80


def _box(b0, b1):
    from rustmspt_amd.mesh import box_mesh
    m = box_mesh(b0, b1)
    return m.vertices, m.faces
