"""Writes the golden fixtures under tests/golden/ from first principles.

Nothing here uses the oracle or the engine: pair counts come from exhaustive
numpy enumeration, shell sizes from a direct lattice count, spline values from
scipy's natural cubic spline, and the remaining entries are the known answers
stated by the reference's own tests (cited per entry).

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def xorshift_fill(dims, state=0x9E3779B97F4A7C15):
    """The reference's packed-vs-bool fixture pattern (src/geometry/s2.rs:1744-1756):
    xorshift64 (13, 7, 17); voxel set when next() % 3 == 0, in flat index order."""
    out = np.zeros(int(np.prod(dims)), np.uint8)
    mask = (1 << 64) - 1
    for i in range(out.size):
        state ^= (state << 13) & mask
        state ^= state >> 7
        state ^= (state << 17) & mask
        out[i] = 1 if state % 3 == 0 else 0
    return out.reshape(dims), state


def brute(occ, d):
    n = occ.shape
    sa, sb = [], []
    for a in range(3):
        if abs(d[a]) >= n[a]:
            return 0, 0
        if d[a] >= 0:
            sa.append(slice(0, n[a] - d[a]))
            sb.append(slice(d[a], n[a]))
        else:
            sa.append(slice(-d[a], n[a]))
            sb.append(slice(0, n[a] + d[a]))
    A = occ[tuple(sa)].astype(bool)
    B = occ[tuple(sb)].astype(bool)
    return int(np.count_nonzero(A & B)), int(A.size)


def tiny_pair_counts():
    cases = []
    # src/geometry/s2.rs:1750 packed_pair_counts_equal_the_bool_scan grids, random fill
    state = 0x9E3779B97F4A7C15
    for dims in ([2, 3, 1], [3, 2, 63], [2, 2, 64], [3, 1, 65], [2, 2, 130]):
        occ, state = xorshift_fill(dims, state)
        reach = [min(n - 1, 70) for n in dims]
        rows = []
        for dx in range(-reach[0], reach[0] + 1):
            for dy in range(-reach[1], reach[1] + 1):
                for dz in range(-reach[2], reach[2] + 1):
                    h, v = brute(occ, (dx, dy, dz))
                    rows.append([dx, dy, dz, h, v])
        cases.append({"dims": dims, "fill": "xorshift%3", "occ": occ.reshape(-1).tolist(), "counts": rows})
    # src/geometry/s2.rs:1836 fft_scratch_matches_integer_pairs grids, occ[i] = i % 7 < 3
    for dims in ([1, 2, 9], [6, 4, 3], [8, 2, 5], [7, 11, 13], [1, 1, 40], [2, 37, 3], [4, 1, 1]):
        occ = (np.arange(int(np.prod(dims))) % 7 < 3).astype(np.uint8).reshape(dims)
        rows = []
        for dx in range(-(dims[0] - 1), dims[0]):
            for dy in range(-(dims[1] - 1), dims[1]):
                for dz in range(-(dims[2] - 1), dims[2]):
                    h, v = brute(occ, (dx, dy, dz))
                    rows.append([dx, dy, dz, h, v])
        cases.append({"dims": dims, "fill": "i%7<3", "occ": occ.reshape(-1).tolist(), "counts": rows})
    # tests/gpu_s2_boundaries_tests.rs:49-99: [3,2,4], occ[i] = i % 3 != 0
    occ = (np.arange(24) % 3 != 0).astype(np.uint8).reshape(3, 2, 4)
    rows = []
    for dx in range(-2, 3):
        for dy in range(-1, 2):
            for dz in range(-3, 4):
                h, v = brute(occ, (dx, dy, dz))
                rows.append([dx, dy, dz, h, v])
    cases.append({"dims": [3, 2, 4], "fill": "i%3!=0", "occ": occ.reshape(-1).tolist(), "counts": rows})
    return cases


def shell_sizes(r_max=64):
    """Lattice points d != 0 with (r - 1/2)^2 <= |d|^2 < (r + 1/2)^2 (pitch 1), i.e.
    r^2 - r + 1 <= |d|^2 <= r^2 + r, counted per radius."""
    lim = r_max + 2
    ax = np.arange(-lim, lim + 1)
    d2 = (ax[:, None, None] ** 2 + ax[None, :, None] ** 2 + ax[None, None, :] ** 2).reshape(-1)
    out = []
    for r in range(1, r_max + 1):
        out.append(int(np.count_nonzero((d2 >= r * r - r + 1) & (d2 <= r * r + r))))
    return out


def spline_cases():
    from scipy.interpolate import CubicSpline

    rng = np.random.default_rng(20260925)
    cases = []
    for n, pattern in [(12, [0, 1, 2, 5, 9, 11]), (9, [0, 4, 8]), (7, [0, 6]), (16, [0, 1, 3, 7, 8, 15]),
                       (10, [0, 2, 4, 6, 8, 9]), (20, [0, 5, 6, 13, 19])]:
        vals = np.zeros(n)
        sup = np.zeros(n, np.uint8)
        vf = float(rng.uniform(0.1, 0.6))
        for i in pattern:
            sup[i] = 1
            vals[i] = float(rng.uniform(0.0, 0.8)) if i else vf
        vals[0] = vf
        xs = np.array(pattern, float)
        ys = vals[pattern]
        expect = vals.copy()
        if len(pattern) == 2:
            xl, xr = pattern
            for i in range(xl + 1, xr):
                t = (i - xl) / (xr - xl)
                s = t * t * (3 - 2 * t)
                expect[i] = min(1.0, max(0.0, ys[0] + (ys[1] - ys[0]) * s))
        elif len(pattern) >= 3:
            cs = CubicSpline(xs, ys, bc_type="natural")
            for i in range(n):
                if not sup[i]:
                    if i < pattern[-1]:
                        expect[i] = min(1.0, max(0.0, float(cs(i))))
        cases.append({"values": vals.tolist(), "support": sup.tolist(), "vf": vf, "expect": expect.tolist()})
    return cases


def known_answers():
    return {
        # tests/measure_execution_tests.rs:13-24 + src/pipeline/optimize_execution.rs:560: a
        # [0.2,1.4]^3 box in a 4^3 domain at pitch 1 covers only voxel (0,0,0)
        "single_voxel_cube": {"box": [[0.2, 0.2, 0.2], [1.4, 1.4, 1.4]], "domain": [4, 4, 4], "pitch": 1.0,
                              "r_max": 2, "s2": [1.0 / 64.0, 0.0, 0.0]},
        # tests/mesh_query_tests.rs:219-246: [0.1,1.9]^3 in a 4x3x2 box -> exact[0] = 8/24
        "two_by_two_cube": {"box": [[0.1, 0.1, 0.1], [1.9, 1.9, 1.9]], "domain": [4, 3, 2], "pitch": 1.0,
                            "s2_0": 8.0 / 24.0},
        # tests/measure_execution_tests.rs:139-153: a solid 32^3 box gives S2(r) = 1 for r <= 16
        "solid": {"domain": [32, 32, 32], "r_max": 16, "value": 1.0},
        # src/gpu/voxel.rs:732-737 analytic box occupancy x<4 and y<2 and z<2
        "box_occupancy": {"box": [[0.25, 0.25, 0.25], [4.25, 2.25, 2.25]], "domain": [6, 5, 4], "pitch": 1.0},
        # tests/gpu_s2_boundaries_tests.rs:26-35: [0.25,2.25]x[0.25,1.25]^2 in 4x3x2 -> x<2, y=0, z=0
        "translated_box": {"box": [[0.25, 0.25, 0.25], [2.25, 1.25, 1.25]], "domain": [4, 3, 2], "pitch": 1.0,
                           "origins": [0.0, 1e9]},
        # src/geometry/s2.rs:1892-1893
        "padded_fft_dims": [[[7, 11, 13], [15, 24, 25]], [[1, 1, 40], [1, 1, 80]]],
        # Random123 known-answer vectors for Philox4x32-10
        "philox4x32_10": [
            {"ctr": [0, 0, 0, 0], "key": [0, 0], "out": [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]},
            {"ctr": [0xFFFFFFFF] * 4, "key": [0xFFFFFFFF] * 2,
             "out": [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]},
            {"ctr": [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], "key": [0xA4093822, 0x299F31D0],
             "out": [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]},
        ],
        # exact_shell_work at the benchmark configurations (repo BASELINE.md table)
        "shell_work": [[[128, 128, 128], 32, 143760], [[256, 256, 256], 64, 1123772], [[96, 96, 96], 24, 61564]],
    }


def main():
    (HERE / "tiny_pair_counts.json").write_text(json.dumps(tiny_pair_counts()))
    (HERE / "shell_sizes_pitch1.json").write_text(json.dumps(shell_sizes()))
    (HERE / "spline_cases.json").write_text(json.dumps(spline_cases(), indent=1))
    (HERE / "known_answers.json").write_text(json.dumps(known_answers(), indent=1))


if __name__ == "__main__":
    main()
