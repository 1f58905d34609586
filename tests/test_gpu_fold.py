"""The exact curve's fold on the device (64 terms per step on the binade grid)
against the sequential left fold it must reproduce bit for bit: adversarial
segments with ties, binade crossings, zeros, wide dynamic range, negative
terms and long shells.  Requires the GPU (marker gpu)."""
import math
import random

import numpy as np
import pytest

from rustmspt_amd import s2

pytestmark = pytest.mark.gpu


def seq_sum(ts):
    s = 0.0
    for t in ts:
        s = s + float(t)
    return s


def segments(seed):
    rng = random.Random(seed)
    out = []
    for trial in range(240):
        kind = trial % 8
        n = rng.choice([1, 2, 63, 64, 65, 511, 512, 513, rng.randint(1, 3000)])
        if kind == 0:
            ts = [rng.random() for _ in range(n)]
        elif kind == 1:   # hits / valid style ratios
            ts = [min(1.0, rng.randint(0, 50) / rng.randint(1, 60)) for _ in range(n)]
        elif kind == 2:   # dyadic values: exact ties on the grid
            ts = [rng.choice([0.0, 0.5, 0.25, 1.0, 0.125, 2.0 ** -40, 3 * 2.0 ** -41]) for _ in range(n)]
        elif kind == 3:   # wide dynamic range
            ts = [math.ldexp(rng.randint(0, 2 ** 20), -rng.randint(0, 60)) for _ in range(n)]
        elif kind == 4:
            ts = [rng.random() * 2.0 ** rng.randint(-30, 5) for _ in range(n)]
        elif kind == 5:   # leading zeros, then values
            z = rng.randint(0, n)
            ts = [0.0] * z + [rng.random() for _ in range(n - z)]
        elif kind == 6:   # a negative term switches to one-at-a-time
            ts = [rng.random() for _ in range(n)]
            ts[rng.randrange(n)] = -rng.random()
        else:
            ts = [rng.randint(0, 10 ** 6) / rng.randint(1, 10 ** 6 + 1) for _ in range(n)]
        out.append(ts)
    return out


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_segments_equal_sequential_fold(ctx, seed):
    segs = segments(seed)
    terms = np.concatenate([np.asarray(s, np.float64) for s in segs])
    starts = np.concatenate([[0], np.cumsum([len(s) for s in segs])]).astype(np.uint64)
    got = s2.ordered_sum(ctx, terms, starts)
    want = np.array([seq_sum(s) for s in segs])
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


def test_long_shell_like_segment(ctx):
    # 51,000 ratios like the largest r = 64 shell at 256^3, plus an empty segment
    rng = np.random.default_rng(0)
    valid = rng.integers(10 ** 6, 1.7 * 10 ** 7, 51000)
    hits = (rng.random(51000) * 0.4 * valid).astype(np.int64)
    terms = hits / valid
    got = s2.ordered_sum(ctx, terms, np.array([0, 51000, 51000], np.uint64))
    assert got[0] == seq_sum(terms) and got[1] == 0.0
