"""The block rule behind the device fold, restated on the CPU and checked
against the sequential left fold of the reference's curve
(src/geometry/s2.rs:1187-1207: sum = sum + t in shell order).

Inside one binade of the running sum S (S = a*u, u = 2^(ilogb(S) - 52)), adding
t >= 0 rounds to (a + n + [f > 1/2])*u where t/u = n + f, unless f == 1/2 (a
tie: the even neighbour depends on a) or S leaves the binade.  So a block whose
terms have no tie, no t/u >= 2^43 and whose increments G keep a + G < 2^53
moves S to exactly (a + G)*u -- whatever a is.  The device's per-radius fold
commits such blocks with one add and runs every other block term by term; the
guess of each block's starting binade comes from plain (unordered) block sums.
This file restates that procedure and checks it bit for bit against the plain
sequential sum on adversarial segments (ties, binade crossings, zeros, wide
dynamic range, negative terms)."""
import math
import random

TWO53 = 2.0 ** 53
TWO43 = 2.0 ** 43
BLOCK = 128  # 64 lanes x 2 terms, as on the device


def seq_sum(ts, s=0.0):
    for t in ts:
        s = s + t
    return s


def ilogb(x):
    return math.frexp(x)[1] - 1


def block_increment(ts, m):
    """(G, flagged) of a block on the grid of binade m, as the device computes them."""
    inv_u = math.ldexp(1.0, 52 - m)
    g, flagged = 0.0, False
    for t in ts:
        y = t * inv_u
        big = not (y < TWO43)
        nfl = 0.0 if big else math.floor(y)
        f = y - nfl
        g += nfl + (1.0 if f > 0.5 else 0.0)
        flagged = flagged or big or f == 0.5
    return g, flagged


MAX_TIES = 4


def block_ties(ts, m):
    """(G, ties, big) of a block on the grid of binade m, as the device's
    increment pass computes them: G sums every term's increment with ties
    counted at n (their floor), ties lists (P, n) in shell order -- the
    increments of the block's terms before the tie, and the tie's floor."""
    inv_u = math.ldexp(1.0, 52 - m)
    g, ties, big = 0.0, [], False
    for t in ts:
        y = t * inv_u
        b = not (y < TWO43)
        big = big or b
        nfl = 0.0 if b else math.floor(y)
        f = y - nfl
        if not b and f == 0.5:
            ties.append((g, nfl))
        g += nfl + (1.0 if f > 0.5 else 0.0)
    return g, ties, big


def tie_block(a, g, ties):
    """S / u after a block with ties, from S / u = a before it: each tie rounds
    to the even neighbour of (a + increments before it + n)."""
    extra = 0.0
    for p, n in ties:
        x = a + p + extra + n
        extra += float(int(x) & 1)
    return a + g + extra


def wide_fold(ts, use_ties=True):
    """The device's per-radius fold: guessed binades, short path (blocks
    without ties, or with at most MAX_TIES of them) or term by term."""
    blocks = [ts[i:i + BLOCK] for i in range(0, len(ts), BLOCK)]
    bad = [any(not (0.0 <= t <= 1.7976931348623157e308) for t in b) for b in blocks]
    approx = [sum(b) for b in blocks]  # any order
    s, seq, short = 0.0, False, 0
    pre = 0.0
    guess = []
    for p in approx:
        guess.append(ilogb(pre) if pre > 0.0 else None)
        pre += p
    for b, m, bd in zip(blocks, guess, bad):
        if not seq and not bd and m is not None and s > 0.0 and ilogb(s) == m:
            g, flagged = block_increment(b, m)
            u = math.ldexp(1.0, m - 52)
            a = s / u
            if not flagged and a + g < TWO53:
                s = (a + g) * u
                short += 1
                continue
            g, ties, big = block_ties(b, m)
            if use_ties and not big and 0 < len(ties) <= MAX_TIES and a + (g + len(ties)) < TWO53:
                s = tie_block(a, g, ties) * u
                short += 1
                continue
        if bd:
            seq = True  # after a negative or non-finite term: one at a time (the same adds)
        s = seq_sum(b, s)
    return s, short


def segments(seed, count):
    rng = random.Random(seed)
    out = []
    for trial in range(count):
        kind = trial % 7
        n = rng.choice([1, 127, 128, 129, 256, 513, rng.randint(1, 6000)])
        if kind == 0:   # hits / valid ratios
            ts = [rng.randint(0, 10 ** 6) / rng.randint(10 ** 6, 2 * 10 ** 6) for _ in range(n)]
        elif kind == 1:  # dyadic values: ties on the grid
            ts = [rng.choice([0.0, 0.5, 0.25, 1.0, 0.125, 2.0 ** -40, 3 * 2.0 ** -41]) for _ in range(n)]
        elif kind == 2:  # wide dynamic range
            ts = [math.ldexp(rng.randint(0, 2 ** 20), -rng.randint(0, 60)) for _ in range(n)]
        elif kind == 3:  # leading zeros
            z = rng.randint(0, n)
            ts = [0.0] * z + [rng.random() for _ in range(n - z)]
        elif kind == 4:  # a negative term
            ts = [rng.random() for _ in range(n)]
            ts[rng.randrange(n)] = -rng.random()
        elif kind == 5:  # terms that push S across many binades
            ts = [rng.random() * 2.0 ** (i // 300) for i in range(n)]
        elif kind == 6 and trial % 2:  # hits / valid with power-of-two valid counts: ties
            ts = [rng.randint(0, 2 ** 12) / 2.0 ** rng.randint(8, 14) for _ in range(n)]
        else:
            ts = [rng.random() * 1e-3 for _ in range(n)]
        out.append(ts)
    return out


def test_block_rule_is_exact_whatever_the_running_sum():
    rng = random.Random(11)
    checked = 0
    for _ in range(400):
        m = rng.randint(-20, 20)
        s = math.ldexp(1.0 + rng.random(), m)  # any S in binade m
        ts = [rng.random() * math.ldexp(1.0, m - rng.randint(9, 30)) for _ in range(rng.randint(1, BLOCK))]
        g, flagged = block_increment(ts, m)
        u = math.ldexp(1.0, m - 52)
        if flagged or s / u + g >= TWO53:
            continue
        assert seq_sum(ts, s) == (s / u + g) * u
        checked += 1
    assert checked > 300


def test_wide_fold_equals_sequential_fold():
    short = 0
    for ts in segments(5, 140):
        got, k = wide_fold(ts)
        want = seq_sum(ts)
        assert math.copysign(1.0, got) == math.copysign(1.0, want) and got == want
        short += k
    assert short > 10  # the short path is taken, not only the term-by-term one


def test_tie_rule_is_exact_whatever_the_running_sum():
    # blocks of dyadic terms that land exactly halfway on the grid, from any S
    rng = random.Random(17)
    checked = 0
    for _ in range(600):
        m = rng.randint(-10, 10)
        s = math.ldexp(1.0 + rng.random(), m)
        u = math.ldexp(1.0, m - 52)
        ts = []
        for _ in range(rng.randint(1, BLOCK)):
            if rng.random() < 0.03:
                ts.append((rng.randint(0, 2 ** 20) + 0.5) * u)      # a tie
            else:
                ts.append(rng.random() * math.ldexp(1.0, m - rng.randint(9, 30)))
        g, ties, big = block_ties(ts, m)
        a = s / u
        if big or not ties or len(ties) > MAX_TIES or a + g + len(ties) >= TWO53:
            continue
        assert seq_sum(ts, s) == tie_block(a, g, ties) * u
        checked += 1
    assert checked > 150


def test_dropping_the_parity_breaks_it():
    # the tie rule is not the floor: a planted mistake (ties never round up) fails
    rng = random.Random(5)
    wrong = 0
    for _ in range(300):
        m = 0
        u = math.ldexp(1.0, m - 52)
        s = math.ldexp(1.0 + rng.random(), m)
        ts = [(rng.randint(0, 2 ** 20) + 0.5) * u for _ in range(3)]
        g, ties, big = block_ties(ts, m)
        if (s / u + g) * u != seq_sum(ts, s):
            wrong += 1
    assert wrong > 50


def test_long_shell_takes_the_short_path():
    # 51,000 ratios like the largest r = 64 shell at 256^3
    rng = random.Random(3)
    ts = [rng.randint(0, 10 ** 6) / rng.randint(10 ** 7, 17 * 10 ** 6) for _ in range(51000)]
    got, short = wide_fold(ts)
    assert got == seq_sum(ts)
    assert short >= 320  # of 399 blocks
