"""Voxel Monte Carlo on the GPU: integer counts equal the oracle's Philox stream
exactly, and estimates meet the family-wise 3-sigma rule (stats_rules) against
the exact estimand
sum_d C(d) / sum_d valid(d) over each full shell.  Requires the GPU."""
import numpy as np
import pytest

from oracle import oracle as o
from rustmspt_amd import s2, synth
from stats_rules import mc_family_rule

pytestmark = pytest.mark.gpu


def test_counts_equal_oracle_stream(ctx):
    rng = np.random.default_rng(4)
    occ = (rng.random((30, 23, 70)) < 0.35).astype(np.uint8)
    vol = s2.Volume.from_bytes(ctx, occ)
    words = o.pack(occ)
    for r_max, pitch, samples, seed in [(6, 1.0, 20000, 1), (9, 0.5, 5000, 2 ** 40 + 3), (4, 1.5, 777, 0)]:
        got = vol.mc_counts(r_max, pitch, samples, seed)
        want = o.mc_counts(words, occ.shape, r_max, pitch, samples, seed)
        assert np.array_equal(got, want), (r_max, pitch)


def test_within_three_sigma_of_exact_estimand(ctx):
    occ = synth.grf_volume(48, 3.0, 0.3, seed=21)
    vol = s2.Volume.from_bytes(ctx, occ)
    r_max, samples = 12, 200000
    hv = vol.mc_counts(r_max, 1.0, samples, seed=12345)
    words = o.pack(occ)
    z = []
    for r in range(1, r_max + 1):
        shell = o.shell_offsets(float(r), 0.5)
        inside = np.all(np.abs(shell) < np.array(occ.shape), axis=1)
        h = o.hits_list(words, occ.shape, shell[inside]).astype(float)
        valid = np.prod(np.array(occ.shape)[None, :] - np.abs(shell[inside]), axis=1).astype(float)
        p = h.sum() / valid.sum()
        n_valid = hv[r, 1]
        est = hv[r, 0] / n_valid
        sigma = np.sqrt(p * (1 - p) / n_valid)
        z.append((est - p) / sigma)
    print(f"max |z| = {np.max(np.abs(z)):.2f}")
    mc_family_rule(z)


def test_curve_shape(ctx):
    occ = np.zeros((20, 20, 20), np.uint8)
    occ[5:15, 5:15, 5:15] = 1
    vol = s2.Volume.from_bytes(ctx, occ)
    c = vol.mc_s2(5, 1.0, 50000, 7)
    assert c[0] == occ.mean() and np.all(np.diff(c[1:]) < 0.02) and np.all(c <= c[0] + 0.02)
