"""The allocation-free Monte Carlo entry point (s2h_mc_counts_noalloc) through
the C ABI: the same counts as s2h_mc_counts, bit for bit, on two calls in a row
(the second smaller), with the device memory the context holds unchanged by
the calls; shells beyond the scratch are refused before any launch.  Requires
the GPU (marker gpu)."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as o
from rustmspt_amd import _lib as L
from rustmspt_amd import s2, synth

pytestmark = pytest.mark.gpu


def noalloc_counts(ctx, vol, r_max, pitch, samples, seed):
    out = np.zeros(2 * (r_max + 1), np.uint64)
    L.check(L.lib().s2h_mc_counts_noalloc(ctx.handle, vol.handle, r_max, float(pitch), samples, seed,
                                          L.ptr(out, L._u64p), len(out)))
    return out.reshape(-1, 2)


def test_noalloc_equals_mc_counts_twice_then_smaller():
    with s2.Context(0) as ctx:
        big = synth.grf_volume(96, 4.0, 0.3, seed=41)
        small = (np.random.default_rng(2).random((30, 23, 41)) < 0.4).astype(np.uint8)
        vb = s2.Volume.from_bytes(ctx, big)
        vs = s2.Volume.from_bytes(ctx, small)
        held = ctx.bytes_in_use
        first = noalloc_counts(ctx, vb, 64, 1.0, 20000, 5)
        second = noalloc_counts(ctx, vs, 9, 0.75, 3000, 2 ** 40 + 7)
        assert ctx.bytes_in_use == held          # no device allocation by the calls
        assert np.array_equal(first, vb.mc_counts(64, 1.0, 20000, 5))
        assert np.array_equal(second, vs.mc_counts(9, 0.75, 3000, 2 ** 40 + 7))
        # and the oracle's Philox stream for the small case
        assert np.array_equal(second, o.mc_counts(o.pack(small), small.shape, 9, 0.75, 3000, 2 ** 40 + 7))
        # the second call left nothing of the first behind: radii above 9 do not exist there
        assert second.shape == (10, 2) and second[0].tolist() == [0, 0]
        # again after a larger call, and r_max = 0
        assert np.array_equal(noalloc_counts(ctx, vb, 96, 1.0, 4096, 9), vb.mc_counts(96, 1.0, 4096, 9))
        assert np.array_equal(noalloc_counts(ctx, vs, 0, 1.0, 100, 1), np.zeros((1, 2), np.uint64))
        vb.free()
        vs.free()


def test_noalloc_refuses_shells_beyond_the_scratch():
    with s2.Context(0) as ctx:
        vol = s2.Volume.from_bytes(ctx, np.ones((8, 8, 8), np.uint8))
        out = np.zeros(2 * 129, np.uint64)
        st = L.lib().s2h_mc_counts_noalloc(ctx.handle, vol.handle, 128, 1.0, 100, 1, L.ptr(out, L._u64p), len(out))
        assert st == 3 and b"no-allocation scratch" in L.lib().s2h_last_error()
        out = np.zeros(2 * 97, np.uint64)
        st = L.lib().s2h_mc_counts_noalloc(ctx.handle, vol.handle, 96, 0.5, 100, 1, L.ptr(out, L._u64p), len(out))
        assert st == 3
        vol.free()
