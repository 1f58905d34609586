"""Serving loop: every result equals the same request run alone and the plain
entry point, whatever else is in flight; admission respects the device budget.
Requires the GPU (marker gpu)."""
import numpy as np
import pytest

from rustmspt_amd import s2, serve, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pool():
    rng = synth.rng_for(5, 1)
    out = []
    for n in (40, 52, 64, 80, 96, 45):
        occ = serve.sphere_packing(n, float(rng.uniform(0.15, 0.45)), rng)
        out.append((occ, serve.pack(occ)))
    return out


REQS = [(0, 8, 0), (1, 12, 1), (2, 16, 0), (3, 20, 0), (4, 24, 1), (5, 10, 0), (2, 30, 1), (4, 40, 0),
        (0, 39, 0), (1, 5, 1)]


def test_alone_equals_mix_equals_plain(ctx, pool):
    alone = []
    with serve.Server(slots=4) as sv:
        for v, r, m in REQS:
            occ, words = pool[v]
            t = sv.submit(words, occ.shape, r, m)
            info, curve = sv.poll(t, wait=True)
            alone.append(curve)
    with serve.Server(slots=4) as sv:
        tickets = []
        for k in range(3):  # filler traffic around the measured requests
            for v, r, m in REQS[::-1]:
                occ, words = pool[v]
                sv.submit(words, occ.shape, r, 1 - m)
            for v, r, m in REQS:
                occ, words = pool[v]
                tickets.append(sv.submit(words, occ.shape, r, m))
        mix = [sv.poll(t, wait=True)[1] for t in tickets]
        st = sv.stats()
    for k in range(3):
        for a, b in zip(alone, mix[k * len(REQS):(k + 1) * len(REQS)]):
            assert np.array_equal(a, b)
    for (v, r, m), curve in zip(REQS, alone):
        vol = s2.Volume.from_bytes(ctx, pool[v][0])
        plain = vol.exact_s2(r) if m == serve.EXACT else vol.fft32_s2(r)
        assert np.array_equal(plain, curve)
        vol.free()
    assert st["failed"] == 0


def test_budget_defers_and_drains(ctx):
    # 160^3 r40 exact requests.  Under a budget B below 1 GiB four slots keep
    # a 64 MiB plan allowance and a B/16 FFT arena each, and requests are
    # admitted against the rest, 3B/4 - 256 MiB: B is set so that two
    # footprints fit, so with four slots free the third waiting request must be
    # deferred
    occ = serve.sphere_packing(160, 0.3, synth.rng_for(11, 2))
    words = serve.pack(occ)
    with serve.Server(slots=1) as sv:
        fp = sv.poll(sv.submit(words, occ.shape, 40, serve.EXACT), wait=True)[0]["footprint"]
    budget = int(((256 << 20) + 2.5 * fp) / 0.75)
    assert budget < 1 << 30
    with serve.Server(slots=4, budget_bytes=budget) as sv:
        tickets = [sv.submit(words, occ.shape, 40, serve.EXACT) for _ in range(12)]
        res = [sv.poll(t, wait=True) for t in tickets]
        st = sv.stats()
    infos = [i for i, _ in res]
    assert all(i["footprint"] == fp for i in infos)
    assert st["completed"] == 12 and st["failed"] == 0
    assert st["deferred"] > 0 and any(i["deferrals"] > 0 for i in infos), st
    assert st["peak_bytes"] <= st["budget"] == budget
    # device memory, not just the estimates: no slot held more than its share
    # (footprint + plan allowance + arena), and between requests the slots hold
    # at most their allowances
    assert st["over_budget_requests"] == 0, st
    assert st["device_bytes_held"] <= budget, st
    assert all(i["t_done"] >= i["t_start"] >= i["t_submit"] for i in infos)
    vol = s2.Volume.from_bytes(ctx, occ)
    exact = vol.exact_s2(40)
    vol.free()
    for _, curve in res:
        assert np.array_equal(curve, exact)
    with serve.Server(slots=2, budget_bytes=64 << 20) as sv:
        big = serve.pack(np.ones((200, 200, 200), np.uint8))
        with pytest.raises(Exception):
            sv.submit(big, (200, 200, 200), 50, serve.FFT32)


def test_stream_reports_percentiles(pool):
    vols = [serve.Volume(o.shape, w, float(o.mean())) for o, w in pool]
    sched = serve.poisson_schedule(100.0, 0.5, vols, seed=3)
    with serve.Server(slots=4) as sv:
        rep = serve.run_stream(sv, vols, sched, timeout=120)
    assert rep["requests"] == len(sched) and rep["failed"] == 0
    assert 0 < rep["latency_p50_ms"] <= rep["latency_p99_ms"] <= rep["latency_max_ms"]
