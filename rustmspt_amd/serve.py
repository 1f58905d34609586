"""Serving front end (s2h_serve_*): submit S2 requests as they arrive, collect
results with their admission and completion times, and report latency
percentiles beside throughput.

Workload generator for the served stream: Poisson arrivals; each request is a
seeded overlapping-sphere packing with grid side log-uniform in [64, 384],
volume fraction uniform in [0.1, 0.5], r_max uniform in {16..96} (clamped to
the grid), 70 % exact / 30 % FP32 FFT.  Volumes come from a pre-generated pool
of such packings so host-side generation stays out of the measured stream.
"""
from __future__ import annotations

import ctypes as C
import math
import threading
import time
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import _lib as L

EXACT, FFT32 = 0, 1


class Server:
    def __init__(self, device: int = 0, budget_bytes: int = 0, slots: int = 4):
        h = C.c_void_p()
        L.check(L.lib().s2h_serve_create(int(device), int(budget_bytes), int(slots), C.byref(h)))
        self.handle = h
        self._keep = {}

    def submit(self, words: np.ndarray, dims, r_max: int, method: int, pitch: float = 1.0, tag: int = 0) -> int:
        words = np.ascontiguousarray(words, np.uint64)
        req = L.Request()
        req.words = L.ptr(words, L._u64p)
        req.nwords = len(words)
        for a in range(3):
            req.dims[a] = int(dims[a])
        req.r_max = int(r_max)
        req.method = int(method)
        req.pitch = float(pitch)
        req.tag = int(tag)
        t = C.c_uint64()
        L.check(L.lib().s2h_serve_submit(self.handle, C.byref(req), C.byref(t)))
        self._keep[t.value] = int(r_max)
        return t.value

    def poll(self, ticket: int, wait: bool = False):
        res = L.Result()
        n = self._keep[ticket] + 1
        out = np.zeros(n, np.float64)
        L.check(L.lib().s2h_serve_poll(self.handle, ticket, 1 if wait else 0, C.byref(res), L.ptr(out, L._f64p), n))
        info = {"ready": bool(res.ready), "tag": res.tag, "deferrals": res.deferrals, "footprint": res.footprint,
                "t_submit": res.t_submit, "t_start": res.t_start, "t_done": res.t_done}
        if res.ready:
            self._keep.pop(ticket, None)
            return info, out
        return info, None

    def stats(self) -> dict:
        s = np.zeros(10, np.uint64)
        L.check(L.lib().s2h_serve_stats(self.handle, L.ptr(s, L._u64p), 10))
        keys = ["admitted", "deferred", "completed", "failed", "waiting", "bytes_in_use", "peak_bytes", "budget",
                "device_bytes_held", "over_budget_requests"]
        return {k: int(v) for k, v in zip(keys, s)}

    def close(self):
        if self.handle:
            L.lib().s2h_serve_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def sphere_packing(n: int, vf: float, rng: np.random.Generator) -> np.ndarray:
    """Overlapping random balls (radius U[n/24, n/10] voxels) added until the
    occupied fraction reaches vf; uint8 grid, x-major, z fastest."""
    occ = np.zeros((n, n, n), np.uint8)
    target = vf * occ.size
    filled = 0
    ax = np.arange(n, dtype=np.float32) + 0.5
    while filled < target:
        r = rng.uniform(n / 24.0, n / 10.0)
        c = rng.uniform(0, n, 3)
        lo = np.maximum(np.floor(c - r).astype(int), 0)
        hi = np.minimum(np.ceil(c + r).astype(int) + 1, n)
        if np.any(hi <= lo):
            continue
        dx = (ax[lo[0]:hi[0]] - c[0]) ** 2
        dy = (ax[lo[1]:hi[1]] - c[1]) ** 2
        dz = (ax[lo[2]:hi[2]] - c[2]) ** 2
        ball = (dx[:, None, None] + dy[None, :, None] + dz[None, None, :]) <= r * r
        sub = occ[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
        filled += int(np.count_nonzero(ball & (sub == 0)))
        sub |= ball.astype(np.uint8)
    return occ


def pack(occ: np.ndarray) -> np.ndarray:
    nx, ny, nz = occ.shape
    rw = (nz + 63) // 64 + 1
    padded = np.zeros((nx, ny, rw * 64), np.uint8)
    padded[:, :, :nz] = occ
    bits = np.packbits(padded.reshape(nx, ny, rw * 8, 8)[..., ::-1], axis=-1).reshape(nx, ny, rw * 8)
    return np.ascontiguousarray(bits.view(np.uint8)).view("<u8").reshape(-1).copy()


@dataclass
class Volume:
    dims: tuple
    words: np.ndarray
    vf: float


def volume_pool(count: int, seed: int, nmin: int = 64, nmax: int = 384) -> List[Volume]:
    from .synth import rng_for
    rng = rng_for(seed, 41)
    pool = []
    for _ in range(count):
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        vf = float(rng.uniform(0.1, 0.5))
        occ = sphere_packing(n, vf, rng)
        pool.append(Volume((n, n, n), pack(occ), float(occ.mean())))
    return pool


@dataclass
class Arrival:
    t: float
    volume: int
    r_max: int
    method: int


def poisson_schedule(rate: float, duration: float, pool: List[Volume], seed: int) -> List[Arrival]:
    from .synth import rng_for
    rng = rng_for(seed, 43)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return out
        v = int(rng.integers(0, len(pool)))
        r = int(rng.integers(16, 97))
        r = max(1, min(r, pool[v].dims[0] - 1))
        m = EXACT if rng.random() < 0.7 else FFT32
        out.append(Arrival(t, v, r, m))


def run_stream(server: Server, pool: List[Volume], schedule: List[Arrival], timeout: float = 600.0) -> dict:
    """Submit at the scheduled arrival times; collect every result; latency =
    completion - submission on the server's clock."""
    tickets: List[Optional[int]] = [None] * len(schedule)
    done = []
    start = time.perf_counter()

    def submitter():
        for i, a in enumerate(schedule):
            delay = a.t - (time.perf_counter() - start)
            if delay > 0:
                time.sleep(delay)
            v = pool[a.volume]
            tickets[i] = server.submit(v.words, v.dims, a.r_max, a.method, tag=i)

    st0 = server.stats()
    th = threading.Thread(target=submitter)
    th.start()
    pending = set()
    deadline = start + timeout
    i_seen = 0
    while True:
        while i_seen < len(schedule) and tickets[i_seen] is not None:
            pending.add(i_seen)
            i_seen += 1
        for i in list(pending):
            info, curve = server.poll(tickets[i], wait=False)
            if info["ready"]:
                done.append((i, info))
                pending.discard(i)
        if i_seen == len(schedule) and not pending:
            break
        if time.perf_counter() > deadline:
            raise TimeoutError("served stream did not drain")
        time.sleep(0.002)
    th.join()
    lat = np.array([info["t_done"] - info["t_submit"] for _, info in done])
    busy = float(sum(info["t_done"] - info["t_start"] for _, info in done))
    t_first = min(info["t_submit"] for _, info in done)
    t_last = max(info["t_done"] for _, info in done)
    st = server.stats()
    exact = sum(1 for i, _ in done if schedule[i].method == EXACT)
    return {"requests": len(done), "exact": exact, "fft32": len(done) - exact,
            "offered_rate": len(schedule) / (schedule[-1].t if schedule else 1.0),
            "throughput_evals_per_s": len(done) / max(t_last - t_first, 1e-9),
            "latency_p50_ms": float(np.percentile(lat, 50) * 1e3), "latency_p99_ms": float(np.percentile(lat, 99) * 1e3),
            "latency_max_ms": float(lat.max() * 1e3), "slot_busy_s": busy, "makespan_s": t_last - t_first,
            "admitted": st["admitted"] - st0["admitted"], "deferred": st["deferred"] - st0["deferred"],
            "failed": st["failed"] - st0["failed"], "peak_bytes": st["peak_bytes"], "budget": st["budget"]}
