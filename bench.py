"""Benchmark: exact-S2 evaluations per second at 256^3, r_max = 64 on one MI355X.

A step is one exact S2 evaluation (exact pair counts of all 1,123,772 in-domain
shell offsets + the bit-exact f64 curve) of a device-resident packed volume --
the reference's VoxelS2::calculate(r_max, "exact") on a prepared grid.  The
volume is a seeded synthetic packing of 500 particles (250 level-3 icospheres,
250 ellipsoids with axis ratios U[1,3] and random orientation) voxelized on the
device before the timed region.

Prints one JSON line.  Options: --gpus (only 1 is supported), --steps, --warmup,
--cpu-seconds (target length of the CPU baseline sample), --no-cpu.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# int32 lane-ops/s at one wave64 VALU instruction per 4 cycles per SIMD: 256 CUs x
# 64 lanes x 2.4 GHz.  The counting loop's own instruction pair issues faster than
# that (the on-device probe measures it), so the roofline's peak is the probe's
# figure and this one is reported beside it.
VALU_4CYCLE_OPS = 256 * 64 * 2.4e9
# HBM bytes per evaluation, from rocprofv3 --pmc on this workload (256^3,
# r_max 64, the 500-particle packing): reads = 32 B x TCC_EA0_RDREQ_32B + 64 B x
# TCC_EA0_RDREQ_64B + 128 B x TCC_EA0_RDREQ_128B, writes = WRITE_SIZE.  On this
# part FETCH_SIZE counts a 128-byte read request as 64 bytes: a 512 MiB device
# copy shows 4,194,304 128-byte requests (512 MiB) and FETCH_SIZE 262,156 KiB,
# WRITE_SIZE 524,288 KiB; the request counts give the bytes read directly.
# Direct counting (exact_count_kernel, launch shape band 3 / 65536 waves):
COUNT_PHASE_TRAFFIC = 64 * 59351.5 + 128 * 177855 + 909816 * 1024.0
# The FP64 FFT passes (z fwd, y fwd, x fwd + |F|^2 + inv, y inv, z inv; the two-stage
# kernels at 320 = 16 x 20), steady-state dispatches (the tile choice's trial runs
# excluded), read then write MB per pass:
FFT64_PHASE_TRAFFIC = (3.3 + 169.9 + 176.3 + 212.3 + 220.3 + 53.9 + 56.0 + 21.7 + 22.7 + 8.7) * 1e6
# The FP32 FFT passes, the same five:
FFT32_PHASE_TRAFFIC = (3.2 + 86.0 + 92.3 + 107.5 + 115.4 + 27.3 + 29.4 + 11.0 + 11.9 + 4.4) * 1e6


def fft64_bytes(dims, offs, r_max, complex_bytes=16.0):
    """(unpruned, compulsory) HBM bytes of one FFT evaluation (complex128
    spectra by default; 8.0 for the FP32 path's complex64), on the kernels'
    padded lengths f_a >= n_a + R_a (R_a: the largest |offset| on axis a).
    Unpruned: 12 x f_x f_y (f_z/2+1) x c -- three axes, forward and inverse,
    read and write of the half spectrum of a transform that keeps every
    output.  Compulsory (the roofline's algorithmic bytes): what the pruned
    passes must move -- each pass reads its input and writes its output once
    (z forward: packed rows in, n_x n_y rows of half spectra out; y forward:
    n_x planes in, f_y rows each out; x forward + |F|^2 + inverse keeping the
    R_x+1 planes dx >= 0 -- the correlation is even, C(d) = C(-d); y inverse
    keeping 2R_y+1 rows; z inverse keeping 2R_z+1 real values)."""
    import numpy as np

    from rustmspt_amd import s2

    nx, ny, nz = dims
    fx, fy, fz = s2.fft_padded_dims(dims, r_max, 1.0)
    h = fz // 2 + 1
    rx, ry, rz = (int(v) for v in np.abs(offs).max(axis=0))
    c = float(complex_bytes)
    row_words = (nz + 63) // 64 + 1
    z_fwd = nx * ny * row_words * 8 + nx * ny * h * c
    y_fwd = nx * ny * h * c + nx * fy * h * c
    x_pow = nx * fy * h * c + (rx + 1) * fy * h * c
    y_inv = (rx + 1) * fy * h * c + (rx + 1) * (2 * ry + 1) * h * c
    z_inv = (rx + 1) * (2 * ry + 1) * h * c + (rx + 1) * (2 * ry + 1) * (2 * rz + 1) * c / 2.0
    return 12.0 * fx * fy * h * c, float(z_fwd + y_fwd + x_pow + y_inv + z_inv)


def exact_timed(ctx, vol, r_max, method, steps, warmup):
    """Evaluations of the exact curve with one kernel choice ("auto": the
    model's pick): evals/s of the plain calls as a caller makes them (no
    profiling events), then count-phase and device ms from HIP events over
    further profiled calls, and the curve."""
    import numpy as np

    ctx.set_exact_method(method)
    ctx.profile(False)
    curve = None
    for _ in range(warmup):
        curve = vol.exact_s2(r_max, 1.0)
    t0 = time.perf_counter()
    for _ in range(steps):
        curve = vol.exact_s2(r_max, 1.0)
    el = time.perf_counter() - t0
    ctx.profile(True)
    count_ms, dev_ms = [], []
    for _ in range(max(3, min(steps, 10))):
        again = vol.exact_s2(r_max, 1.0)
        p = ctx.profile_read()
        count_ms.append(p["count_ms"])
        dev_ms.append(p["device_ms"])
    assert again.tobytes() == curve.tobytes()
    ctx.set_exact_method("auto")
    return {"evals_per_s": steps / el, "ms_per_step": 1e3 * el / steps, "count_ms": float(np.mean(count_ms)),
            "device_ms": float(np.mean(dev_ms)), "launches": ctx.profile_read()["launches"], "curve": curve}


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(occ_words, dims, r_max, offs, pair_work_total, target_s):
    """Time the oracle's packed direct kernel (C restatement of
    direct_pair_counts_bits, OpenMP over offsets) on every k-th offset in shell
    order and extrapolate by pair work -- with all the threads OpenMP gets on
    this host, then with 8 threads.  The reference's own planner picks this
    direct kernel at 256^3 r_max=64 (its FFT working set exceeds the 768 MiB
    CPU budget)."""
    import numpy as np

    from oracle import oracle as o

    valid = np.prod(np.array(dims, dtype=np.float64)[None, :] - np.abs(offs), axis=1)

    def timed(k):
        sample = offs[::k]
        t0 = time.perf_counter()
        o.hits_list(occ_words, dims, sample)
        return time.perf_counter() - t0, float(valid[::k].sum()), len(sample)

    # calibrate: a small stride-sample first, then size the real sample
    dt, pairs, _ = timed(4001)
    rate = pairs / max(dt, 1e-4)
    k = max(1, int(pair_work_total / max(rate * target_s, 1.0)))
    dt, pairs, n = timed(k)
    rate = pairs / dt
    threads = o.num_threads()
    out = {
        "value": rate / pair_work_total,
        "unit": "evals/s",
        "cores": threads,
        "kind": "port",
        "sample": f"every {k}th of {len(offs)} in-domain offsets in shell order ({n} offsets, "
                  f"{pairs:.3e} of {pair_work_total:.3e} pair-tests) timed in {dt:.1f} s, extrapolated by pair work",
        "seconds_per_eval": pair_work_total / rate,
        "cpu_model": _cpu_model(),
        "host_cpus": os.cpu_count(),
    }
    if threads != 8:
        o.set_threads(8)
        k8 = max(1, int(k * threads / 8))
        dt8, pairs8, n8 = timed(k8)
        o.set_threads(threads)
        out["seconds_per_eval_8_threads"] = pair_work_total / (pairs8 / dt8)
    return out


def unpack_occupancy(words, dims):
    """uint8 grid from packed rows (row_words = nz / 64 + 1 words, bit z % 64 of word z / 64)."""
    import numpy as np

    nx, ny, nz = dims
    rw = (nz + 63) // 64 + 1
    bits = np.unpackbits(np.ascontiguousarray(words, "<u8").view(np.uint8), bitorder="little")
    return np.ascontiguousarray(bits.reshape(nx, ny, rw * 64)[:, :, :nz])


def cpu_legs(ctx, c2_words, c2_dims, c2_r, c2_hits, ct, target_s):
    """The further CPU timings beside the GPU: (1) the f64 FFT kernel forced at
    the single large measure, where the reference's planner would not take it
    (its working set exceeds the 768 MiB budget) -- autocorrelation on the
    reference's padding, counts at every shell offset, the curve; (2) the
    plain measure (64 icospheres, 128^3, r_max 32) with the kernel the
    reference's planner picks (the FFT there), all threads and 8; (3) the
    CT-like volume (768^3, r_max 128): the direct kernel on every 997th
    in-domain offset in shell order, extrapolated by pair work.  Each leg
    checks its counts or curve against the GPU's."""
    import numpy as np

    from oracle import oracle as o
    from rustmspt_amd import s2, synth

    out = {}
    threads = o.num_threads()
    offs, rs = s2.shell_order_offsets(c2_dims, c2_r, 1.0)
    occ = unpack_occupancy(c2_words, c2_dims)
    padded = o.padded_fft_dims(c2_dims)
    t0 = time.perf_counter()
    corr = o.fft_autocorr(occ, padded)
    hits = o.fft_hits(corr, offs)
    del corr
    vf = float(occ.sum()) / occ.size
    o.curve_from_hits(c2_dims, c2_r, offs, rs, hits, vf)
    dt = time.perf_counter() - t0
    plan = o.plan_exact_cpu(c2_dims, c2_r, 1.0, threads)
    out["single_large_fft_forced"] = {
        "seconds_per_eval": dt, "threads": threads, "padded": padded, "counts_equal_gpu": bool(np.array_equal(hits, c2_hits)),
        "planner_pick": plan["method"], "planner_fft_bytes": plan["fft_bytes"],
        "note": "f64 FFT kernel of the oracle (mixed-radix 2,3,5, OpenMP over lines) on the reference's padding"}
    progress(f"cpu: forced FFT at {c2_dims[0]}^3 {dt:.1f} s")
    del occ, hits
    # plain measure with the planner's kernel
    mesh, bbox, _ = synth.config_plain(seed=1, n=64, grid=128)
    occ1 = o.voxelize(mesh.vertices, mesh.faces, [0.0, 0.0, 0.0], 1.0, [128, 128, 128])
    vol = s2.Volume.from_bytes(ctx, occ1)
    gpu_curve = vol.exact_s2(32, 1.0)
    vol.free()
    offs1, rs1 = s2.shell_order_offsets([128] * 3, 32, 1.0)
    leg = {}
    for th in sorted({threads, 8}):
        o.set_threads(th)
        pick = o.plan_exact_cpu([128] * 3, 32, 1.0, th)
        t0 = time.perf_counter()
        if pick["method"] == "fft":
            corr = o.fft_autocorr(occ1)
            h1 = o.fft_hits(corr, offs1)
            del corr
            curve = o.curve_from_hits([128] * 3, 32, offs1, rs1, h1, float(occ1.mean()))
        else:
            curve = o.exact_s2(occ1, 32)
        leg[f"seconds_per_eval_{th}_threads"] = time.perf_counter() - t0
        leg["planner_pick"] = pick["method"]
        leg["curve_equal_gpu"] = bool(np.array_equal(curve, gpu_curve))
    o.set_threads(threads)
    out["plain_measure_planner"] = leg
    progress("cpu: plain measure done")
    # CT-like volume: stride sample of the direct kernel
    if ct is not None:
        ct_words, ct_dims, ct_r, ct_hits_of = ct
        offs5, _ = s2.shell_order_offsets(ct_dims, ct_r, 1.0)
        valid = np.prod(np.array(ct_dims, dtype=np.float64)[None, :] - np.abs(offs5), axis=1)
        idx = np.arange(0, len(offs5), 997)
        t0 = time.perf_counter()
        h5 = o.hits_list(ct_words, ct_dims, offs5[idx])
        dt = time.perf_counter() - t0
        W, w = float(valid.sum()), float(valid[idx].sum())
        out["ct_volume_direct_sample"] = {
            "seconds_per_eval": dt * W / w, "sample_seconds": dt, "threads": threads,
            "sample": f"every 997th of {len(offs5)} in-domain offsets in shell order ({len(idx)} offsets, "
                      f"{w:.3e} of {W:.3e} pair-tests), extrapolated by pair work",
            "counts_equal_gpu": bool(np.array_equal(h5, ct_hits_of(idx)))}
        progress(f"cpu: CT volume sample {dt:.1f} s")
    return out


def batched_leg(ctx, ncand: int, seed: int, plain_per_grid: int = 8):
    """256 one-move candidates over 4 resident seed packings (grids 96, 128, 192,
    256; 150-400 ellipsoids each; r_max = grid/4), evaluated in one batch, vs the
    plain per-call path (s2h_exact_s2 on each candidate's resident volume: the
    candidate volumes are voxelized outside the timed region, and each grid's
    plan is built and tuned before the timing).  Per seed: its candidates alone
    as one batch, and the plain path on a sample of them."""
    import numpy as np

    from rustmspt_amd import batch, s2, synth

    packings, moves = synth.config_batched(seed, ncand)
    seeds = [batch.Seed(ctx, [p.mesh() for p in ps], s2.BoundingBox.from_size((g, g, g)), 1.0, g // 4)
             for g, ps in packings]
    parts = [ps for _, ps in packings]
    # first call: code objects, graph capture and the one-time row-chunk tuning
    t0 = time.perf_counter()
    batch.batch_eval(ctx, seeds, moves)
    first = time.perf_counter() - t0
    t0 = time.perf_counter()
    curves = batch.batch_eval(ctx, seeds, moves)
    bt = time.perf_counter() - t0
    prof = ctx.profile_read()
    # plain path: per grid, a sample of its candidates (one untimed call per grid first)
    per_seed, plain_s_per_eval, mism = [], {}, 0
    for si, sd in enumerate(seeds):
        mine = [i for i, mv in enumerate(moves) if mv[0] == si]
        sample = mine[:plain_per_grid]
        vols = [s2.voxelize(ctx, sd.moved_mesh(moves[i][1], moves[i][2]), sd.bbox, 1.0) for i in sample]
        vols[0].exact_s2(sd.r_max)
        t0 = time.perf_counter()
        got = [v.exact_s2(sd.r_max) for v in vols]
        plain_s = (time.perf_counter() - t0) / len(vols)
        plain_s_per_eval[si] = plain_s
        mism += sum(int(not np.array_equal(c, curves[i])) for c, i in zip(got, sample))
        for v in vols:
            v.free()
        # this seed's candidates alone, as one batch (second call timed)
        sub = [moves[i] for i in mine]
        batch.batch_eval(ctx, seeds, sub)
        t0 = time.perf_counter()
        batch.batch_eval(ctx, seeds, sub)
        sb = time.perf_counter() - t0
        per_seed.append({"grid": packings[si][0], "r_max": sd.r_max, "candidates": len(mine),
                         "particles": len(parts[si]), "batched_evals_per_s": len(mine) / sb,
                         "plain_evals_per_s": 1.0 / plain_s, "batched_over_plain": plain_s * len(mine) / sb})
    # the plain path's rate on the whole mix: its per-grid time weighted by the candidates
    plain_total = sum(plain_s_per_eval[mv[0]] for mv in moves)
    out = {"candidates": ncand, "batched_s": bt, "batched_evals_per_s": ncand / bt,
           "plain_evals_per_s": ncand / plain_total, "plain_sample_per_grid": plain_per_grid,
           "batched_over_plain": plain_total / bt, "bit_exact_mismatches": mism,
           "batched_device_ms": prof["device_ms"], "first_call_s": first, "grids": [96, 128, 192, 256],
           "particles": [len(p) for p in parts], "per_seed": per_seed}
    for sd in seeds:
        sd.free()
    return out


def per_radius_sums(values, rs, r_max):
    import numpy as np

    out = np.zeros(r_max + 1)
    for r in range(1, r_max + 1):
        out[r] = float(np.sum(values[rs[r]:rs[r + 1]], dtype=np.float64))
    return out


def fft_leg(ctx, vol, r_max, hits, offs, rs, curve_exact, steps):
    """FP32 FFT autocorrelation on the same resident volume: evals/s, and its
    error against the exact integer counts per radius bin (max |count error|
    over the bin's offsets, relative to the largest count of the bin) and in
    the finished curve."""
    import numpy as np

    vol.fft32_s2(r_max, 1.0)  # plan + warm-up
    ctx.profile(False)
    t0 = time.perf_counter()
    for _ in range(steps):
        curve = vol.fft32_s2(r_max, 1.0)
    el = time.perf_counter() - t0
    ctx.profile(True)
    dev, passes = [], []
    for _ in range(3):
        vol.fft32_s2(r_max, 1.0)
        p = ctx.profile_read()
        dev.append(p["device_ms"])
        passes.append(p["count_ms"])
    fc = vol.fft32_counts(r_max, 1.0).astype(np.float64)
    err = np.abs(fc - hits.astype(np.float64))
    per_bin_abs = [0.0] + [float(err[rs[r]:rs[r + 1]].max()) for r in range(1, r_max + 1)]
    per_bin_rel = [0.0] + [float(err[rs[r]:rs[r + 1]].max() / max(1.0, hits[rs[r]:rs[r + 1]].max()))
                           for r in range(1, r_max + 1)]
    d_s2 = np.abs(curve - curve_exact)
    from rustmspt_amd import s2

    f = s2.fft_padded_dims(vol.dims, r_max, 1.0)
    unpruned, comp = fft64_bytes(vol.dims, offs, r_max, 8.0)
    sec = float(np.mean(passes)) * 1e-3  # the FFT passes (z fwd .. z inv)
    return {"evals_per_s": steps / el, "device_ms_per_eval": float(np.mean(dev)),
            "roofline": {"bound": "hbm", "achieved": comp / sec / 1e9, "peak": 8000.0, "unit": "GB/s",
                         "frac": comp / sec / 8e12, "kernel": "FP32 FFT passes (z fwd, y fwd, x, y inv, z inv)",
                         "kernel_ms": sec * 1e3, "algorithmic_bytes_per_eval": comp,
                         "traffic": FFT32_PHASE_TRAFFIC if (list(vol.dims), r_max) == ([256] * 3, 64) else None,
                         "unpruned_bytes_per_eval": unpruned, "frac_unpruned": unpruned / sec / 8e12,
                         "padded": f},
            "max_abs_count_error": float(err.max()), "max_rel_count_error": float(max(per_bin_rel)),
            "rounds_to_exact": bool(np.all(np.rint(fc) == hits)),
            "max_abs_s2_error": float(d_s2.max()),
            "per_bin_max_abs_count_error": [float(f"{v:.3g}") for v in per_bin_abs],
            "per_bin_abs_s2_error": [float(f"{v:.3g}") for v in d_s2]}


def mc_leg(ctx, vol, r_max, hits, offs, rs, total_samples, seed):
    """Philox voxel Monte Carlo with total_samples pair samples spread evenly
    over the r_max radii; each radius' estimate hits/valid is compared with the
    exact estimand sum(hits)/sum(valid) of the same radius."""
    import numpy as np

    per_r = max(200, total_samples // r_max)
    vol.mc_counts(r_max, 1.0, per_r, seed)  # warm-up
    t0 = time.perf_counter()
    hv = vol.mc_counts(r_max, 1.0, per_r, seed).astype(np.float64)
    el = time.perf_counter() - t0
    dims = np.array(vol.dims, dtype=np.float64)
    valid = np.prod(dims[None, :] - np.abs(offs), axis=1)
    # the estimand samples the shell offsets uniformly then a uniform voxel, so the
    # exact value is sum over the shell of hits / sum of valid
    ex_h = per_radius_sums(hits.astype(np.float64), rs, r_max)
    ex_v = per_radius_sums(valid, rs, r_max)
    z = []
    for r in range(1, r_max + 1):
        p = ex_h[r] / ex_v[r]
        n = float(hv[r, 1])
        est = hv[r, 0] / n
        z.append((est - p) / max(np.sqrt(p * (1 - p) / n), 1e-300))
    z = np.abs(np.array(z))
    return {"samples_total": per_r * r_max, "samples_per_radius": per_r, "seconds": el,
            "samples_per_s": per_r * r_max / el, "max_abs_z": float(z.max()),
            "radii_beyond_3sigma": int((z > 3).sum()), "fraction_beyond_2sigma": float((z > 2).mean())}


def mesh_mc_leg(ctx, mesh, bbox, r_max, total_samples, seed):
    """Continuous mesh Monte Carlo on the same 500-particle mesh (no voxels):
    total_samples point pairs spread evenly over the r_max radii."""
    import numpy as np

    from rustmspt_amd import s2

    per_r = max(200, total_samples // r_max)
    from rustmspt_amd.devmesh import DeviceMesh

    s2.mesh_mc_counts(ctx, mesh, bbox, r_max, 1000, seed)  # warm-up
    t0 = time.perf_counter()
    hv = s2.mesh_mc_counts(ctx, mesh, bbox, r_max, per_r, seed).astype(np.float64)
    el = time.perf_counter() - t0
    prof = ctx.profile_read()
    vf = s2.geometric_volume_fraction(mesh, bbox)
    # resident mesh: the optimize pattern (move one particle, re-evaluate)
    dm = DeviceMesh.from_mesh(ctx, mesh)
    dm.mesh_mc_counts(bbox, r_max, 1000, seed)
    n_part = 642 if len(mesh.vertices) >= 642 else len(mesh.vertices)
    moved = mesh.vertices[:n_part] + 0.25
    t0 = time.perf_counter()
    dm.set_vertices(0, moved)
    hv2 = dm.mesh_mc_counts(bbox, r_max, per_r, seed)
    el2 = time.perf_counter() - t0
    prof2 = ctx.profile_read()
    dm.free()
    return {"samples_total": per_r * r_max, "samples_per_radius": per_r, "seconds": el,
            "device_ms": prof["device_ms"], "bin_ms": prof["count_ms"], "mc_kernel_ms": prof["fold_ms"],
            "pair_samples_per_s": per_r * r_max / el, "triangles": int(len(mesh.faces)),
            "resident_move_and_eval_s": el2, "resident_device_ms": prof2["device_ms"],
            "geometric_vf": vf, "s2_r1_2_4_8": [round(float(hv[r, 0] / hv[r, 1]), 6) for r in (1, 2, 4, 8)],
            "moved_s2_r1": round(float(hv2[1, 0]) / float(hv2[1, 1]), 6)}


def ingest_leg(ctx, n, seed, keep=False):
    """The CT-like volume (n^3 uint8 Gaussian random field, correlation length
    12, thresholded to VF 0.35) as a folder of z-major RAW slices: read through
    pinned staging, classified and bit-packed on the device.  Returns the
    figures and, with keep, the resident packed volume (else None)."""
    import shutil
    import tempfile

    import numpy as np

    from rustmspt_amd import synth
    from rustmspt_amd import volume as V

    occ = synth.grf_volume(n, 12.0, 0.35, seed)          # x-major, z fastest
    zyx = np.ascontiguousarray(np.transpose(occ, (2, 1, 0)))
    del occ
    folder = tempfile.mkdtemp(prefix="ct_raw_")
    try:
        for z in range(n):
            with open(os.path.join(folder, f"slice_{z:05d}.raw"), "wb") as fh:
                fh.write(zyx[z].tobytes())
        src = V.VolumeSource(folder, "raw", width=n, height=n, bits=8)
        phase = V.Phase.foreground(0)
        V.load(ctx, src, phase).volume.free()                # warm-up (page cache, allocations)
        t0 = time.perf_counter()
        lv = V.load(ctx, src, phase)
        load_s = time.perf_counter() - t0
        pack_ms = ctx.profile_read()["count_ms"]
        nwords = lv.volume.nwords
        occupied = lv.occupied
        kept = lv.volume if keep else None
        if not keep:
            lv.volume.free()
        t0 = time.perf_counter()
        lv2 = V.from_samples(ctx, zyx, phase)
        mem_s = time.perf_counter() - t0
        lv2.volume.free()
    finally:
        shutil.rmtree(folder, ignore_errors=True)
    sample_bytes = n ** 3
    alg = sample_bytes + nwords * 8
    return kept, {"grid": n, "sample_bytes": sample_bytes, "packed_bytes": nwords * 8, "vf": occupied / n ** 3,
            "raw_folder_load_s": load_s, "raw_folder_GB_per_s": sample_bytes / load_s / 1e9,
            "from_memory_s": mem_s, "pack_kernel_ms": pack_ms,
            "pack_roofline": {"bound": "hbm", "achieved": alg / (pack_ms * 1e-3) / 1e9, "peak": 8000.0,
                              "unit": "GB/s", "frac": alg / (pack_ms * 1e-3) / 1e9 / 8000.0,
                              "algorithmic_bytes": alg}}


def ct_leg(ctx, vol, r_max, steps):
    """Exact S2 of the ingested CT-like volume (the RAW folder, packed on the
    device) at r_max: evals/s, the count kernel against the VALU roofline, and
    the FP32 FFT path's per-bin error against the exact counts."""
    import numpy as np

    from rustmspt_amd import s2

    dims = list(vol.dims)
    offs, rs = s2.shell_order_offsets(dims, r_max, 1.0)
    valid = np.prod(np.array(dims, dtype=np.float64)[None, :] - np.abs(offs), axis=1)
    half_work = float(valid.sum()) / 2.0
    t0 = time.perf_counter()
    vol.exact_s2(r_max, 1.0)  # plan (offsets, columns, launch shape) + first run
    first_s = time.perf_counter() - t0
    auto = exact_timed(ctx, vol, r_max, "auto", max(steps, 3), 1)
    kernel = ctx.exact_tune_info(dims, r_max, 1.0)["kernel"]
    direct = exact_timed(ctx, vol, r_max, "direct", steps, 1)
    achieved = half_work / 32.0 * 2.0 / (direct["count_ms"] * 1e-3)
    peak = ctx.valu_probe()
    out = {"grid": dims, "r_max": r_max, "offsets": int(len(offs)), "pair_work": 2.0 * half_work,
           "exact_evals_per_s": auto["evals_per_s"], "exact_s_per_eval": auto["ms_per_step"] * 1e-3,
           "exact_kernel": kernel, "first_call_s": first_s, "count_phase_ms": auto["count_ms"],
           "device_ms_per_eval": auto["device_ms"],
           "curves_identical": auto["curve"].tobytes() == direct["curve"].tobytes(),
           "direct": {"evals_per_s": direct["evals_per_s"], "count_kernel_ms": direct["count_ms"],
                      "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": peak / 1e12,
                                   "unit": "TOP/s (int32 lane-ops)", "frac": achieved / peak,
                                   "peak_4cycle": VALU_4CYCLE_OPS / 1e12,
                                   "frac_of_4cycle": achieved / VALU_4CYCLE_OPS},
                      "launch_shape": ctx.exact_tune_info(dims, r_max, 1.0)},
           "s2_head": [round(float(v), 6) for v in auto["curve"][:5]]}
    if kernel == "fft64":
        unpruned, comp = fft64_bytes(dims, offs, r_max)
        sec = auto["count_ms"] * 1e-3
        out["roofline"] = {"bound": "hbm", "achieved": comp / sec / 1e9, "peak": 8000.0, "unit": "GB/s",
                           "frac": comp / sec / 8e12, "algorithmic_bytes": comp, "unpruned_bytes": unpruned,
                           "frac_unpruned": unpruned / sec / 8e12}
    hits = vol.exact_counts(r_max, 1.0)
    out["fft32"] = fft_leg(ctx, vol, r_max, hits, offs, rs, auto["curve"], steps=2)
    return out


def stl_ingest_leg(ctx, mesh, bbox, reps=3):
    """The single-measure packing as a binary STL (32 MB, 640,000 faces) to a
    256^3 occupancy grid: host weld + granules + host-prepared voxelization vs
    weld, granules and voxelization on the device (same bits)."""
    import shutil
    import tempfile

    import numpy as np

    from rustmspt_amd import mesh as M
    from rustmspt_amd import s2
    from rustmspt_amd.devmesh import DeviceMesh

    d = tempfile.mkdtemp(prefix="stl_")
    try:
        path = os.path.join(d, "packing.stl")
        M.save_stl(path, mesh)
        host_s, dev_s, dev_load_s = [], [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            hm = M.load_stl(path)
            hv = s2.voxelize(ctx, hm, bbox, 1.0)   # granules + host preparation inside
            hw = hv.read_packed() if not host_s else None
            hv.free()
            host_s.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            dm = DeviceMesh.load_stl(ctx, path)
            t1 = time.perf_counter()
            dv = dm.voxelize(bbox, 1.0)          # ends with a device popcount: synchronous
            dev_s.append(time.perf_counter() - t0)
            dev_load_s.append(t1 - t0)
            if hw is not None:
                same = bool(np.array_equal(dv.read_packed(), hw))
            dv.free()
            dm.free()
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return {"faces": int(len(mesh.faces)), "file_bytes": 84 + 50 * int(len(mesh.faces)),
            "host_stl_to_grid_s": min(host_s), "device_stl_to_grid_s": min(dev_s),
            "device_load_weld_granules_s": min(dev_load_s), "speedup": min(host_s) / min(dev_s),
            "grids_identical": same}


def served_leg(rate: float, duration: float, pool_size: int, slots: int, budget_gb: float, seed: int,
               compare_slots: int = 4, compare_s: float = 20.0, sustained_s: float = 30.0, burst: int = 600,
               plain_sample: int = 300):
    """Poisson stream of mixed requests (grid 64^3-384^3 log-uniform, VF
    U[0.1, 0.5], r_max U{16..96}, 70 % exact / 30 % FP32 FFT) through the serving
    loop; p50/p99 latency from submission to completion, evals/s, admission.
    Runs: the offered rate for `duration` s on `slots` slots (the stated
    workload) and for `compare_s` s on `compare_slots` slots; capacity: a burst
    of `burst` requests submitted at once, on both slot counts; a stream at
    70 % of the larger capacity for `sustained_s` s.  Beside them, the plain
    per-call path (one context, one request at a time) on the first
    `plain_sample` requests of the burst."""
    from rustmspt_amd import s2, serve

    pool = serve.volume_pool(pool_size, seed)
    out = {}
    bursts = serve.poisson_schedule(1e6, burst / 1e6, pool, seed + 7)

    def run(label, nslots, sched, lam, secs):
        with serve.Server(slots=nslots, budget_bytes=int(budget_gb * (1 << 30))) as sv:
            for m in (serve.EXACT, serve.FFT32):  # warm-up: one small request per method
                t = sv.submit(pool[0].words, pool[0].dims, 16, m)
                sv.poll(t, wait=True)
            rep = serve.run_stream(sv, pool, sched, timeout=900)
        rep.update({"rate": lam, "duration_s": secs, "slots": nslots})
        out[label] = rep
        progress(f"served {label}: {rep['requests']} requests, {rep['throughput_evals_per_s']:.1f} evals/s, "
                 f"p50 {rep['latency_p50_ms']:.1f} ms, p99 {rep['latency_p99_ms']:.1f} ms")

    run("offered", slots, serve.poisson_schedule(rate, duration, pool, seed), rate, duration)
    if compare_s > 0:
        run(f"offered_{compare_slots}_slots", compare_slots, serve.poisson_schedule(rate, compare_s, pool, seed + 1),
            rate, compare_s)
    for n in sorted({slots, compare_slots}):
        run(f"capacity_{n}_slots", n, bursts, None, 0.0)
    cap = max(out[f"capacity_{n}_slots"]["throughput_evals_per_s"] for n in {slots, compare_slots})
    best = max((n for n in {slots, compare_slots}), key=lambda n: out[f"capacity_{n}_slots"]["throughput_evals_per_s"])
    if sustained_s > 0:
        lam = round(0.7 * cap, 1)
        run("sustained", best, serve.poisson_schedule(lam, sustained_s, pool, seed + 2), lam, sustained_s)
    # the plain path on the burst's requests, one at a time
    ctx = s2.Context(0)
    ctx.set_autotune(False)
    vols, n = {}, min(plain_sample, len(bursts))
    try:
        for a in bursts[:n]:  # volumes resident before the timing, as in the server
            if a.volume not in vols:
                vols[a.volume] = s2.Volume.from_packed(ctx, pool[a.volume].words, pool[a.volume].dims)
        t0 = time.perf_counter()
        for a in bursts[:n]:
            vol = vols[a.volume]
            if a.method == serve.EXACT:
                vol.exact_s2(a.r_max, 1.0)
            else:
                vol.fft32_s2(a.r_max, 1.0)
        plain_s = time.perf_counter() - t0
    finally:
        for vol in vols.values():
            vol.free()
        ctx.close()
    out["plain"] = {"requests": n, "evals_per_s": n / plain_s,
                    "note": "the burst's first requests in order through the plain per-call path on one context "
                            "(volumes resident, plans built on first use as in the server)"}
    out["capacity_evals_per_s"] = cap
    out["capacity_slots"] = best
    out["served_over_plain"] = cap / out["plain"]["evals_per_s"]
    out.update({"pool_volumes": pool_size, "budget_gb": budget_gb, "grids": sorted({v.dims[0] for v in pool})})
    return out


def progress(msg: str) -> None:
    """A progress line on stderr (the result line is the only stdout output)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--r-max", type=int, default=64)
    ap.add_argument("--particles", type=int, default=500)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cpu-legs", action="store_true",
                    help="skip the forced-FFT, planner-pick and CT-volume CPU timings")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-candidates", type=int, default=256)
    ap.add_argument("--no-fft", action="store_true")
    ap.add_argument("--no-mc", action="store_true")
    ap.add_argument("--mc-samples", type=int, default=10_000_000, help="total pair samples over all radii")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-ingest", action="store_true")
    ap.add_argument("--ingest-grid", type=int, default=768, help="CT-like RAW volume edge for the ingest leg")
    ap.add_argument("--ct-r-max", type=int, default=128,
                    help="r_max of the exact and FFT32 S2 of the ingested CT-like volume (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--tune", type=str, default="",
                    help="band,waves: force the exact launch shape instead of the autotuner's pick")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="first run a plain device copy of this many MiB (counter calibration for profiling)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU sharding is a design note only "
                                   "(offset ranges per GPU, one gather of per-offset counts)", "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np

    from rustmspt_amd import s2, synth

    if args.copy_probe_mb:
        import torch
        n = args.copy_probe_mb * 1024 * 1024 // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    grid, r_max = args.grid, args.r_max
    mesh, bbox, parts = synth.config_large(seed=2, n=args.particles, grid=grid)
    ctx = s2.Context(0)
    t0 = time.perf_counter()
    vol = s2.voxelize(ctx, mesh, bbox, 1.0)
    vox_s = time.perf_counter() - t0
    occupied = vol.occupied
    ctx.profile(True)
    offs, rs = s2.shell_order_offsets([grid] * 3, r_max, 1.0)
    valid = np.prod(np.array([grid] * 3, dtype=np.float64)[None, :] - np.abs(offs), axis=1)
    pair_work = float(valid.sum())          # W: ordered pair-tests of all shell offsets
    half_work = pair_work / 2.0             # C(d) = C(-d): unordered pair-tests

    if args.tune:
        band, waves = (int(v) for v in args.tune.split(","))
        ctx.exact_tune_set([grid] * 3, r_max, 1.0, band, waves)
    # the exact path as callers get it: the kernel the model picks (direct
    # counting or the FP64 FFT rounded to the same integer counts) ...
    auto = exact_timed(ctx, vol, r_max, "auto", args.steps, args.warmup)
    kernel = ctx.exact_tune_info([grid] * 3, r_max, 1.0)["kernel"]
    progress(f"exact ({kernel}): {auto['evals_per_s']:.1f} evals/s")
    curve = auto["curve"]
    # ... and the direct counting kernel forced, beside it
    direct = exact_timed(ctx, vol, r_max, "direct", args.steps, max(args.warmup, 1))
    count_avg = direct["count_ms"]
    # algorithmic VALU work: one v_and_b32 + one v_bcnt_u32_b32 per 32 unordered pair-tests
    alg_ops = half_work / 32.0 * 2.0
    achieved = alg_ops / (count_avg * 1e-3)
    valu_measured = ctx.valu_probe()  # issue ceiling of the counting loop's and/popcount pair
    shape = ctx.exact_tune_info([grid] * 3, r_max, 1.0)
    profiled = (grid, r_max, args.particles, shape["band"], shape["waves"]) == (256, 64, 500, 3, 65536)
    direct_roofline = {"bound": "valu", "achieved": achieved / 1e12, "peak": valu_measured / 1e12,
                       "unit": "TOP/s (int32 lane-ops)", "frac": achieved / valu_measured,
                       "traffic": COUNT_PHASE_TRAFFIC if profiled else None,
                       "peak_source": "on-device probe of the counting loop's v_and_b32 (SGPR operand) + "
                                      "v_bcnt_u32_b32 pair, every SIMD occupied (no published int32 VALU figure)",
                       "peak_4cycle": VALU_4CYCLE_OPS / 1e12, "frac_of_4cycle": achieved / VALU_4CYCLE_OPS,
                       "kernel": "exact_count_kernel", "kernel_ms": count_avg,
                       "launches_per_eval": direct["launches"], "device_ms_per_eval": direct["device_ms"],
                       "algorithmic_ops_per_eval": alg_ops, "launch_shape": shape}
    if kernel == "fft64":
        unpruned_b, comp_b = fft64_bytes([grid] * 3, offs, r_max)
        sec = auto["count_ms"] * 1e-3
        roofline = {"bound": "hbm", "achieved": comp_b / sec / 1e9, "peak": 8000.0, "unit": "GB/s",
                    "frac": comp_b / sec / 8e12,
                    "traffic": FFT64_PHASE_TRAFFIC if (grid, r_max, args.particles) == (256, 64, 500) else None,
                    "kernel": "FP64 FFT passes (z fwd, y fwd, x fwd+|F|^2+inv, y inv, z inv, gather)",
                    "kernel_ms": auto["count_ms"], "launches_per_eval": auto["launches"],
                    "device_ms_per_eval": auto["device_ms"], "algorithmic_bytes_per_eval": comp_b,
                    "unpruned_bytes_per_eval": unpruned_b, "frac_unpruned": unpruned_b / sec / 8e12,
                    "note": "algorithmic bytes: what the pruned passes must move (each reads its input and "
                            "writes its output once); unpruned: 12 passes over the complex128 half spectrum "
                            "of a transform keeping every output, which the pruning beats"}
    else:
        roofline = direct_roofline

    out = {
        "metric": "exact-S2 evals/s at 256^3 r_max=64, batched vs plain; p50/p99 served latency",
        "value": auto["evals_per_s"],
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": auto["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 integer counts (direct: u32 popcounts; fft64: f64 spectra rounded), f64 curve",
        "data": "synthetic: seeded packing, voxelized on device (bit-exact f64 predicate)",
        "config": {"workload": f"single large measure: {args.particles} particles (icospheres + ellipsoids), "
                               f"{grid}^3 grid, r_max={r_max}, exact S2, plain per-call path",
                   "exact_kernel": kernel,
                   "grid": grid, "r_max": r_max, "offsets": int(len(offs)), "pair_work": pair_work,
                   "volume_fraction": occupied / grid ** 3, "voxelize_s": vox_s},
        "roofline": roofline,
        "direct": {"evals_per_s": direct["evals_per_s"], "ms_per_step": direct["ms_per_step"],
                   "curves_identical": direct["curve"].tobytes() == curve.tobytes(), "roofline": direct_roofline},
        "s2_head": [round(float(v), 6) for v in curve[:5]],
    }
    if not (args.no_fft and args.no_mc):
        hits = vol.exact_counts(r_max, 1.0)
        if not args.no_fft:
            out["fft32"] = fft_leg(ctx, vol, r_max, hits, offs, rs, curve, steps=5)
            progress("fft32 done")
        if not args.no_mc:
            out["monte_carlo"] = mc_leg(ctx, vol, r_max, hits, offs, rs, args.mc_samples, seed=7)
            progress("monte carlo done")
    if not args.no_mc:
        out["mesh_monte_carlo"] = mesh_mc_leg(ctx, mesh, bbox, r_max, args.mc_samples, seed=11)
        progress("mesh monte carlo done")
    ct_cpu = None
    if not args.no_ingest:
        out["stl_ingest"] = stl_ingest_leg(ctx, mesh, bbox)
        ct_vol, out["ingest"] = ingest_leg(ctx, args.ingest_grid, seed=5, keep=args.ct_r_max > 0)
        if ct_vol is not None:
            out["ct_volume"] = ct_leg(ctx, ct_vol, args.ct_r_max, steps=2)
            if not args.no_cpu:
                ct_hits = ct_vol.exact_counts(args.ct_r_max, 1.0)
                ct_cpu = (ct_vol.read_packed(), list(ct_vol.dims), args.ct_r_max, lambda idx: ct_hits[idx])
            ct_vol.free()
        progress("ingest and CT volume done")
    if not args.no_batch:
        out["batched"] = batched_leg(ctx, args.batch_candidates, seed=3)
        progress(f"batched: {out['batched'].get('batched_over_plain', 0):.2f}x plain")
    if not args.no_serve:
        out["served"] = served_leg(args.serve_rate, args.serve_seconds, 24, args.serve_slots, 64.0, seed=4)
    if not args.no_cpu:
        words = vol.read_packed()
        out["cpu_baseline"] = cpu_baseline(words, [grid] * 3, r_max, offs, pair_work, args.cpu_seconds)
        progress("cpu: direct sample done")
        if not args.no_cpu_legs:
            out["cpu_baseline"]["legs"] = cpu_legs(ctx, words, [grid] * 3, r_max, vol.exact_counts(r_max, 1.0),
                                                   ct_cpu, args.cpu_seconds)
        out["cpu_baseline"]["gpu_speedup"] = out["value"] / out["cpu_baseline"]["value"]
        out["cpu_baseline"]["gpu_direct_speedup"] = direct["evals_per_s"] / out["cpu_baseline"]["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
