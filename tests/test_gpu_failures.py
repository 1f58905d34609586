"""Error paths under injected device failures (S2HIP_FAIL_AT=alloc|launch|readback):
the library returns S2H_ERR_HIP from its host wrapper before the failing step's
device work, so no kernel is broken.  Checked: the measure pipeline's policy
(tests/measure_execution_tests.rs:7-72 and src/pipeline/measure.rs:225-246:
without cpu_fallback an error naming the method, with it the warning and the
host's CPU path), every entry point's status, and that the context and the
server keep working afterwards.  Requires the GPU (marker gpu)."""
import json

import numpy as np
import pytest

from rustmspt_amd import _lib as L
from rustmspt_amd import batch, config as cfg
from rustmspt_amd import mesh as M
from rustmspt_amd import pipeline, s2, serve, synth

pytestmark = pytest.mark.gpu

SITES = ["alloc", "launch", "readback"]


def measure(tmp_path, method, fallback, env=None, host_ingest=True):
    m = M.merge_meshes([M.icosphere_mesh(np.array([6.0, 6.5, 5.5]), 3.0, 2),
                        M.icosphere_mesh(np.array([13.0, 12.0, 14.0]), 4.0, 2)])
    M.save_stl(tmp_path / "in.stl", m)
    doc = {"measurement": {"stl_path": str(tmp_path / "in.stl"), "output_path": str(tmp_path / "out.txt"),
                           "r_max": 6, "voxel_pitch": 1.0, "mc_method": method, "mc_samples": 5000,
                           "bounding_box": [20, 20, 20],
                           "acceleration": {"mode": "hip", "cpu_fallback": fallback}}}
    (tmp_path / "c.json").write_text(json.dumps(doc))
    env = dict(env or {}, S2HIP_HOST_INGEST="1" if host_ingest else "0")
    pipeline.MeasurePipeline(cfg.load_measurement_config(tmp_path / "c.json"), env=env).run()
    return (tmp_path / "out.txt").read_text()


@pytest.mark.parametrize("site", SITES)
@pytest.mark.parametrize("method", ["exact", "monte_carlo"])
def test_measure_policy_on_device_failure(tmp_path, monkeypatch, capfd, site, method):
    want = measure(tmp_path, method, False, {"RUSTMSPT_S2_SEED": "3"})
    monkeypatch.setenv("S2HIP_FAIL_AT", site)
    with pytest.raises(cfg.GpuError, match=f"measure {method}: .*injected"):
        measure(tmp_path, method, False, {"RUSTMSPT_S2_SEED": "3"})
    monkeypatch.setenv("S2HIP_FAIL_AT", site + ":1 ")  # re-armed: the variable changed
    with pytest.raises(pipeline.CpuPathUnavailable):
        measure(tmp_path, method, True, {"RUSTMSPT_S2_SEED": "3"})
    assert f"[Warning] measure {method}: " in capfd.readouterr().err
    # the next run, with the injection spent, is the same as before
    assert measure(tmp_path, method, False, {"RUSTMSPT_S2_SEED": "3"}) == want


def test_device_ingest_failure_loads_on_the_host(tmp_path, monkeypatch, capfd):
    want = measure(tmp_path, "exact", False, host_ingest=False)
    monkeypatch.setenv("S2HIP_FAIL_AT", "alloc:2")  # the first mesh buffer after the context's own
    assert measure(tmp_path, "exact", False, host_ingest=False) == want
    assert "[Warning] device STL ingest: " in capfd.readouterr().err


@pytest.mark.parametrize("site", SITES)
def test_entry_points_report_and_recover(ctx, monkeypatch, site):
    rng = np.random.default_rng(4)
    occ = (rng.random((24, 26, 22)) < 0.3).astype(np.uint8)
    vol = s2.Volume.from_bytes(ctx, occ)
    calls = [lambda: vol.exact_s2(7), lambda: vol.exact_counts(7), lambda: vol.fft32_s2(7),
             lambda: vol.mc_s2(7, 1.0, 2000, 5)]
    want = [c() for c in calls]
    for k, call in enumerate(calls):
        monkeypatch.setenv("S2HIP_FAIL_AT", site)
        if site == "alloc":
            # every allocation passes the site, cached block or not
            with pytest.raises(L.S2HipError, match="injected allocation failure"):
                s2.Volume.from_bytes(ctx, occ[:, :, :20 - k].copy())
        else:
            name = "launch" if site == "launch" else "read-back"
            with pytest.raises(L.S2HipError, match=f"injected {name} failure"):
                call()
        monkeypatch.delenv("S2HIP_FAIL_AT")
        for c, w in zip(calls, want):
            assert np.array_equal(c(), w)
    vol.free()


def test_batch_and_server_recover(ctx, monkeypatch):
    rng = synth.rng_for(8, 2)
    parts = synth.place_particles(8, 32, rng, synth.ellipsoid_maker(2.5, 4.0), overlap=True)
    sd = batch.Seed(ctx, [p.mesh() for p in parts], s2.BoundingBox.from_size((32, 32, 32)), 1.0, 8)
    moves = [(0, i, parts[i].moved(rng.uniform(-2, 2, 3), synth.random_rotation(rng)).mesh().vertices)
             for i in range(5)]
    want = batch.batch_eval(ctx, [sd], moves)
    for site in ("launch", "readback"):
        monkeypatch.setenv("S2HIP_FAIL_AT", site)
        with pytest.raises(L.S2HipError, match="injected"):
            batch.batch_eval(ctx, [sd], moves)
        monkeypatch.delenv("S2HIP_FAIL_AT")
        for a, b in zip(batch.batch_eval(ctx, [sd], moves), want):
            assert np.array_equal(a, b)
    sd.free()
    occ = serve.sphere_packing(40, 0.3, synth.rng_for(3, 1))
    words = serve.pack(occ)
    with serve.Server(slots=2) as sv:
        good = sv.poll(sv.submit(words, occ.shape, 10, serve.EXACT), wait=True)[1]
        monkeypatch.setenv("S2HIP_FAIL_AT", "launch")
        t = sv.submit(words, occ.shape, 10, serve.EXACT)
        with pytest.raises(L.S2HipError, match="request failed: injected"):
            sv.poll(t, wait=True)
        monkeypatch.delenv("S2HIP_FAIL_AT")
        again = sv.poll(sv.submit(words, occ.shape, 10, serve.EXACT), wait=True)[1]
        st = sv.stats()
    assert np.array_equal(good, again) and st["failed"] == 1 and st["completed"] == 2
