"""`measure` pipeline on the HIP engine (src/pipeline/measure.rs:49-291).

Same inputs (YAML), bbox precedence, grid normalisation, working-set gate,
per-method backend resolution with the reference's fallback policy, report file
format and [Timing] lines.  The exact / Monte Carlo curves come from
libs2hip.so.  A method that resolves to the CPU backend belongs to the host
crate's own CPU implementation; this engine does not contain one, so such a
resolution raises CpuPathUnavailable instead of computing anything.
"""
from __future__ import annotations

import math
import os
import sys
import time
from pathlib import Path

import numpy as np

from . import config as cfg
from . import report
from . import s2
from ._lib import S2HipError
from .devmesh import DeviceMesh
from .mesh import BoundingBox, granules, load_stl_or_merge_folder, mesh_bbox

DEFAULT_CPU_EXACT_BUDGET_BYTES = 768 * 1024 * 1024


class CpuPathUnavailable(RuntimeError):
    """The policy chose the CPU backend, which lives in the host crate."""


StageTimer = report.StageTimer
l2_error = report.l2_error


class MeasurePipeline:
    def __init__(self, params: cfg.MeasurementParams, env=None, out=sys.stdout):
        self.p = params
        self.env = os.environ if env is None else env
        self.out = out

    def log(self, msg):
        print(msg, file=self.out)

    def run(self):
        self._res = {"grid": None, "dmesh": None, "ctx": None}
        try:
            return self._run()
        finally:
            r = self._res
            if r["grid"] is not None:
                r["grid"].volume.free()
            if r["dmesh"] is not None:
                r["dmesh"].free()
            if r["ctx"] is not None:
                r["ctx"].close()

    def _run(self):
        p = self.p
        res = self._res
        if not math.isfinite(p.voxel_pitch):
            raise cfg.ConfigError("measurement requires finite pitch and representable radius count")
        requested = cfg.configured_mode(p.acceleration, self.env)
        timer = StageTimer("measure", self.out)
        ctx, dmesh = None, None
        # With a HIP-capable mode the STL is welded and split into granules on the
        # device (the resident mesh is voxelized without host preparation); the
        # host copy read back serves the volume fraction and the mesh MC.
        mesh = particle_count = None
        if requested != cfg.AccelerationMode.CPU and self.env.get("S2HIP_HOST_INGEST", "0") != "1":
            try:
                budget = (p.acceleration.gpu_memory_limit_mb or 0) * 1024 * 1024
                ctx = res["ctx"] = s2.Context(0, budget)
                dmesh = res["dmesh"] = DeviceMesh.load_stl(ctx, p.stl_path)
                mesh, _ = dmesh.read()
                particle_count = dmesh.ngranules
            except S2HipError as err:
                # the host loader reads the same file into the same mesh
                print(f"[Warning] device STL ingest: {err}; loading on the host", file=sys.stderr)
                if res["dmesh"] is not None:
                    res["dmesh"].free()
                dmesh = res["dmesh"] = None
                mesh = particle_count = None
        if mesh is None:
            mesh = load_stl_or_merge_folder(p.stl_path)
        timer.stage("load")
        bbox = None
        for box in (p.bounding_box, p.stl_bounding_box):
            if box:
                bbox = cfg.parse_box_dimensions(box)
                break
        if bbox is None:
            bbox = (dmesh.bbox if dmesh is not None else mesh_bbox(mesh)) or BoundingBox.from_size((1.0, 1.0, 1.0))
        self.log(f"[Info] STL file(s) loaded from: {p.stl_path}")
        self.log("[Info] Measurement bbox: min=({:.4f},{:.4f},{:.4f}), max=({:.4f},{:.4f},{:.4f})".format(
            *bbox.min, *bbox.max))
        # particle count and the geometric volume fraction in the box (granules
        # crossing it clipped), before any S2 work, as the reference's split_vf stage
        if particle_count is None:
            _, particle_count = granules(mesh)
        vf = s2.geometric_volume_fraction(mesh, bbox)
        timer.stage("split_vf")
        raw = p.mc_method.strip().lower()
        method = "exact" if raw == "exact" else ("both" if raw == "both" else ("fft32" if raw == "fft32" else
                                                                              "monte_carlo"))
        samples = p.mc_samples if p.mc_samples is not None else 10_000
        self.log(f"[Info] S2 config: method={method}, r_max={p.r_max}, mc_samples={samples}, "
                 f"voxel_pitch={p.voxel_pitch:.6f}")
        pitch = 1.0 if p.voxel_pitch <= 0.0 else p.voxel_pitch
        dims, voxel_count = [], 1
        for a in range(3):
            extent = bbox.max[a] - bbox.min[a]
            n = max(math.ceil(extent / pitch), 1.0) if math.isfinite(extent / pitch) else float("inf")
            if not math.isfinite(extent) or extent <= 0.0 or not math.isfinite(n):
                raise cfg.ConfigError("measurement grid dimensions are unsupported")
            dims.append(int(n))
            voxel_count *= int(n)
        methods = ["exact", "monte_carlo"] if method == "both" else [method]
        curves, backends, grid = [], [], None
        timer.restart()
        for m in methods:
            exactish = m in ("exact", "fft32")
            # monte_carlo with voxel_pitch <= 0 is the continuous mesh MC, as calculate_s2
            # routes it (src/geometry/s2.rs:1351-1356; measure, src/pipeline/measure.rs:182-215)
            mesh_mc = m == "monte_carlo" and p.voxel_pitch <= 0.0
            est = cfg.hip_exact_bytes(dims, p.r_max, pitch, len(mesh.faces), fft=(m == "fft32")) if exactish \
                else (len(mesh.faces) * 144 + len(mesh.vertices) * 24 + 16 * (p.r_max + 1) if mesh_mc else
                      dims[0] * dims[1] * s2.row_words(dims[2]) * 8 + len(mesh.faces) * 136 + 64 * (p.r_max + 2))
            sel = cfg.resolve_execution(p.acceleration, requested, voxel_count, p.acceleration.gpu_min_voxels,
                                        True, est)
            if sel.fallback_reason:
                print(f"[Info] {m}: {sel.fallback_reason}", file=sys.stderr)
            if sel.backend != "hip":
                raise CpuPathUnavailable(
                    f"measure {m}: the policy selected the CPU backend, which is the host crate's own "
                    f"implementation and not part of this engine")
            seed = int(self.env.get("RUSTMSPT_S2_SEED", "0"))
            try:
                if ctx is None:
                    budget = (p.acceleration.gpu_memory_limit_mb or 0) * 1024 * 1024
                    ctx = res["ctx"] = s2.Context(0, budget)
                if mesh_mc:
                    hv = s2.mesh_mc_counts(ctx, mesh, bbox, p.r_max, samples, seed).astype(np.float64)
                    curve = np.array([vf] + [hv[r, 0] / hv[r, 1] if hv[r, 1] else 0.0
                                             for r in range(1, p.r_max + 1)])
                else:
                    if grid is None:
                        self.log("[Info] Preparing shared HIP S2 occupancy grid")
                        if dmesh is not None:
                            grid = s2.VoxelS2Hip.from_volume(dmesh.voxelize(bbox, max(pitch, 1e-9)), pitch)
                        else:
                            grid = s2.VoxelS2Hip(ctx, mesh, bbox, pitch)
                        res["grid"] = grid
                    curve = grid.calculate_seeded(p.r_max, m, samples, seed)
            except S2HipError as err:
                # a failing device run: an error without cpu_fallback, else the
                # reference's warning and its CPU path (src/pipeline/measure.rs:225-246)
                if not p.acceleration.cpu_fallback:
                    raise cfg.GpuError(f"measure {m}: {err}") from err
                print(f"[Warning] measure {m}: {err}; falling back to CPU {m}", file=sys.stderr)
                raise CpuPathUnavailable(
                    f"measure {m}: the HIP run failed and the CPU fallback is the host crate's own "
                    f"implementation, not part of this engine") from err
            backends.append(f"{m}=hip/{ctx.device_name}")
            self.log(f"[Info] Measure execution: method={m}, backend=hip, device={ctx.device_name}")
            curves.append(curve)
            timer.stage(f"s2_{m}")
        out = Path(p.output_path)
        out.parent.mkdir(parents=True, exist_ok=True)
        out.write_text(report.measure_report(vf, backends, method, methods, curves))
        timer.stage("write")
        self.log(f"[Info] Measurement completed: {particle_count} particles, VF {vf:.6f}; output {p.output_path}")
        timer.total("total_in_pool")
        timer.report_resources()
        return curves


def main(argv=None):
    """`python -m rustmspt_amd.pipeline measure --config cfg.yaml [--input X] [--output Y]`
    (src/main.rs:224-253)."""
    import argparse
    ap = argparse.ArgumentParser(prog="rustmspt-hip")
    sub = ap.add_subparsers(dest="cmd", required=True)
    m = sub.add_parser("measure")
    m.add_argument("--config", required=True)
    m.add_argument("--input")
    m.add_argument("--output")
    args = ap.parse_args(argv)
    params = cfg.load_measurement_config(args.config)
    if args.input:
        params.stl_path = args.input
    if args.output:
        params.output_path = args.output
    MeasurePipeline(params).run()


if __name__ == "__main__":
    main()
