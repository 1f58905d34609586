"""Report, history and timing output in the reference's formats, so tools that
parse RustMSPT's output (scripts/perf_matrix.py, the measure report readers)
read this engine's output unchanged.

- measure report: MeasurePipeline::run's output file (src/pipeline/measure.rs:253-282);
- S2 history lines: format_s2_series / push_history_s2 and the per-iteration
  best line of run_sa_island (src/pipeline/optimize.rs:139-150, 761-764);
- [Timing] lines: StageTimer, format_stage_line, report_workers,
  report_peak_rss, parse_vm_hwm (src/pipeline/timing.rs:3-81).
"""
from __future__ import annotations

import math
import sys
import time
from typing import List, Optional, Sequence


# ------------------------------------------------------------------ measure report

def l2_error(a: Sequence[float], b: Sequence[float]) -> float:
    """MeasurePipeline::l2_error (src/pipeline/measure.rs:27-39): over the common
    prefix; 0 when either curve is empty."""
    n = min(len(a), len(b))
    if n == 0:
        return 0.0
    return math.sqrt(sum((float(a[i]) - float(b[i])) ** 2 for i in range(n)))


def measure_report(vf: float, backends: Sequence[str], requested_method: str, methods: Sequence[str],
                   curves: Sequence[Sequence[float]]) -> str:
    """The measure output file, byte for byte (src/pipeline/measure.rs:253-282)."""
    out = [f"Volume Fraction: {vf:.6f}", f"Compute Backend: {', '.join(backends)}", f"Method: {requested_method}"]
    for m, curve in zip(methods, curves):
        out.append(f"S2(0)-VF diff [{m}]: {abs(float(curve[0]) - vf):.6f}")
        if len(methods) == 1:
            out.append("S2 Values:")
        elif m == "exact":
            out.append("S2 Values [exact]:")
        else:
            out.append("S2 Values [monte_carlo]:")
        out += [f"{r}: {float(v):.6f}" for r, v in enumerate(curve)]
    if len(curves) == 2:
        out.append(f"L2 error [exact vs monte_carlo]: {l2_error(curves[0], curves[1]):.6f}")
    return "\n".join(out) + "\n"


# ------------------------------------------------------------------ S2 history

def format_s2_series(values: Sequence[float]) -> str:
    """format_s2_series (src/pipeline/optimize.rs:139-145): space-separated, six decimals."""
    return " ".join(f"{float(v):.6f}" for v in values)


def push_history_s2(history: List[str], label: str, values: Sequence[float]) -> None:
    """push_history_s2 (src/pipeline/optimize.rs:147-150)."""
    history.append(f"{label}: {format_s2_series(values)}")


def push_history_best(history: List[str], iteration: int, loss: float, s2: Sequence[float]) -> None:
    """A new island best in run_sa_island (src/pipeline/optimize.rs:761-764)."""
    history.append(f"Iter {iteration}: Loss {loss:.6f} | S2 {format_s2_series(s2)}")


def write_history(path, history: Sequence[str]) -> None:
    """s2_history.txt: the lines joined by newlines, no trailing newline
    (src/pipeline/optimize.rs:1052)."""
    with open(path, "w") as fh:
        fh.write("\n".join(history))


# ------------------------------------------------------------------ [Timing]

def format_stage_line(pipeline: str, stage: str, seconds: float) -> str:
    """format_stage_line (src/pipeline/timing.rs:49-51): nine decimals."""
    return f"[Timing] {pipeline} stage={stage} seconds={seconds:.9f}"


def parse_vm_hwm(status: str) -> Optional[int]:
    """parse_vm_hwm (src/pipeline/timing.rs:73-81): the VmHWM line in kB as bytes;
    None when absent or malformed."""
    for line in status.splitlines():
        if line.startswith("VmHWM:"):
            fields = line[len("VmHWM:"):].split()
            if len(fields) < 2 or fields[1] != "kB":
                return None
            try:
                value = int(fields[0])
            except ValueError:
                return None
            return value * 1024 if value >= 0 else None
    return None


def peak_rss_bytes() -> Optional[int]:
    """peak_rss_bytes (src/pipeline/timing.rs:67-71): VmHWM of /proc/self/status."""
    try:
        with open("/proc/self/status") as fh:
            return parse_vm_hwm(fh.read())
    except OSError:
        return None


class StageTimer:
    """StageTimer (src/pipeline/timing.rs:3-47): `[Timing] <pipeline> stage=<name>
    seconds=<f>` per stage, the total, and the workers / peak RSS lines.  The host
    side of this engine runs on one thread, so workers is 1."""

    def __init__(self, pipeline: str, out=sys.stdout, workers: int = 1):
        self.pipeline = pipeline
        self.out = out
        self.workers = workers
        self.t0 = self.last = time.perf_counter()

    def restart(self) -> None:
        self.last = time.perf_counter()

    def stage(self, name: str) -> float:
        now = time.perf_counter()
        seconds = now - self.last
        print(format_stage_line(self.pipeline, name, seconds), file=self.out)
        self.last = time.perf_counter()
        return seconds

    def report(self, name: str, seconds: float) -> None:
        print(format_stage_line(self.pipeline, name, seconds), file=self.out)

    def total(self, name: str) -> float:
        seconds = time.perf_counter() - self.t0
        print(format_stage_line(self.pipeline, name, seconds), file=self.out)
        return seconds

    def report_resources(self) -> None:
        print(f"[Timing] {self.pipeline} workers={self.workers}", file=self.out)
        rss = peak_rss_bytes()
        print(f"[Timing] {self.pipeline} peak_rss_bytes={rss if rss is not None else 'unavailable'}", file=self.out)
