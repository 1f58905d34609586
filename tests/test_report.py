"""Output formats shared with the reference's tooling: the measure report, the
S2 history lines and the [Timing] lines (known answers from
src/pipeline/timing.rs:84-110 and the line patterns scripts/perf_matrix.py:30-32
parses)."""
import io
import re

from rustmspt_amd import report as R

# the three line shapes scripts/perf_matrix.py reads
TIMING_STAGE = re.compile(r"^\[Timing\] (\S+) stage=(\S+) seconds=([-+0-9.eE]+)\s*$")
TIMING_WORKERS = re.compile(r"^\[Timing\] (\S+) workers=(\d+)\s*$")
TIMING_RSS = re.compile(r"^\[Timing\] (\S+) peak_rss_bytes=(\d+|unavailable)\s*$")


def test_stage_line_matches_the_shared_format():
    assert R.format_stage_line("crop", "load", 1.5) == "[Timing] crop stage=load seconds=1.500000000"


def test_vm_hwm_parses_kb_and_refuses_anything_else():
    assert R.parse_vm_hwm("Name:\tx\nVmHWM:\t  2048 kB\nVmRSS:\t1 kB\n") == 2048 * 1024
    assert R.parse_vm_hwm("Name:\tx\n") is None
    assert R.parse_vm_hwm("VmHWM:\tabc kB\n") is None
    assert R.parse_vm_hwm("VmHWM:\t12 MB\n") is None
    assert R.peak_rss_bytes() > 0


def test_timer_lines_parse_with_the_perf_matrix_patterns():
    buf = io.StringIO()
    t = R.StageTimer("measure", buf)
    assert t.stage("load") >= 0.0
    t.restart()
    t.stage("s2_exact")
    t.report("views", 0.25)
    t.total("total_in_pool")
    t.report_resources()
    lines = buf.getvalue().splitlines()
    stages = [TIMING_STAGE.match(l) for l in lines[:4]]
    assert all(stages) and [m.group(2) for m in stages] == ["load", "s2_exact", "views", "total_in_pool"]
    assert stages[2].group(3) == "0.250000000"
    assert TIMING_WORKERS.match(lines[4]).group(2) == "1"
    assert TIMING_RSS.match(lines[5])


def test_measure_report_layout():
    one = R.measure_report(0.25, ["exact=hip/MI355X"], "exact", ["exact"], [[0.25, 0.125, 0.0]])
    assert one == ("Volume Fraction: 0.250000\nCompute Backend: exact=hip/MI355X\nMethod: exact\n"
                   "S2(0)-VF diff [exact]: 0.000000\nS2 Values:\n0: 0.250000\n1: 0.125000\n2: 0.000000\n")
    both = R.measure_report(0.5, ["exact=hip/x", "monte_carlo=hip/x"], "both", ["exact", "monte_carlo"],
                            [[0.5, 0.3], [0.5, 0.25, 0.1]])
    assert "S2 Values [exact]:\n0: 0.500000\n1: 0.300000\n" in both
    assert "S2 Values [monte_carlo]:\n0: 0.500000\n1: 0.250000\n2: 0.100000\n" in both
    # l2 over the common prefix (src/pipeline/measure.rs:27-39)
    assert both.endswith("L2 error [exact vs monte_carlo]: 0.050000\n")
    assert R.l2_error([], [1.0]) == 0.0


def test_history_lines(tmp_path):
    h = []
    R.push_history_s2(h, "Target S2", [0.3, 0.1234564, 0.05])
    R.push_history_best(h, 17, 0.0012345, [0.3, 0.12])
    assert h == ["Target S2: 0.300000 0.123456 0.050000", "Iter 17: Loss 0.001234 | S2 0.300000 0.120000"]
    R.write_history(tmp_path / "s2_history.txt", h)
    assert (tmp_path / "s2_history.txt").read_text() == "\n".join(h)
