"""bench.py's rocprofv3 summaries on synthetic CSV runs: the profiled pass is
cut out between the two groups of window-marker dispatches, kernel names are
shortened and mapped to their families, the traffic table is scaled by the
copy calibration, and a table stamped with other kernel sources is not used.
"""
import csv
import json
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402

TRACE_COLS = ["Kind", "Agent_Id", "Queue_Id", "Stream_Id", "Thread_Id", "Dispatch_Id", "Kernel_Id", "Kernel_Name",
              "Correlation_Id", "Start_Timestamp", "End_Timestamp", "LDS_Block_Size", "Scratch_Size", "VGPR_Count",
              "Accum_VGPR_Count", "SGPR_Count", "Workgroup_Size_X", "Workgroup_Size_Y", "Workgroup_Size_Z",
              "Grid_Size_X", "Grid_Size_Y", "Grid_Size_Z"]
PMC_COLS = ["Correlation_Id", "Dispatch_Id", "Agent_Id", "Queue_Id", "Process_Id", "Thread_Id", "Grid_Size", "Kernel_Id",
            "Kernel_Name", "Workgroup_Size", "LDS_Block_Size", "Scratch_Size", "VGPR_Count", "Accum_VGPR_Count",
            "SGPR_Count", "Counter_Name", "Counter_Value", "Start_Timestamp", "End_Timestamp"]

MARK = "(anonymous namespace)::k_copy_probe(HIP_vector_type<float, 4u> const*, HIP_vector_type<float, 4u>*, unsigned long)"
DELTA = ("void pbf::(anonymous namespace)::k_delta<256, false, false, false>(pbf::SceneConsts const*, "
         "HIP_vector_type<unsigned int, 4u> const*)")
LAMBDA = "void pbf::(anonymous namespace)::k_lambda<512, false, false>(pbf::SceneConsts const*)"
TARE = "pbf::k_tare(unsigned int)"


def _dispatches():
    """(name, start ns, end ns): a delta launch before the window (60 us), the
    two-dispatch marker groups, and inside them two deltas (50 and 70 us), a
    lambda (40 us) and the empty tare kernel (4 us)."""
    d = [(DELTA, 1_000, 61_000), (MARK, 100_000, 101_000), (MARK, 102_000, 103_000),
         (DELTA, 200_000, 250_000), (LAMBDA, 260_000, 300_000), (DELTA, 310_000, 380_000),
         (TARE, 390_000, 394_000), (MARK, 900_000, 901_000), (MARK, 902_000, 903_000),
         (DELTA, 1_000_000, 1_100_000)]
    return d


def _write_trace(path: Path):
    path.mkdir(parents=True)
    with open(path / "kt_kernel_trace.csv", "w", newline="") as fh:
        w = csv.writer(fh, quoting=csv.QUOTE_NONNUMERIC)
        w.writerow(TRACE_COLS)
        for q, (name, t0, t1) in enumerate(_dispatches()):
            w.writerow(["KERNEL_DISPATCH", "Agent 1", 2, 0, 1, q + 1, 7, name, q + 1, t0, t1, 0, 0, 8, 0, 16,
                        256, 1, 1, 1024, 1, 1])


def _write_pmc(path: Path, counter: str, value_of):
    path.mkdir(parents=True)
    with open(path / "p_counter_collection.csv", "w", newline="") as fh:
        w = csv.writer(fh, quoting=csv.QUOTE_NONNUMERIC)
        w.writerow(PMC_COLS)
        for q, (name, t0, t1) in enumerate(_dispatches()):
            w.writerow([q + 1, q + 1, "Agent 1", 2, 1, 1, 1024, 7, name, 256, 0, 0, 8, 0, 16, counter,
                        value_of(name, q), t0, t1])


def _write_copy(path: Path, counter: str, kib: float):
    path.mkdir(parents=True)
    with open(path / "c_counter_collection.csv", "w", newline="") as fh:
        w = csv.writer(fh, quoting=csv.QUOTE_NONNUMERIC)
        w.writerow(PMC_COLS)
        for q in range(4):   # warm-up, then three timed copies
            w.writerow([q + 1, q + 1, "Agent 1", 2, 1, 1, 524288, 7, MARK, 256, 0, 0, 8, 0, 16, counter,
                        kib * (1.5 if q == 0 else 1.0), 1000 * q, 1000 * q + 500])


def _line(tmp_path: Path) -> Path:
    line = {"metric": "m", "value": 1.0,
            "roofline": {"kernel": "computeDelta", "algorithmic_bytes_per_launch": 6.0e6, "peak": 8000.0,
                         "frac": 6.0e6 / 60e-6 / 1e9 / 8000.0, "avg_launch_ms": 0.061, "empty_kernel_ms": 0.004,
                         "kernel_source_sha16": "feedfacecafebeef",
                         "by_kernel": {"computeDelta": {"avg_launch_ms": 0.061},
                                       "computeScaling": {"avg_launch_ms": 0.040}}}}
    p = tmp_path / "line.json"
    p.write_text("[bench] progress\n" + json.dumps(line) + "\n")
    return p


def test_kernel_names_shorten_to_their_templates():
    assert bench._kernel_short(DELTA) == "k_delta<256, false, false, false>"
    assert bench._kernel_short(MARK) == "k_copy_probe"
    assert bench._kernel_short(TARE) == "k_tare"
    assert bench._family_of_kernel("k_friends_direct<256, true, false>") == "buildFriendsList"
    assert bench._family_of_kernel("k_friends<64, true>") == "buildFriendsList"
    assert bench._family_of_kernel("k_gather_cells") == "sortParticles"
    assert bench._family_of_kernel("k_delta_b") is None   # the adjoint's kernel, not a forward stage
    assert bench._family_of_kernel("k_copy_probe") is None


def test_window_is_between_the_marker_groups(tmp_path):
    _write_trace(tmp_path / "kt")
    rows = bench._read_dispatches(str(tmp_path / "kt"))
    lo, hi = bench._window(rows)
    assert (lo, hi) == (103_000, 900_000)
    win = bench._in_window(rows)
    assert [d["kernel"].split("<")[0] for d in win] == ["k_delta", "k_lambda", "k_delta"]
    assert [d["kernel"] for d in bench._in_window(rows, keep="k_tare")] == ["k_tare"]


def test_trace_summary_compares_the_pass_with_the_line(tmp_path):
    _write_trace(tmp_path / "kt")
    out = bench.trace_summary(str(tmp_path / "kt"), str(_line(tmp_path)))
    # computeDelta: (50 + 70) / 2 = 60 us in the pass (not the 60 and 100 us outside it)
    assert "| computeDelta | 2 | 60.000 | 61.000 | 1.000 | 0.9836 |" in out
    assert "| computeScaling | 1 | 40.000 | 40.000 | 0.000 | 1.0000 |" in out
    assert "4.000 us each in the trace, 4.000 us by the line" in out
    assert "(ratio 1.0000)" in out


def test_pmc_summary_scales_by_the_copy_calibration(tmp_path):
    # FETCH_SIZE reads half the bytes a copy moves, WRITE_SIZE all of them
    _write_pmc(tmp_path / "pf", "FETCH_SIZE", lambda name, q: 1000.0 if "k_delta" in name else 500.0)
    _write_pmc(tmp_path / "pw", "WRITE_SIZE", lambda name, q: 100.0 if "k_delta" in name else 50.0)
    _write_copy(tmp_path / "cf", "FETCH_SIZE", 131072.0)
    _write_copy(tmp_path / "cw", "WRITE_SIZE", 262144.0)
    out = bench.pmc_summary(str(tmp_path / "pf"), str(tmp_path / "pw"), str(tmp_path / "cf"), str(tmp_path / "cw"),
                            256, str(_line(tmp_path)))
    assert "FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000" in out
    table = json.loads(out.split("PMC_TRAFFIC = ", 1)[1].split("```", 1)[0])
    assert table["source_sha16"] == "feedfacecafebeef"
    assert table["bytes_per_launch"] == {"computeDelta": (1000 * 2 + 100) * 1024,
                                         "computeScaling": (500 * 2 + 50) * 1024}


def test_traffic_table_of_other_sources_is_not_used(monkeypatch):
    monkeypatch.setitem(bench.PMC_TRAFFIC, "source_sha16", "0000000000000000")
    assert bench.pmc_traffic("computeDelta", 1048576) is None
    monkeypatch.setitem(bench.PMC_TRAFFIC, "source_sha16", bench.kernel_source_sha16())
    monkeypatch.setitem(bench.PMC_TRAFFIC, "bytes_per_launch", {"computeDelta": 123})
    assert bench.pmc_traffic("computeDelta", 1048576) == 123
    assert bench.pmc_traffic("computeDelta", 4096) is None


def test_source_stamp_ignores_comments_and_indentation_only():
    a = '#define K 4\nint f(int x) {  // twice\n    return x * 2; /* a\n block */\n}\nconst char* s = "// kept";\n'
    b = '#define K 4\n\n// a new comment line\nint f(int x) {\n  return x * 2;\n}\nconst char* s = "// kept";\n'
    assert bench.source_code_tokens(a) == bench.source_code_tokens(b)
    assert bench.source_code_tokens(a).splitlines()[0] == "#define K 4"
    for changed in (a.replace("x * 2", "x * 3"), a.replace('"// kept"', '"// kept "'),
                    a.replace("#define K 4\nint", "#define K 4 int")):
        assert bench.source_code_tokens(changed) != bench.source_code_tokens(a)
    assert bench.source_code_tokens("char c = '\\'';  // it's\nint y;") == "char c = '\\'';\nint y;"


def test_a_run_without_markers_is_refused(tmp_path):
    d = tmp_path / "kt"
    d.mkdir()
    with open(d / "x_kernel_trace.csv", "w", newline="") as fh:
        w = csv.writer(fh, quoting=csv.QUOTE_NONNUMERIC)
        w.writerow(TRACE_COLS)
        w.writerow(["KERNEL_DISPATCH", "Agent 1", 2, 0, 1, 1, 7, DELTA, 1, 10, 20, 0, 0, 8, 0, 16, 256, 1, 1,
                    1024, 1, 1])
    with pytest.raises(SystemExit):
        bench.trace_summary(str(d), None)
