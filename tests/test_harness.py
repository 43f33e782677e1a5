"""The C++ harness (CPU only): CLI, run loop, statistics and CSV output against
the reference's behaviour, through a fake device and fake benchmarks."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "sycl-bench_amd", "harness")
GOLDEN = open(os.path.join(ROOT, "tests", "golden", "reference_csv_header.txt")).read().strip()


@pytest.fixture(scope="module")
def harness_exe(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("harness") / "harness_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", HARNESS,
                    os.path.join(ROOT, "tests", "harness", "harness_check.cpp"), "-o", out, "-lpthread"], check=True)
    return out


def run(harness_exe, scenario, *args, env=None):
    e = dict(os.environ, HARNESS_SCENARIO=scenario, **(env or {}))
    return subprocess.run([harness_exe, *args], capture_output=True, text=True, env=e)


def parse_stdout(text):
    res = {}
    for line in text.splitlines():
        if ": " in line and not line.startswith("*"):
            k, v = line.split(": ", 1)
            res[k] = v
    return res


def test_csv_header_matches_reference(harness_exe, tmp_path):
    csv = tmp_path / "out.csv"
    r = run(harness_exe, "plain", "--size=64", "--num-runs=3", f"--output={csv}")
    assert r.returncode == 0, r.stderr
    lines = csv.read_text().splitlines()
    assert lines[0] == GOLDEN
    row = lines[1].split(",")
    cols = GOLDEN.split(",")
    assert len(row) == len(cols) or '"' in lines[1]  # the samples column is quoted with spaces
    fields = dict(zip(cols, _split_csv(lines[1])))
    assert fields["# Benchmark name"] == "Plain"
    assert fields["Verification"] == "PASS"
    assert fields["problem-size"] == "64" and fields["local-size"] == "256"
    assert fields["device-name"] == "Fake Device"
    assert fields["kernel-time-mean"] == "N/A" and fields["throughput-metric"] == "N/A"
    assert fields["run-time-samples"].count(" ") == 2  # three samples


def _split_csv(line):
    out, cur, quoted = [], "", False
    for ch in line:
        if ch == '"':
            quoted = not quoted
            cur += ch
        elif ch == "," and not quoted:
            out.append(cur)
            cur = ""
        else:
            cur += ch
    out.append(cur)
    return out


def test_csv_appends_one_header_per_flush(harness_exe, tmp_path):
    csv = tmp_path / "out.csv"
    for _ in range(2):
        assert run(harness_exe, "plain", "--num-runs=1", f"--output={csv}").returncode == 0
    lines = csv.read_text().splitlines()
    assert lines[0] == lines[2] == GOLDEN and len(lines) == 4


def test_queue_profiling_fills_kernel_times(harness_exe):
    r = run(harness_exe, "timed", "--num-runs=2", "--queue-profiling")
    res = parse_stdout(r.stdout)
    assert res["kernel-time-mean"] == "0.000100 [s]"
    assert res["submit-time-min"] == "0.000001 [s]"
    assert res["throughput-metric"] == "1000000000.000000 [FLOP]"
    assert res["kernel-time-throughput"].endswith("[FLOP/s]")
    assert res["Verification"] == "N/A"  # no verify() method
    off = parse_stdout(run(harness_exe, "timed", "--num-runs=2").stdout)
    assert off["kernel-time-mean"] == "N/A" and off["system-time-median"] == "N/A"


def test_statistics_values(harness_exe):
    res = parse_stdout(run(harness_exe, "stats").stdout)
    # samples 1,2,3,4,10 ms: mean 4, sample stddev sqrt(50/4), median element 2 of the sorted list
    assert res["run-time-mean"] == "0.004000 [s]"
    assert res["run-time-median"] == "0.003000 [s]"
    assert res["run-time-min"] == "0.001000 [s]"
    assert res["run-time-stddev"] == "0.003536 [s]"
    assert res["run-time-samples"] == '"0.001000 0.002000 0.003000 0.004000 0.010000"'
    assert res["run-time-throughput"] == "1000000000000.000000 [FLOP/s]"  # 1e9 FLOP / min
    assert res["kernel-time-mean"] == "N/A"
    # --warmup-run drops the first sample (4 ms): 1,2,3,10 -> mean 4, median 3
    w = parse_stdout(run(harness_exe, "stats", env={"HARNESS_WARMUP": "1"}).stdout)
    assert w["run-time-samples"] == '"0.001000 0.002000 0.003000 0.010000"'


def test_verification_failure_stops_runs(harness_exe):
    r = run(harness_exe, "plain", "--num-runs=5", "--fail-at=2")
    assert parse_stdout(r.stdout)["Verification"] == "FAIL"
    assert "runs=2 verifies=2" in r.stderr
    assert r.returncode == 1


def test_no_verification_flag(harness_exe):
    r = run(harness_exe, "plain", "--num-runs=2", "--no-verification")
    assert parse_stdout(r.stdout)["Verification"] == "N/A"
    assert "verifies=0" in r.stderr
    r = run(harness_exe, "plain", "--num-runs=2", "--verification-range=0,1,1")
    assert parse_stdout(r.stdout)["Verification"] == "N/A"


def test_verification_range_parses_by_element_count(harness_exe):
    r = run(harness_exe, "plain", "--num-runs=1", "--verification-range=1,1,1", "--verification-begin=0,0")
    assert r.returncode == 0, r.stderr
    assert parse_stdout(r.stdout)["Verification"] == "PASS"


def test_fresh_object_per_run(harness_exe):
    r = run(harness_exe, "plain", "--num-runs=4")
    assert "setups=4 runs=4 verifies=4" in r.stderr


def test_throwing_benchmark_discards_row_and_continues(harness_exe, tmp_path):
    csv = tmp_path / "out.csv"
    r = run(harness_exe, "throwing", "--num-runs=1", f"--output={csv}")
    assert "setup exploded" in r.stderr
    assert r.returncode == 1
    lines = csv.read_text().splitlines()
    assert [l.split(",")[0] for l in lines[1:]] == ["Plain"]


def test_duplicate_names_rejected(harness_exe):
    r = run(harness_exe, "duplicate", "--num-runs=1")
    assert "has already been run" in r.stderr and r.returncode == 1
    assert r.stdout.count("Results for Plain") == 1


def test_bad_command_lines(harness_exe):
    r = run(harness_exe, "plain", "--size=10", "--size=20")
    assert "several times" in r.stderr and r.returncode == 1
    r = run(harness_exe, "plain", "--size=abc")
    assert r.returncode == 1 and "--size" in r.stderr
    r = run(harness_exe, "plain", "--device=tpu")
    assert r.returncode == 1 and "unknown device type" in r.stderr
    r = run(harness_exe, "plain", "--num-runs=0")
    assert r.returncode == 1


def test_several_benchmarks_share_one_csv_header(harness_exe, tmp_path):
    csv = tmp_path / "out.csv"
    assert run(harness_exe, "all", "--num-runs=1", f"--output={csv}").returncode == 0
    lines = csv.read_text().splitlines()
    headers = [l for l in lines if l.startswith("#")]
    assert all(h == GOLDEN for h in headers) and len(lines) == 6
