"""The suite runner (sycl-bench_amd/run-suite) on the CPU, with stand-in
programs: the reference's profile defaults and range helpers
(bin/run-suite:10-24, :41-49), the argument order, the size-growth and
abort-on-failure rules, the refusal to overwrite a results file, the sweep of
list-valued options, and the one-table merge of the results file."""
import importlib.machinery
import importlib.util
import os
import stat
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PATH = os.path.join(ROOT, "sycl-bench_amd", "run-suite")


def load():
    loader = importlib.machinery.SourceFileLoader("run_suite", PATH)
    spec = importlib.util.spec_from_loader("run_suite", loader)
    mod = importlib.util.module_from_spec(spec)
    loader.exec_module(mod)
    return mod


rs = load()

# A stand-in program: appends its arguments to a CSV block like the real
# programs' CSV output, fails for --size above FAIL_ABOVE.
FAKE = '''#!{py}
import sys
args = dict(a.split("=", 1) for a in sys.argv[1:] if "=" in a)
flags = sorted(a for a in sys.argv[1:] if "=" not in a)
if int(args["--size"]) > {fail_above}:
    sys.exit(3)
with open(args["--output"], "a") as f:
    f.write("# Benchmark name,problem-size,precision,flags,Verification\\n")
    f.write("{name}_%s,%s,%s,%s,PASS\\n" % (args.get("--precision", "fp32"), args["--size"],
                                           args.get("--precision", "-"), " ".join(flags)))
'''


def make_bin(tmp_path, specs):
    d = tmp_path / "bin"
    d.mkdir()
    for name, fail_above in specs.items():
        p = d / name
        p.write_text(FAKE.format(py=sys.executable, name=name, fail_above=fail_above))
        p.chmod(p.stat().st_mode | stat.S_IEXEC)
    (d / "notes.txt").write_text("not a program")
    return str(d)


def test_reference_ranges_and_default_profile():
    assert rs.create_log_range(1024, 1024) == [1024]
    assert rs.create_log_range(1024, 1500) == [1024, 2048]
    assert rs.create_log_range(2 ** 10, 2 ** 13) == [1024, 2048, 4096, 8192]
    assert rs.create_linear_range(0, 10, 5) == [0, 2, 4, 6, 8]
    assert rs.create_linear_range(0, 3, 5) == [0, 1, 2, 3, 4]
    opts = rs.profiles["default"]["default-options"]
    assert opts == {"--size": [1024], "--local": [256], "--num-runs": 10, "--output": "./sycl-bench.csv"}
    assert rs.profiles["default"]["max-allowed-runtime"] == 10800.0
    assert rs.profiles["quicktest"]["max-allowed-runtime"] == 1.0
    assert rs.profiles["gpu-warmup"]["default-flags"] == {"--warmup-run"}
    assert rs.profiles["gpu-noverify"]["default-options"]["--device"] == "gpu"


def test_argument_order_and_sweep():
    options = {"--size": [1024], "--local": [256], "--num-runs": 10, "--precision": ["fp32", "bf16"],
               "--output": "x.csv"}
    combos = rs.sweep_values(options)
    assert combos == [["--precision=fp32"], ["--precision=bf16"]]
    args = rs.build_args({"--b", "--a"}, options, 1024, 256, combos[1])
    assert args == ["--a", "--b", "--num-runs=10", "--output=x.csv", "--size=1024", "--local=256",
                    "--precision=bf16"]


def test_is_benchmark(tmp_path):
    d = make_bin(tmp_path, {"gemm": 10 ** 9})
    assert rs.is_benchmark(os.path.join(d, "gemm"))
    assert not rs.is_benchmark(os.path.join(d, "notes.txt"))


def test_sweep_appends_one_results_file_and_merges(tmp_path):
    d = make_bin(tmp_path, {"gemm": 10 ** 9, "syrk": 10 ** 9})
    out = str(tmp_path / "res.csv")
    prof = rs.construct_profile({"--size": [256, 512], "--num-runs": 1}, ["--queue-profiling"],
                                individual_options={"gemm": {"--precision": ["fp32", "split-bf16"]}})
    assert rs.run_profile(prof, d, out) == []
    rows = rs.read_results(out)
    names = [r["Benchmark name"] for r in rows]
    assert names == ["gemm_fp32", "gemm_split-bf16", "gemm_fp32", "gemm_split-bf16", "syrk_fp32", "syrk_fp32"]
    assert [r["problem-size"] for r in rows] == ["256", "256", "512", "512", "256", "512"]
    assert all(r["flags"] == "--queue-profiling" for r in rows)
    table = str(tmp_path / "table.csv")
    rs.write_table(rows, table)
    lines = open(table).read().splitlines()
    assert lines[0] == "Benchmark name,Verification,flags,precision,problem-size"
    assert len(lines) == 7


def test_failure_skips_larger_sizes_and_is_reported(tmp_path, capsys):
    d = make_bin(tmp_path, {"gemm": 300, "syrk": 10 ** 9})
    out = str(tmp_path / "res.csv")
    prof = rs.construct_profile({"--size": [256, 512, 1024], "--num-runs": 1})
    assert rs.run_profile(prof, d, out) == ["gemm"]
    rows = rs.read_results(out)
    assert [(r["Benchmark name"], r["problem-size"]) for r in rows] == [
        ("gemm_fp32", "256"), ("syrk_fp32", "256"), ("syrk_fp32", "512"), ("syrk_fp32", "1024")]
    assert "--size=1024" not in [l for l in capsys.readouterr().out.splitlines() if l.startswith("gemm")][-1]


def test_max_allowed_runtime_stops_growth(tmp_path):
    d = make_bin(tmp_path, {"gemm": 10 ** 9})
    out = str(tmp_path / "res.csv")
    prof = rs.construct_profile({"--size": [256, 512], "--num-runs": 1}, max_allowed_runtime=0.0)
    assert rs.run_profile(prof, d, out) == []
    assert not os.path.exists(out)
    prof = rs.construct_profile({"--size": [256, 512], "--num-runs": 1}, max_allowed_runtime=1e-9)
    rs.run_profile(prof, d, out)
    assert [r["problem-size"] for r in rs.read_results(out)] == ["256"]


def test_size_not_multiple_of_local_is_skipped(tmp_path):
    d = make_bin(tmp_path, {"gemm": 10 ** 9})
    out = str(tmp_path / "res.csv")
    prof = rs.construct_profile({"--size": [300, 512], "--num-runs": 1})
    rs.run_profile(prof, d, out)
    assert [r["problem-size"] for r in rs.read_results(out)] == ["512"]


def test_main_refuses_existing_output_and_cpu_profiles(tmp_path, capsys):
    d = make_bin(tmp_path, {"gemm": 10 ** 9})
    out = tmp_path / "res.csv"
    out.write_text("")
    assert rs.main(["gpu", "--bin-dir", d, "--output", str(out)]) == 255
    assert "already exists" in capsys.readouterr().out
    assert rs.main(["cpu", "--bin-dir", d]) == 255
    assert "GPU only" in capsys.readouterr().out
    assert rs.main(["nonsense", "--bin-dir", d]) == 255
    assert rs.main([]) == 255


def test_main_runs_profile_and_writes_table(tmp_path):
    d = make_bin(tmp_path, {"gemm": 10 ** 9, "3mm": 10 ** 9})
    out, table = str(tmp_path / "r.csv"), str(tmp_path / "t.csv")
    assert rs.main(["gpu-grouped", "--bin-dir", d, "--output", out, "--table", table]) == 0
    rows = rs.read_results(out)
    # gpu-grouped runs 2mm / 3mm / matmulchain only, three precisions each
    assert [r["Benchmark name"] for r in rows] == ["3mm_fp32", "3mm_split-bf16", "3mm_bf16"]
    assert all("--grouped" in r["flags"] and "--queue-profiling" in r["flags"] for r in rows)
    assert len(open(table).read().splitlines()) == 4


def test_precision_lists_match_what_the_library_supports():
    # FP8 is built for plain gemm only (no syrk / syr2k / statistics); the
    # grouped launch takes FP32 and the BF16 forms.
    for name in ("syrk", "syr2k", "covariance", "correlation"):
        assert not any(p.startswith("fp8") for p in rs.PRECISIONS[name])
    assert rs.profiles["gpu-grouped"]["default-options"]["--precision"] == ["fp32", "split-bf16", "bf16"]


def test_dry_run_lists_every_command(tmp_path, capsys):
    d = make_bin(tmp_path, {"gemm": 10 ** 9, "mfma_peak": 10 ** 9, "hbm_copy": 10 ** 9})
    assert rs.main(["gpu-precisions", "--bin-dir", d, "--dry-run"]) == 0
    cmds = [l for l in capsys.readouterr().out.splitlines() if l.split(" ")[0] in ("gemm", "mfma_peak", "hbm_copy")]
    assert len(cmds) == 1 + 5 + 3
