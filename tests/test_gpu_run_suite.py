"""The suite runner over the real programs on the GPU: the gpu-smoke profile
(every program once at --size=256, two runs each) writes one results file in
which every benchmark verified, and the one-table merge holds every row."""
import os

import pytest

from test_run_suite import rs

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "sycl-bench_amd", "bin")


def test_gpu_smoke_profile(tmp_path):
    out, table = str(tmp_path / "sycl-bench.csv"), str(tmp_path / "table.csv")
    print("run-suite gpu-smoke", flush=True)
    assert rs.main(["gpu-smoke", "--bin-dir", BIN, "--output", out, "--table", table, "--timeout", "120"]) == 0
    rows = rs.read_results(out)
    names = {r["Benchmark name"] for r in rows}
    expected = {"Polybench_Gemm", "Polybench_2mm", "Polybench_3mm", "Polybench_Syrk", "Polybench_Syr2k",
                "MatmulChain", "Polybench_Covariance", "Polybench_Correlation", "MicroBench_MFMA_fp32",
                "MicroBench_DRAM_float_1", "MicroBench_DRAM_double_1"}
    assert names == expected
    assert all(r["Verification"] == "PASS" for r in rows)
    assert all(r["problem-size"] == "256" for r in rows)
    assert all(r["kernel-time-mean"] not in ("", "N/A") for r in rows)
    assert len(open(table).read().splitlines()) == len(rows) + 1
