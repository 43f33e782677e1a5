"""The six benchmark programs end to end on the GPU: reference names, the
reference's own inputs and verify rule, CSV header, and the seeded variants.
Each program verifies itself (FP64 on the GPU); these tests check that it
reports PASS, and cross-check one output column against the CPU oracle's
closed form."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "sycl-bench_amd", "bin")
GOLDEN = open(os.path.join(ROOT, "tests", "golden", "reference_csv_header.txt")).read().strip()


def bench(name, *args, timeout=240):
    print(f"running {name} {' '.join(args)}", flush=True)
    r = subprocess.run([os.path.join(BIN, name), *args], capture_output=True, text=True, timeout=timeout)
    res = {}
    for line in r.stdout.splitlines():
        if line.startswith("*"):
            res.setdefault("_names", []).append(line.split("Results for ")[1].rstrip("*"))
        elif ": " in line:
            k, v = line.split(": ", 1)
            res[k] = v
    return r, res


REFERENCE_RUNS = [
    ("gemm", "Polybench_Gemm", 1024),
    ("2mm", "Polybench_2mm", 512),
    ("3mm", "Polybench_3mm", 512),
    ("syrk", "Polybench_Syrk", 512),
    ("syr2k", "Polybench_Syr2k", 512),
    ("matmulchain", "MatmulChain", 512),
    ("covariance", "Polybench_Covariance", 512),
    ("correlation", "Polybench_Correlation", 512),
]


@pytest.mark.parametrize("exe,name,n", REFERENCE_RUNS)
def test_reference_configuration_passes_reference_verify(exe, name, n):
    r, res = bench(exe, f"--size={n}", "--num-runs=2", "--queue-profiling")
    assert r.returncode == 0, r.stderr
    assert res["_names"] == [name]
    assert res["Verification"] == "PASS"
    assert res["device-name"]
    assert res["kernel-time-mean"] != "N/A"
    assert res["throughput-metric"].endswith("[FLOP]")


@pytest.mark.parametrize("exe,n", [("gemm", 300), ("syrk", 300), ("syr2k", 257), ("2mm", 100), ("3mm", 130)])
def test_ragged_sizes(exe, n):
    r, res = bench(exe, f"--size={n}", "--num-runs=1")
    assert r.returncode == 0, r.stderr
    assert res["Verification"] == "PASS"


@pytest.mark.parametrize("prec", ["fp32", "split-bf16", "bf16"])
@pytest.mark.parametrize("exe", ["gemm", "2mm", "3mm", "syrk", "syr2k"])
def test_seeded_variants(exe, prec):
    r, res = bench(exe, "--size=512", "--num-runs=1", "--init=uniform", f"--precision={prec}", "--metrics=extended")
    assert r.returncode == 0, r.stderr
    assert res["Verification"] == "PASS", res
    assert res["init"] == "uniform" and res["seed"] == "1"
    err = float(res["rel-frobenius-error"])
    assert err < {"fp32": 1e-5, "split-bf16": 2e-5, "bf16": 3e-2}[prec]
    suffix = "" if prec == "fp32" else "_" + prec.replace("-", "")
    assert res["_names"][0].endswith(suffix + "_uniform")


def test_split_bf16_on_reference_inputs_passes_percent_diff():
    r, res = bench("gemm", "--size=1024", "--num-runs=1", "--precision=split-bf16", "--metrics=extended")
    assert r.returncode == 0, r.stderr
    assert res["Verification"] == "PASS" and res["percent-diff-failures"] == "0"
    assert res["_names"] == ["Polybench_Gemm_splitbf16"]


def test_matmulchain_uniform_chain_of_eight():
    r, res = bench("matmulchain", "--size=512", "--num-runs=1", "--init=uniform", "--chain-length=8",
                   "--precision=split-bf16", "--metrics=extended")
    assert r.returncode == 0, r.stderr
    assert res["Verification"] == "PASS"
    assert float(res["rel-frobenius-error"]) < 1e-5 * 7 ** 0.5


@pytest.mark.parametrize("init", ["polybench", "uniform"])
def test_matmulchain_grouped_launch(init):
    r, res = bench("matmulchain", "--size=512", "--num-runs=2", f"--init={init}", "--chain-length=8", "--grouped",
                   "--precision=split-bf16", "--queue-profiling")
    assert r.returncode == 0, r.stderr
    assert res["Verification"] == "PASS"
    assert res["_names"][0].endswith("_len8_grouped")


@pytest.mark.parametrize("exe,n", [("covariance", 300), ("correlation", 257), ("covariance", 1024),
                                   ("correlation", 1024)])
@pytest.mark.parametrize("prec,init", [("fp32", "polybench"), ("split-bf16", "polybench"), ("bf16", "uniform"),
                                       ("split-bf16", "uniform")])
def test_covariance_correlation_programs(exe, n, prec, init):
    r, res = bench(exe, f"--size={n}", "--num-runs=2", f"--precision={prec}", f"--init={init}", "--metrics=extended")
    assert r.returncode == 0, r.stderr
    assert res["Verification"] == "PASS", res


@pytest.mark.parametrize("exe", ["2mm", "3mm"])
@pytest.mark.parametrize("init,prec", [("polybench", "fp32"), ("uniform", "split-bf16"), ("uniform", "fp32")])
def test_2mm_3mm_grouped_launch(exe, init, prec):
    r, res = bench(exe, "--size=384", "--num-runs=2", f"--init={init}", f"--precision={prec}", "--grouped",
                   "--metrics=extended")
    assert r.returncode == 0, r.stderr
    assert res["Verification"] == "PASS", res
    assert res["_names"][0].endswith("_grouped")


def test_csv_output_header(tmp_path):
    csv = tmp_path / "out.csv"
    r, _ = bench("gemm", "--size=256", "--num-runs=2", f"--output={csv}")
    assert r.returncode == 0, r.stderr
    lines = csv.read_text().splitlines()
    assert lines[0] == GOLDEN
    assert lines[1].startswith("Polybench_Gemm,PASS,")


def test_warmup_run_drops_first_sample():
    _, res = bench("gemm", "--size=256", "--num-runs=3", "--warmup-run")
    assert res["run-time-samples"].count(" ") == 1  # 3 runs, first dropped


def test_device_cpu_is_rejected():
    r, _ = bench("gemm", "--size=64", "--device=cpu")
    assert r.returncode != 0 and "oracle" in r.stderr


@pytest.mark.parametrize("exe", ["gemm", "2mm", "3mm", "syrk", "syr2k", "matmulchain", "covariance", "correlation"])
def test_edge_sizes(exe):
    # sizes below one MFMA block, one off a block edge on either side, and a seeded odd size
    for n in (1, 2, 7, 31, 33, 65):
        r, res = bench(exe, f"--size={n}", "--num-runs=1")
        assert r.returncode == 0, (n, r.stderr)
        assert res["Verification"] == "PASS", n
    r, res = bench(exe, "--size=33", "--num-runs=1", "--init=uniform", "--precision=split-bf16")
    assert r.returncode == 0 and res["Verification"] == "PASS", r.stderr


@pytest.mark.parametrize("prec", ["fp8-block", "fp8-block2"])
@pytest.mark.parametrize("exe", ["gemm", "2mm", "3mm", "matmulchain"])
def test_fp8_variants(exe, prec):
    r, res = bench(exe, "--size=512", "--num-runs=1", "--init=uniform", f"--precision={prec}", "--metrics=extended")
    assert r.returncode == 0, r.stderr
    assert res["Verification"] == "PASS", res
    # the programs' FP8 tolerances: one term 0.25, two terms 0.05 (per chained product: x sqrt(stages))
    assert float(res["rel-frobenius-error"]) < {"fp8-block": 0.25, "fp8-block2": 0.05}[prec] * 3 ** 0.5
    assert res["_names"][0].endswith("_" + prec.replace("-", "") + "_uniform")


def test_fp8_lognormal_gemm_and_grouped_refusal():
    r, res = bench("gemm", "--size=1024", "--num-runs=1", "--init=lognormal", "--precision=fp8-block2",
                   "--metrics=extended")
    assert r.returncode == 0 and res["Verification"] == "PASS", r.stderr
    # FP8 has no grouped kernel: the grouped form says so instead of running
    r, _ = bench("3mm", "--size=256", "--num-runs=1", "--precision=fp8-block", "--grouped")
    assert r.returncode != 0
    assert "grouped" in (r.stderr + r.stdout).lower()
