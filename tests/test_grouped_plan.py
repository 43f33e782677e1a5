"""The grouped-launch planner, checked on the CPU: tests/grouped_plan_check.cpp
plans many problem lists (the batched uneven workload with its chain, ragged
and transposed problems, 2mm / 3mm DAGs, an odd chain, empty and K = 0
problems) for every precision and tile policy and checks that every output
tile is planned exactly once, every operand packed exactly once before its
tiles run (by its segment's pre-pass or the previous segment's pre-pack
items), every dependency ordered, and the queue, job and block tables
consistent.  Built with hipcc against libmfx.so; runs without a GPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "sycl-bench_amd")
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_grouped_plans_are_consistent(tmp_path):
    lib = os.path.join(PKG, "libmfx.so")
    if not os.path.exists(lib):
        pytest.fail("libmfx.so is not built")
    exe = str(tmp_path / "grouped_plan_check")
    cmd = [HIPCC, "--offload-arch=gfx950", "-O1", "-std=c++17", "-I", os.path.join(PKG, "csrc"),
           os.path.join(ROOT, "tests", "grouped_plan_check.cpp"), "-o", exe, "-L", PKG, "-lmfx",
           "-Wl,-rpath," + PKG]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stderr[-3000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    print(r.stdout, r.stderr[-3000:])
    assert r.returncode == 0 and r.stdout.startswith("OK")
