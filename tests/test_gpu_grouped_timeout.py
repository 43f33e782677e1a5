"""The grouped launch's give-up path: a bounded in-kernel wait that expires
ends every other wait of the launch promptly, every kernel exits normally,
and mfx_grouped_status reports MFX_ERROR_TIMEOUT.

The condition is forced from the host: MFX_SPIN_LIMIT=0 makes each wait give
up at its first unsatisfied poll, so a consumer that polls before its
producers have finished times out.  Runs in a fresh process (the limit is read
when a handle is created) under a time limit, printing each step.
"""
import os
import subprocess
import sys
import textwrap

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = textwrap.dedent(r"""
    import importlib.util, os, sys, time
    import torch
    spec = importlib.util.spec_from_file_location("mfx", os.path.join(sys.argv[1], "sycl-bench_amd", "mfx.py"))
    mfx = importlib.util.module_from_spec(spec); spec.loader.exec_module(mfx)
    n = 512
    mats = [torch.empty((n, n), device="cuda") for _ in range(8)]
    h = mfx.Handle(0)
    for j, x in enumerate(mats):
        h.fill_uniform(x, 3, j)
    scratch = [torch.empty((n, n), device="cuda") for _ in range(6)]
    out = torch.empty((n, n), device="cuda")
    # policy 3: 256x128 tiles, one persistent kernel with in-kernel dependency waits
    big, small, bound, per_level = mfx.grouped_policy_info(3)
    assert per_level == 0
    h.force_grouped_policy(3)
    timeouts = 0
    t0 = time.perf_counter()
    for prec in ("fp32", "split-bf16"):
        for rep in range(4):
            h.chain(mats, scratch, out, prec=prec)
            try:
                h.grouped_status()
                print(f"limit 0 {prec} launch {rep}: success", flush=True)
            except mfx.MfxError as e:
                assert e.status == 6, e  # MFX_ERROR_TIMEOUT
                timeouts += 1
                print(f"limit 0 {prec} launch {rep}: timeout reported", flush=True)
    elapsed = time.perf_counter() - t0
    torch.cuda.synchronize()
    print(f"8 launches with limit 0 in {elapsed:.2f} s, {timeouts} timeouts", flush=True)
    h.close()
    # a handle with the default limit recovers: same chain, correct result
    os.environ.pop("MFX_SPIN_LIMIT")
    h = mfx.Handle(0)
    h.force_grouped_policy(3)
    h.chain(mats, scratch, out, prec="split-bf16")
    h.grouped_status()
    ref = [m.double() for m in mats]
    while len(ref) > 1:
        ref = [ref[i] @ ref[i + 1] for i in range(0, len(ref), 2)]
    err = (torch.linalg.norm(out.double() - ref[0]) / torch.linalg.norm(ref[0])).item()
    print(f"default limit: chain rel err {err:.3e}", flush=True)
    h.close()
    print(f"RESULT timeouts={timeouts} elapsed={elapsed:.3f} err={err:.3e}", flush=True)
""")


def test_expired_wait_ends_the_launch_promptly():
    env = dict(os.environ, MFX_SPIN_LIMIT="0")
    r = subprocess.run([sys.executable, "-u", "-c", CHILD, ROOT], env=env, capture_output=True, text=True,
                       timeout=100)
    print(r.stdout)
    print(r.stderr[-2000:])
    assert r.returncode == 0
    line = [x for x in r.stdout.splitlines() if x.startswith("RESULT")][-1]
    fields = dict(kv.split("=") for kv in line.split()[1:])
    # consumers poll while their producers still run, so with a zero limit
    # the launches give up; each gives up within a fraction of a second
    assert int(fields["timeouts"]) >= 1
    assert float(fields["elapsed"]) < 20.0
    # three chained split-BF16 levels: sqrt(3) x the single-product bar of 1e-5
    assert float(fields["err"]) < 1e-5 * 3 ** 0.5
