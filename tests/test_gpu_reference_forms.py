"""Reference-form parity on the GPU, through the C ABI.

- 2mm and 3mm exactly as polybench/2mm.cpp:94-123 and polybench/3mm.cpp:115-158
  submit them (C += A*B then E = C*D; E = A*B, F = C*D, G = E*F, outputs
  accumulated from their initial contents), on the reference's own inputs,
  one mfx_gemm per kernel and as one grouped launch with the dependency
  edges the SYCL accessors imply.  Each result must pass the reference's
  verify rule (percentDiff > 0.05 % fails, polybenchUtilFuncts.h:27-42)
  against both the restated serial CPU loop (polybench/2mm.cpp:126-145,
  polybench/3mm.cpp:161-183) and the FP64 closed form.
- The gemm workload exactly as quoted: N = 1024, A/B/C seeded U[-1,1),
  alpha = 1.5, beta = 1.2, FP32, against the FP64 product (2e-6).
- The lognormal input generator of the FP8 workload against the oracle's
  restatement (same keyed hash and Box-Muller formula, FP64 there, FP32
  here: elementwise to 2e-5 relative, signs exact).
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = "cuda"


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


_CACHE = {}


def _cached(oracle, name, n):
    """The restated serial loop's answer and the closed form, computed once per (name, n)."""
    if (name, n) not in _CACHE:
        ref = oracle.mm2_reference(n) if name == "2mm" else oracle.mm3_reference(n)
        _CACHE[(name, n)] = (ref, oracle.closed_form(name, n))
    return _CACHE[(name, n)]


def _assert_reference_verify(oracle, got, cpu_answer, closed):
    assert oracle.percent_diff_count(cpu_answer, got) == 0
    assert oracle.percent_diff_count(closed.astype(np.float32), got) == 0


@pytest.mark.parametrize("grouped", [False, True], ids=["per-call", "grouped"])
@pytest.mark.parametrize("prec", ["fp32", "split-bf16"])
@pytest.mark.parametrize("n", [64, 256, 1024])
def test_2mm_reference_form(handle, mfx, oracle, n, prec, grouped):
    (A, B, Cin, D, E_cpu), closed = _cached(oracle, "2mm", n)
    Ad, Bd, Dd = to_dev(A), to_dev(B), to_dev(D)
    Cd = to_dev(Cin)
    Ed = torch.full((n, n), float("nan"), device=DEV)  # discard_write: never read
    print(f"2mm reference form n={n} prec={prec} grouped={grouped}", flush=True)
    if grouped:
        handle.grouped([mfx.problem(Ad, Bd, Cd, alpha=1.0, beta=1.0), mfx.problem(Cd, Dd, Ed, alpha=1.0, beta=0.0)],
                       deps=[(1, 0)], prec=prec)
        handle.grouped_status()
    else:
        handle.gemm(Ad, Bd, Cd, alpha=1.0, beta=1.0, prec=prec)
        handle.gemm(Cd, Dd, Ed, alpha=1.0, beta=0.0, prec=prec)
    torch.cuda.synchronize()
    _assert_reference_verify(oracle, Ed.cpu().numpy(), E_cpu, closed)


@pytest.mark.parametrize("grouped", [False, True], ids=["per-call", "grouped"])
@pytest.mark.parametrize("prec", ["fp32", "split-bf16"])
@pytest.mark.parametrize("n", [64, 256, 1024])
def test_3mm_reference_form(handle, mfx, oracle, n, prec, grouped):
    (A, B, C, D, G_cpu), closed = _cached(oracle, "3mm", n)
    Ad, Bd, Cd, Dd = (to_dev(x) for x in (A, B, C, D))
    # E, F, G are zero-initialised and accumulated (+=) in the reference
    Ed, Fd, Gd = (torch.zeros((n, n), device=DEV) for _ in range(3))
    print(f"3mm reference form n={n} prec={prec} grouped={grouped}", flush=True)
    probs = [mfx.problem(Ad, Bd, Ed, beta=1.0), mfx.problem(Cd, Dd, Fd, beta=1.0), mfx.problem(Ed, Fd, Gd, beta=1.0)]
    if grouped:
        handle.grouped(probs, deps=[(2, 0), (2, 1)], prec=prec)
        handle.grouped_status()
    else:
        handle.gemm(Ad, Bd, Ed, alpha=1.0, beta=1.0, prec=prec)
        handle.gemm(Cd, Dd, Fd, alpha=1.0, beta=1.0, prec=prec)
        handle.gemm(Ed, Fd, Gd, alpha=1.0, beta=1.0, prec=prec)
    torch.cuda.synchronize()
    _assert_reference_verify(oracle, Gd.cpu().numpy(), G_cpu, closed)


def test_gemm_1024_seeded_workload(handle, oracle):
    """gemm N=1024, C = 1.5 A B + 1.2 C, A/B/C seeded U[-1,1) (seed 1, matrices
    0, 1, 2, generated on the device), plain FP32 method."""
    n = 1024
    A, B, C = (torch.empty((n, n), device=DEV) for _ in range(3))
    for i, x in enumerate((A, B, C)):
        handle.fill_uniform(x, 1, i)
    a, b, c = (oracle.fill_uniform(n, n, 1, i) for i in range(3))
    np.testing.assert_array_equal(C.cpu().numpy(), c)
    handle.gemm(A, B, C, alpha=1.5, beta=1.2, prec="fp32")
    torch.cuda.synchronize()
    want = oracle.gemm_f64(a.astype(np.float64), b.astype(np.float64), c.astype(np.float64), 1.5, 1.2)
    err = oracle.rel_frobenius(C.cpu().numpy(), want)
    print(f"gemm 1024 seeded fp32: rel Frobenius error {err:.3e}", flush=True)
    assert err < 2e-6


@pytest.mark.parametrize("rows,cols,seed,mid,mu,sigma", [(512, 384, 7, 0, 0.0, 2.0), (300, 257, 123, 5, 1.5, 0.5),
                                                         (1, 1000, 2**40 + 3, 2**31, -2.0, 3.0)])
def test_fill_lognormal_matches_oracle(handle, oracle, rows, cols, seed, mid, mu, sigma):
    x = torch.full((rows, cols + 3), float("nan"), device=DEV)[:, :cols]
    handle.fill_lognormal(x, seed, mid, mu, sigma)
    y = torch.empty((rows, cols), device=DEV)
    handle.fill_lognormal(y, seed, mid, mu, sigma)
    got = x.cpu().numpy()
    np.testing.assert_array_equal(got, y.cpu().numpy())  # deterministic, independent of the leading dimension
    want = oracle.fill_lognormal(rows, cols, seed, mid, mu, sigma)
    np.testing.assert_array_equal(np.signbit(got), np.signbit(want))
    np.testing.assert_allclose(got, want, rtol=2e-5, atol=0)
