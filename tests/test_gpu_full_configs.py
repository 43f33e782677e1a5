"""The workloads at their full sizes, checked through size-independent
properties against an FP64 product on the GPU:

* 2mm D = 1.5 (A B) C + 1.2 D and 3mm G = (A B)(C D), 4096^2, seeded
  U[-1,1): relative Frobenius error per output (one bar per chained product,
  sqrt(stages) x the single-product bar), per-call and grouped launches;
* syrk C = 1.5 A A^T + 1.2 C and syr2k C = 1.5 (A B^T + B A^T) + 1.2 C,
  8192^2, C symmetric seeded, lower triangle only: the strict upper triangle
  is left bit-identical, and the lower triangle's error vs FP64;
* the batched workload (512 uneven problems + an 8-long chain of 2048^2, one
  grouped launch in a hipGraph): chain error vs FP64 and a sampled error of
  the independent problems;
* block-scaled FP8 at 16384^2 on sign x lognormal(0, 2) inputs: error of a
  2048 x 2048 block vs FP64, the two-term form below the one-term form.

Bars (relative Frobenius error vs FP64): FP32 and split BF16 1e-5 per
product, single BF16 2e-2, FP8 reported (one-term < 1, two-term below it)."""
import math
import os
import sys

import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda"
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BAR = {"fp32": 1e-5, "split-bf16": 1e-5, "bf16": 2e-2}
ALPHA, BETA = 1.5, 1.2


def seeded(handle, shape, matrix_id, seed=1):
    x = torch.empty(shape, device=DEV)
    handle.fill_uniform(x, seed, matrix_id)
    return x


@pytest.fixture(scope="module")
def mm_inputs(handle):
    n = 4096
    print(f"2mm / 3mm inputs n={n}", flush=True)
    return [seeded(handle, (n, n), i) for i in range(4)]


@pytest.fixture(scope="module")
def mm_reference(handle, mm_inputs):
    A, B, C, D = (x.double() for x in mm_inputs)
    n = A.shape[0]
    T = torch.empty((n, n), device=DEV, dtype=torch.float64)
    handle.gemm_f64(A, B, T, alpha=ALPHA)
    D2 = D.clone()
    handle.gemm_f64(T, C, D2, alpha=1.0, beta=BETA)
    E = torch.empty_like(T)
    F = torch.empty_like(T)
    G = torch.empty_like(T)
    handle.gemm_f64(A, B, E)
    handle.gemm_f64(C, D, F)
    handle.gemm_f64(E, F, G)
    torch.cuda.synchronize()
    return {"2mm": D2, "3mm": G}


@pytest.mark.parametrize("grouped", [False, True])
@pytest.mark.parametrize("prec", ["fp32", "split-bf16"])
def test_2mm_3mm_4096(handle, mfx, mm_inputs, mm_reference, prec, grouped):
    A, B, C, D = mm_inputs
    n = A.shape[0]
    T = torch.empty((n, n), device=DEV)
    D2 = D.clone()
    E, F, G = (torch.empty((n, n), device=DEV) for _ in range(3))
    print(f"2mm / 3mm n={n} {prec} grouped={grouped}", flush=True)
    if grouped:
        handle.grouped([mfx.problem(A, B, T, alpha=ALPHA), mfx.problem(T, C, D2, alpha=1.0, beta=BETA)],
                       deps=[(1, 0)], prec=prec)
        handle.grouped([mfx.problem(A, B, E), mfx.problem(C, D, F), mfx.problem(E, F, G)],
                       deps=[(2, 0), (2, 1)], prec=prec)
    else:
        handle.gemm(A, B, T, alpha=ALPHA, prec=prec)
        handle.gemm(T, C, D2, alpha=1.0, beta=BETA, prec=prec)
        handle.gemm(A, B, E, prec=prec)
        handle.gemm(C, D, F, prec=prec)
        handle.gemm(E, F, G, prec=prec)
    e2 = handle.rel_frobenius(D2, mm_reference["2mm"])
    e3 = handle.rel_frobenius(G, mm_reference["3mm"])
    print(f"  rel Frobenius error: 2mm {e2:.3e}, 3mm {e3:.3e}", flush=True)
    assert e2 < BAR[prec] * math.sqrt(2)
    assert e3 < BAR[prec] * math.sqrt(2)


@pytest.fixture(scope="module")
def sym_inputs(handle):
    n = 8192
    print(f"syrk / syr2k inputs n={n}", flush=True)
    A = seeded(handle, (n, n), 0)
    B = seeded(handle, (n, n), 1)
    X = seeded(handle, (n, n), 2)
    C0 = ((X + X.T) * 0.5).contiguous()
    del X
    return A, B, C0


@pytest.fixture(scope="module")
def sym_reference(handle, sym_inputs):
    A, B, C0 = sym_inputs
    A64, B64 = A.double(), B.double()
    R1 = C0.double()
    handle.gemm_f64(A64, A64, R1, alpha=ALPHA, beta=BETA, transB=True)
    R2 = C0.double()
    handle.gemm_f64(A64, B64, R2, alpha=ALPHA, beta=BETA, transB=True)
    handle.gemm_f64(B64, A64, R2, alpha=ALPHA, beta=1.0, transB=True)
    torch.cuda.synchronize()
    del A64, B64
    torch.cuda.empty_cache()
    return {"syrk": R1, "syr2k": R2}


@pytest.mark.parametrize("prec", ["fp32", "split-bf16", "bf16"])
@pytest.mark.parametrize("op", ["syrk", "syr2k"])
def test_syrk_syr2k_8192_lower_triangle(handle, mfx, sym_inputs, sym_reference, op, prec):
    A, B, C0 = sym_inputs
    print(f"{op} n={A.shape[0]} {prec} lower triangle", flush=True)
    C = C0.clone()
    if op == "syrk":
        handle.syrk(A, C, alpha=ALPHA, beta=BETA, prec=prec, uplo=mfx.LOWER, mirror=False)
    else:
        handle.syr2k(A, B, C, alpha=ALPHA, beta=BETA, prec=prec, uplo=mfx.LOWER, mirror=False)
    torch.cuda.synchronize()
    assert torch.equal(torch.triu(C, 1), torch.triu(C0, 1)), "strict upper triangle was written"
    err = handle.rel_frobenius(C, sym_reference[op], uplo=mfx.LOWER)
    print(f"  rel Frobenius error (lower triangle): {err:.3e}", flush=True)
    assert err < BAR[prec]
    del C
    torch.cuda.empty_cache()


def test_batched_uneven_and_chain_in_one_graph(handle, mfx):
    sys.path.insert(0, ROOT)
    import bench
    print("batched: 512 uneven problems + 8-long chain of 2048^2, split-BF16, hipGraph", flush=True)
    g, _ = bench.grouped_workload(torch, mfx, handle, 3, 1)
    print(f"  {g['ms']} ms, {g['tflops']} TFLOP/s, chain error {g['chain_rel_err']:.3e}, "
          f"sampled error max {g['sampled_rel_err_max']:.3e}", flush=True)
    # three chained levels
    assert g["chain_rel_err"] < BAR["split-bf16"] * math.sqrt(3)
    assert g["sampled_rel_err_max"] < BAR["split-bf16"]
    torch.cuda.empty_cache()


def test_fp8_16384_sampled_block(handle):
    n, blk = 16384, 2048
    print(f"fp8 n={n}, sign x lognormal(0, 2), error on a {blk}^2 block", flush=True)
    A = torch.empty((n, n), device=DEV)
    B = torch.empty((n, n), device=DEV)
    handle.fill_lognormal(A, 7, 0, 0.0, 2.0)
    handle.fill_lognormal(B, 7, 1, 0.0, 2.0)
    r0, c0 = 5 * 1024, 11 * 1024
    R = A[r0:r0 + blk].double() @ B[:, c0:c0 + blk].double()
    errs = {}
    for prec in ("fp8-block", "fp8-block2"):
        C = torch.empty((n, n), device=DEV)
        handle.gemm(A, B, C, prec=prec)
        torch.cuda.synchronize()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        handle.gemm(A, B, C, prec=prec)
        stop.record()
        torch.cuda.synchronize()
        ms = start.elapsed_time(stop)
        Cb = C[r0:r0 + blk, c0:c0 + blk].double()
        errs[prec] = (torch.linalg.norm(Cb - R) / torch.linalg.norm(R)).item()
        print(f"  {prec}: {ms:.3f} ms ({2.0 * n ** 3 / ms / 1e9:.0f} TFLOP/s), block error {errs[prec]:.3e}",
              flush=True)
        del C, Cb
        torch.cuda.empty_cache()
    assert math.isfinite(errs["fp8-block"]) and errs["fp8-block"] < 1.0
    assert errs["fp8-block2"] < errs["fp8-block"]
