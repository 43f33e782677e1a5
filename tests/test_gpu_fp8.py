"""Block-scaled FP8 (OCP e4m3) path on the GPU against the CPU oracle.

The quantiser must match the oracle bit for bit (codes and block exponents).
The contraction must equal, to FP32-accumulation accuracy, the FP64 product
of the oracle's dequantised operands -- which pins the kernel's use of the
packed codes and of the per-block E8M0 scales exactly; the error against the
unquantised inputs is reported, not targeted (lognormal inputs span more
decades than e4m3 holds)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda"


def lognormal(rng, shape, sigma=2.0):
    return (np.sign(rng.standard_normal(shape)) * np.exp(sigma * rng.standard_normal(shape))).astype(np.float32)


def dequant(oracle, q, e, block=128):
    vals = np.vectorize(oracle.e4m3_to_f32, otypes=[np.float64])(q)
    scale = np.repeat(np.repeat(2.0 ** e.astype(np.float64), block, 0), block, 1)[: q.shape[0], : q.shape[1]]
    return vals * scale


@pytest.mark.parametrize("rows,cols,block", [(128, 128, 128), (300, 200, 128), (64, 96, 32), (130, 70, 64)])
def test_quantize_bit_exact(handle, oracle, rows, cols, block):
    rng = np.random.default_rng(rows + cols)
    x = lognormal(rng, (rows, cols))
    x[0, 0] = 0.0
    q = torch.empty((rows, cols), dtype=torch.uint8, device=DEV)
    tr, tc = -(-rows // block), -(-cols // block)
    e = torch.empty((tr, tc), dtype=torch.int32, device=DEV)
    handle.quantize_fp8_block(torch.from_numpy(x).to(DEV), q, e, block=block)
    want_q, want_e = oracle.quantize_fp8_block(x, block)
    np.testing.assert_array_equal(e.cpu().numpy(), want_e)
    np.testing.assert_array_equal(q.cpu().numpy(), want_q)


def two_term(oracle, x):
    q1, e1 = oracle.quantize_fp8_block(x, 128)
    d1 = dequant(oracle, q1, e1).astype(np.float32)
    r = (x - d1).astype(np.float32)
    q2, e2 = oracle.quantize_fp8_block(r, 128)
    return dequant(oracle, q1, e1), dequant(oracle, q2, e2)


def check_dequantised_product(handle, oracle, prec, m, n, k, tb):
    rng = np.random.default_rng(m * 7 + n + k)
    A = lognormal(rng, (m, k), 1.0)
    B = lognormal(rng, (n, k) if tb else (k, n), 1.0)
    C = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    Bk = B if tb else B.T.copy()  # B-role operand, rows = n
    if prec == "fp8-block":
        qa, ea = oracle.quantize_fp8_block(A, 128)
        qb, eb = oracle.quantize_fp8_block(Bk, 128)
        want = dequant(oracle, qa, ea) @ dequant(oracle, qb, eb).T
    else:
        a1, a2 = two_term(oracle, A)
        b1, b2 = two_term(oracle, Bk)
        want = a1 @ b1.T + a1 @ b2.T + a2 @ b1.T
    want = 0.75 * want + 0.5 * C.astype(np.float64)
    Cd = torch.from_numpy(C).to(DEV)
    handle.gemm(torch.from_numpy(A).to(DEV), torch.from_numpy(B).to(DEV), Cd, alpha=0.75, beta=0.5, prec=prec,
                transB=tb)
    got = Cd.cpu().numpy().astype(np.float64)
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    # the scaled FP8 MFMA sums each 64-wide dot with ~2^-15 relative error (measured)
    assert err < 1e-4, err


@pytest.mark.parametrize("prec", ["fp8-block", "fp8-block2"])
@pytest.mark.parametrize("m,n,k,tb", [(256, 256, 256, False), (384, 200, 300, False), (128, 520, 640, True),
                                      (1024, 1024, 1024, False)])
def test_fp8_contraction_equals_dequantised_product(handle, oracle, prec, m, n, k, tb):
    check_dequantised_product(handle, oracle, prec, m, n, k, tb)


@pytest.mark.parametrize("prec", ["fp8-block", "fp8-block2"])
def test_fp8_every_tile_configuration(handle, oracle, mfx, prec):
    try:
        for config in range(mfx.config_count(prec)):
            handle.force_config(prec, config)
            check_dequantised_product(handle, oracle, prec, 520, 300, 384, config % 2 == 1)
    finally:
        handle.force_config(prec, -1)


def test_fp8_error_vs_fp64_is_reported_and_two_term_helps(handle):
    n = 2048
    A = torch.empty((n, n), device=DEV)
    B = torch.empty((n, n), device=DEV)
    handle.fill_lognormal(A, 5, 0, 0.0, 2.0)
    handle.fill_lognormal(B, 5, 1, 0.0, 2.0)
    R = A.double() @ B.double()
    errs = {}
    for prec in ("fp8-block", "fp8-block2", "split-bf16"):
        C = torch.empty((n, n), device=DEV)
        handle.gemm(A, B, C, prec=prec)
        errs[prec] = (torch.linalg.norm(C.double() - R) / torch.linalg.norm(R)).item()
    assert np.isfinite(errs["fp8-block"]) and errs["fp8-block"] < 1.0
    assert errs["fp8-block2"] < errs["fp8-block"]
    assert errs["split-bf16"] < errs["fp8-block2"]
