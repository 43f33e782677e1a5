"""Parity of the HIP dense contractions with the CPU oracle (GPU, through the C ABI).

Bars: bit-exact for the integer/bit work (input generation, BF16 split);
the reference's own verify rule (percentDiff <= 0.05 %) on the reference's
inputs; relative Frobenius error against FP64 on seeded inputs with the
tolerance written in each test.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = "cuda"

# Tolerances of the relative Frobenius error vs FP64 (uniform [-1, 1) inputs).
TOL = {"fp32": 2e-6, "split-bf16": 1e-5, "bf16": 1e-2}


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def empty_view(rows, cols, pad=0, dtype=torch.float32, fill=float("nan")):
    """rows x cols view with leading dimension cols + pad (pad > 0 exercises ld != cols)."""
    base = torch.full((rows, cols + pad), fill, dtype=dtype, device=DEV)
    return base[:, :cols]


def fp64_ref(oracle, A, B, C, alpha, beta, ta=False, tb=False):
    return oracle.gemm_f64(A.astype(np.float64), B.astype(np.float64),
                           None if C is None else C.astype(np.float64), alpha, beta, ta=ta, tb=tb)


def test_fill_uniform_bit_exact(handle, oracle):
    for rows, cols, seed, mid in ((37, 53, 1, 0), (256, 128, 12345, 7), (1, 1000, 2**63 + 5, 2**31)):
        x = empty_view(rows, cols, pad=3)
        handle.fill_uniform(x, seed, mid)
        want = oracle.fill_uniform(rows, cols, seed, mid)
        np.testing.assert_array_equal(x.cpu().numpy(), want)


def test_split_bf16_bit_exact(handle, oracle):
    rng = np.random.default_rng(11)
    x = np.concatenate([rng.uniform(-1, 1, 4096), rng.standard_normal(4096) * 1e20,
                        [0.0, -0.0, 1 + 2**-8, 1 + 3 * 2**-8, 3.3e38, 1e-40]]).astype(np.float32)
    xd = to_dev(x)
    hi = torch.empty(x.size, dtype=torch.int16, device=DEV)
    lo = torch.empty(x.size, dtype=torch.int16, device=DEV)
    handle.split_bf16(xd, hi, lo)
    h, l = oracle.split_bf16(x)
    np.testing.assert_array_equal(hi.cpu().numpy().view(np.uint16), h)
    np.testing.assert_array_equal(lo.cpu().numpy().view(np.uint16), l)


@pytest.mark.parametrize("m,n,k,ta,tb", [(67, 45, 33, False, False), (64, 64, 64, True, False),
                                         (50, 70, 90, False, True), (33, 17, 129, True, True)])
def test_gemm_f64_matches_oracle(handle, oracle, m, n, k, ta, tb):
    rng = np.random.default_rng(m + n + k)
    A = rng.standard_normal((k, m) if ta else (m, k))
    B = rng.standard_normal((n, k) if tb else (k, n))
    C = rng.standard_normal((m, n))
    Cd = to_dev(C)
    handle.gemm_f64(to_dev(A), to_dev(B), Cd, alpha=0.5, beta=2.0, transA=ta, transB=tb)
    want = oracle.gemm_f64(A, B, C, alpha=0.5, beta=2.0, ta=ta, tb=tb)
    np.testing.assert_allclose(Cd.cpu().numpy(), want, rtol=1e-12, atol=1e-12)


def test_rel_frobenius_and_percent_diff(handle, oracle):
    rng = np.random.default_rng(3)
    R = rng.standard_normal((300, 200))
    X = (R * (1 + 1e-4 * rng.standard_normal(R.shape))).astype(np.float32)
    X[5, 7] = R[5, 7] * 1.01  # one element 1 % off: fails the 0.05 % rule
    got = handle.rel_frobenius(to_dev(X), to_dev(R))
    assert got == pytest.approx(oracle.rel_frobenius(X, R), rel=1e-9)
    cnt = handle.percent_diff_count(to_dev(X), to_dev(R), 0.05)
    assert cnt == oracle.percent_diff_count(R.astype(np.float32), X, 0.05)
    assert cnt >= 1
    # lower-triangle variant covers r >= c only
    low = handle.rel_frobenius(to_dev(X), to_dev(R), uplo=0)
    mask = np.tril(np.ones(R.shape, bool))
    assert low == pytest.approx(np.linalg.norm((X - R)[mask]) / np.linalg.norm(R[mask]), rel=1e-9)


@pytest.mark.parametrize("prec", ["fp32", "split-bf16", "bf16"])
@pytest.mark.parametrize("n", [64, 100, 256, 1024])
def test_gemm_reference_inputs_pass_reference_verify(handle, oracle, prec, n):
    """polybench/gemm.cpp with its own inputs and constants; verify() rule vs the restated CPU loop."""
    A, B, C, want = oracle.gemm_reference(n)
    Cd = to_dev(C)
    handle.gemm(to_dev(A), to_dev(B), Cd, alpha=32412.0, beta=2123.0, prec=prec)
    got = Cd.cpu().numpy()
    if prec == "bf16":
        # single BF16 keeps 8 significand bits: it cannot meet the 0.05 % rule
        assert oracle.rel_frobenius(got, oracle.closed_form("gemm", n)) < 1e-2
        return
    assert oracle.percent_diff_count(want, got) == 0
    assert oracle.percent_diff_count(oracle.closed_form("gemm", n).astype(np.float32), got) == 0


SHAPES = [(128, 128, 128), (77, 130, 51), (256, 256, 1000), (1000, 300, 260), (33, 1030, 64), (512, 512, 4),
          (1, 1, 1), (300, 257, 255)]


@pytest.mark.parametrize("prec", ["fp32", "split-bf16", "bf16"])
@pytest.mark.parametrize("m,n,k", SHAPES)
@pytest.mark.parametrize("ta,tb", [(False, False), (False, True), (True, False), (True, True)])
def test_gemm_seeded_vs_fp64(handle, oracle, prec, m, n, k, ta, tb):
    A = oracle.fill_uniform(k, m, 1, 0) if ta else oracle.fill_uniform(m, k, 1, 0)
    B = oracle.fill_uniform(n, k, 1, 1) if tb else oracle.fill_uniform(k, n, 1, 1)
    C = oracle.fill_uniform(m, n, 1, 2)
    Ad = empty_view(*A.shape, pad=4)
    Ad.copy_(to_dev(A))
    Bd = empty_view(*B.shape, pad=8)
    Bd.copy_(to_dev(B))
    Cd = empty_view(m, n, pad=4)
    Cd.copy_(to_dev(C))
    handle.gemm(Ad, Bd, Cd, alpha=1.5, beta=1.2, prec=prec, transA=ta, transB=tb)
    want = fp64_ref(oracle, A, B, C, 1.5, 1.2, ta, tb)
    err = oracle.rel_frobenius(Cd.cpu().numpy(), want)
    tol = TOL[prec] if k >= 16 else TOL[prec] * 4
    assert err < tol, err
    # padding columns of C untouched
    assert torch.isnan(Cd.as_strided((m, 4), (Cd.stride(0), 1), Cd.storage_offset() + n)).all()


@pytest.mark.parametrize("prec", ["fp32", "split-bf16", "bf16"])
def test_gemm_beta_zero_does_not_read_c(handle, oracle, prec):
    m, n, k = 200, 136, 72
    A, B = oracle.fill_uniform(m, k, 2, 0), oracle.fill_uniform(k, n, 2, 1)
    Cd = empty_view(m, n)  # NaN-filled
    handle.gemm(to_dev(A), to_dev(B), Cd, alpha=1.0, beta=0.0, prec=prec)
    got = Cd.cpu().numpy()
    assert np.isfinite(got).all()
    assert oracle.rel_frobenius(got, fp64_ref(oracle, A, B, None, 1.0, 0.0)) < TOL[prec]


@pytest.mark.parametrize("prec", ["fp32", "split-bf16", "bf16"])
def test_gemm_k_zero_scales_c(handle, oracle, prec):
    m, n = 70, 90
    C = oracle.fill_uniform(m, n, 4, 2)
    Cd = to_dev(C)
    A = torch.zeros((m, 0), device=DEV)
    B = torch.zeros((0, n), device=DEV)
    handle.gemm(A, B, Cd, alpha=2.0, beta=0.5, prec=prec)
    np.testing.assert_array_equal(Cd.cpu().numpy(), (C * np.float32(0.5)).astype(np.float32))


@pytest.mark.parametrize("prec", ["fp32", "split-bf16", "bf16"])
@pytest.mark.parametrize("config", [0, 1, 2, 3])
def test_every_tile_config(handle, oracle, prec, config):
    m, n, k = 300, 520, 200
    A, B, C = oracle.fill_uniform(m, k, 5, 0), oracle.fill_uniform(k, n, 5, 1), oracle.fill_uniform(m, n, 5, 2)
    Cd = to_dev(C)
    handle.force_config(prec, config)
    try:
        handle.gemm(to_dev(A), to_dev(B), Cd, alpha=-0.5, beta=0.25, prec=prec)
    finally:
        handle.force_config(prec, -1)
    assert oracle.rel_frobenius(Cd.cpu().numpy(), fp64_ref(oracle, A, B, C, -0.5, 0.25)) < TOL[prec]


def test_split_bf16_error_at_large_k(handle):
    """N = 2048 uniform: split BF16 within 1e-5 of FP64 (scored on the GPU FP64 oracle)."""
    n = 2048
    A = torch.empty((n, n), device=DEV)
    B = torch.empty((n, n), device=DEV)
    handle.fill_uniform(A, 9, 0)
    handle.fill_uniform(B, 9, 1)
    A64, B64 = A.double(), B.double()
    R = torch.empty((n, n), dtype=torch.float64, device=DEV)
    handle.gemm_f64(A64, B64, R)
    errs = {}
    for prec in ("fp32", "split-bf16", "bf16"):
        C = torch.empty((n, n), device=DEV)
        handle.gemm(A, B, C, prec=prec)
        errs[prec] = handle.rel_frobenius(C, R)
    assert errs["fp32"] < 2e-6
    assert errs["split-bf16"] < 1e-5
    assert 1e-4 < errs["bf16"] < 1e-2
    # the GPU FP64 oracle agrees with torch's FP64 matmul
    assert torch.linalg.norm(R - A64 @ B64) / torch.linalg.norm(R) < 1e-13


@pytest.mark.parametrize("prec", ["fp32", "split-bf16"])
@pytest.mark.parametrize("n", [64, 256, 300])
def test_syrk_reference_inputs_full_square(handle, oracle, prec, n):
    """polybench/syrk.cpp: lower-triangle schedule + mirror reproduces the full square."""
    A, C, want = oracle.syrk_reference(n)
    Cd = to_dev(C)
    handle.syrk(to_dev(A), Cd, alpha=123.0, beta=14512.0, prec=prec, uplo=0, mirror=True)
    got = Cd.cpu().numpy()
    assert oracle.percent_diff_count(want, got) == 0
    np.testing.assert_array_equal(got, got.T)


@pytest.mark.parametrize("prec", ["fp32", "split-bf16"])
@pytest.mark.parametrize("n", [64, 256, 300])
def test_syr2k_reference_inputs_full_square(handle, oracle, prec, n):
    A, B, C, want = oracle.syr2k_reference(n)
    Cd = to_dev(C)
    handle.syr2k(to_dev(A), to_dev(B), Cd, alpha=1.0, beta=1.0, prec=prec, uplo=0, mirror=True)
    got = Cd.cpu().numpy()
    assert oracle.percent_diff_count(want, got) == 0


@pytest.mark.parametrize("prec", ["fp32", "split-bf16", "bf16"])
@pytest.mark.parametrize("uplo", [0, 1, 2])
@pytest.mark.parametrize("n,k", [(200, 72), (513, 300)])
def test_syrk_syr2k_seeded_triangles(handle, oracle, prec, uplo, n, k):
    A, B = oracle.fill_uniform(n, k, 6, 0), oracle.fill_uniform(n, k, 6, 1)
    C0 = oracle.fill_uniform(n, n, 6, 2)
    C0 = np.tril(C0) + np.tril(C0, -1).T  # symmetric C
    want1 = fp64_ref(oracle, A, A, C0, 0.75, 1.25, tb=True)
    want2 = fp64_ref(oracle, B, A, fp64_ref(oracle, A, B, C0, 0.75, 1.25, tb=True), 0.75, 1.0, tb=True)
    mask = {0: np.tril(np.ones((n, n), bool)), 1: np.triu(np.ones((n, n), bool)), 2: np.ones((n, n), bool)}[uplo]
    for which, want in (("syrk", want1), ("syr2k", want2)):
        Cd = to_dev(C0)
        if which == "syrk":
            handle.syrk(to_dev(A), Cd, alpha=0.75, beta=1.25, prec=prec, uplo=uplo, mirror=False)
        else:
            handle.syr2k(to_dev(A), to_dev(B), Cd, alpha=0.75, beta=1.25, prec=prec, uplo=uplo, mirror=False)
        got = Cd.cpu().numpy()
        np.testing.assert_array_equal(got[~mask], C0[~mask])  # other triangle untouched
        err = np.linalg.norm((got - want)[mask]) / np.linalg.norm(want[mask])
        assert err < TOL[prec], (which, err)
