"""Pins the CPU oracle before anything is compared against it (CPU only).

Every oracle function is checked against something it does not itself
compute: closed forms against FP64 brute force on exact rational inputs, the
FP64 brute force against numpy's matmul, the restated reference loops against
the closed forms through the reference's own percentDiff rule, the number
formats against exhaustive nearest-value search and torch's converters, the
generator against the published SplitMix64 sequence.
"""
import math

import numpy as np
import pytest
import torch


def exact_inputs(name, n):
    """The reference inputs computed exactly in float64 (no float rounding)."""
    i = np.arange(n, dtype=np.float64)[:, None]
    j = np.arange(n, dtype=np.float64)[None, :]
    if name == "gemm":  # polybench/gemm.cpp:18-40
        return dict(A=i * j / n, B=(i * j + 1) / n, C=(i * j + 2) / n)
    if name == "2mm":  # polybench/2mm.cpp:16-45
        return dict(A=i * j / n, B=i * (j + 1) / n, C=i * (j + 3) / n, D=i * (j + 2) / n)
    if name == "3mm":  # polybench/3mm.cpp:17-47
        return dict(A=i * j / n, B=i * (j + 1) / n, C=i * (j + 3) / n, D=i * (j + 2) / n)
    if name == "syrk":  # polybench/syrk.cpp:18-31
        return dict(A=i * j / n, C=(i * j + 2) / n)
    if name == "syr2k":  # polybench/syr2k.cpp:18-32
        return dict(A=i * j / n, B=(i * j + 1) / n, C=(i * j + 2) / n)
    raise KeyError(name)


def brute(name, n, oracle):
    x = exact_inputs(name, n)
    g = oracle.gemm_f64
    if name == "gemm":
        return g(x["A"], x["B"], x["C"], alpha=32412.0, beta=2123.0)
    if name == "2mm":
        c = g(x["A"], x["B"], x["C"], beta=1.0)
        return g(c, x["D"])
    if name == "3mm":
        return g(g(x["A"], x["B"]), g(x["C"], x["D"]))
    if name == "syrk":
        return g(x["A"], x["A"], x["C"], alpha=123.0, beta=14512.0, tb=True)
    if name == "syr2k":
        t = g(x["A"], x["B"], x["C"], beta=1.0, tb=True)
        return g(x["B"], x["A"], t, beta=1.0, tb=True)


@pytest.mark.parametrize("m,n,k,ta,tb", [(5, 7, 3, False, False), (17, 9, 33, True, False), (8, 31, 12, False, True),
                                         (19, 4, 25, True, True), (1, 1, 1, False, False), (6, 3, 0, False, False)])
def test_gemm_f64_matches_numpy(oracle, m, n, k, ta, tb):
    rng = np.random.default_rng(m * 1000 + n * 10 + k)
    A = rng.standard_normal((k, m) if ta else (m, k))
    B = rng.standard_normal((n, k) if tb else (k, n))
    C = rng.standard_normal((m, n))
    got = oracle.gemm_f64(A, B, C, alpha=1.5, beta=-0.75, ta=ta, tb=tb)
    want = 1.5 * ((A.T if ta else A) @ (B.T if tb else B)) - 0.75 * C
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("name", ["gemm", "2mm", "3mm", "syrk", "syr2k"])
@pytest.mark.parametrize("n", [7, 16, 33, 64])
def test_closed_forms_match_brute_force(oracle, name, n):
    got = oracle.closed_form(name, n)
    want = brute(name, n, oracle)
    scale = np.abs(want).max()
    assert np.abs(got - want).max() <= 1e-13 * scale


@pytest.mark.parametrize("n", [16, 64, 96])
def test_restated_reference_loops_pass_reference_verify(oracle, n):
    """Each restated serial loop must pass the reference's own verify rule
    (percentDiff > 0.05 fails) against the exact closed form."""
    _, _, _, gemm = oracle.gemm_reference(n)
    assert oracle.percent_diff_count(oracle.closed_form("gemm", n), gemm) == 0
    *_, e = oracle.mm2_reference(n)
    assert oracle.percent_diff_count(oracle.closed_form("2mm", n), e) == 0
    *_, g = oracle.mm3_reference(n)
    assert oracle.percent_diff_count(oracle.closed_form("3mm", n), g) == 0
    _, _, s = oracle.syrk_reference(n)
    assert oracle.percent_diff_count(oracle.closed_form("syrk", n), s) == 0
    _, _, _, s2 = oracle.syr2k_reference(n)
    assert oracle.percent_diff_count(oracle.closed_form("syr2k", n), s2) == 0


def test_restated_loops_exact_for_power_of_two(oracle):
    """At N a power of two every reference input is exact in FP32, so the
    restated init must reproduce the exact rationals bit for bit."""
    n = 32
    A, B, C = oracle.gemm_inputs(n)
    x = exact_inputs("gemm", n)
    for got, key in ((A, "A"), (B, "B"), (C, "C")):
        np.testing.assert_array_equal(got.astype(np.float64), x[key])


def test_planted_mistakes_fail_the_check(oracle):
    """A dropped beta term must fail percentDiff on the reference inputs.  The
    reference's A, B and C are symmetric (i*j based), so a transposed operand
    is invisible to them: that mistake is only caught by the seeded inputs."""
    n = 48
    x = exact_inputs("gemm", n)
    closed = oracle.closed_form("gemm", n)
    no_beta = oracle.gemm_f64(x["A"], x["B"], alpha=32412.0).astype(np.float32)
    assert oracle.percent_diff_count(closed, no_beta) > 0
    transposed = oracle.gemm_f64(x["A"], x["B"], x["C"], alpha=32412.0, beta=2123.0, tb=True).astype(np.float32)
    assert oracle.percent_diff_count(closed, transposed) == 0  # symmetric inputs hide it
    A = oracle.fill_uniform(n, n, 3, 0).astype(np.float64)
    B = oracle.fill_uniform(n, n, 3, 1).astype(np.float64)
    good = oracle.gemm_f64(A, B)
    assert oracle.rel_frobenius(oracle.gemm_f64(A, B, tb=True), good) > 0.5
    assert oracle.rel_frobenius(oracle.gemm_f64(A, B, ta=True), good) > 0.5


def test_percent_diff_known_values(oracle):
    pd = oracle.percent_diff
    assert pd(1.0, 1.0) == 0.0
    assert pd(0.005, 0.009) == 0.0  # both below 0.01: pass by definition
    assert pd(-0.005, 0.0099) == 0.0
    assert pd(200.0, 201.0) == pytest.approx(0.5, rel=1e-6)
    assert pd(100.0, 50.0) == pytest.approx(50.0, rel=1e-6)  # denominator is val1
    assert pd(50.0, 100.0) == pytest.approx(100.0, rel=1e-6)
    assert pd(-4.0, 4.0) == pytest.approx(200.0, rel=1e-6)
    assert pd(0.0, 0.02) > 1e5  # val1 ~ 0: denominator is the 1e-8 guard


def test_matmulchain_int_accumulator(oracle):
    n = 24
    eye = oracle.matmulchain_identity(n)
    np.testing.assert_array_equal(eye, np.eye(n, dtype=np.float32))
    p = oracle.matmulchain_multiply_int(eye, eye)
    q = oracle.matmulchain_multiply_int(eye, eye)
    np.testing.assert_array_equal(oracle.matmulchain_multiply_int(p, q), np.eye(n, dtype=np.float32))
    # `auto sum = 0` truncates toward zero after every step: 0.75 + 0.75 -> 0 + 0
    a = np.array([[0.75, 0.75]], np.float32).reshape(1, 2)
    b = np.array([[1.0], [1.0]], np.float32)
    a2 = np.zeros((2, 2), np.float32)
    a2[0] = a
    b2 = np.zeros((2, 2), np.float32)
    b2[:, 0] = b[:, 0]
    assert oracle.matmulchain_multiply_int(a2, b2)[0, 0] == 0.0
    a2[0] = [1.5, -1.25]
    assert oracle.matmulchain_multiply_int(a2, b2)[0, 0] == 0.0  # int(1.5)=1, int(1 - 1.25)=0


def test_splitmix64_published_sequence(oracle):
    """SplitMix64 seeded with 0 (Steele, Lea & Flood 2014) starts with these outputs."""
    golden = 0x9E3779B97F4A7C15
    want = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    got = [oracle.splitmix64_mix((golden * (i + 1)) & 0xFFFFFFFFFFFFFFFF) for i in range(3)]
    assert got == want


def test_uniform_grid_and_determinism(oracle):
    a = oracle.fill_uniform(64, 48, seed=7, matrix=3)
    b = oracle.fill_uniform(64, 48, seed=7, matrix=3)
    np.testing.assert_array_equal(a, b)
    assert a.min() >= -1.0 and a.max() < 1.0
    grid = (a.astype(np.float64) + 1.0) * 2.0 ** 23
    np.testing.assert_array_equal(grid, np.round(grid))
    # keyed: a different matrix id or seed gives different values
    assert not np.array_equal(a, oracle.fill_uniform(64, 48, seed=7, matrix=4))
    assert not np.array_equal(a, oracle.fill_uniform(64, 48, seed=8, matrix=3))
    # the value of element (r, c) depends only on its linear index r*cols + c
    h = oracle.keyed_hash(7, 3, 5 * 48 + 11)
    assert a[5, 11] == np.float32((h >> 40) * 2.0 ** -23 - 1.0)
    # mean and variance of U[-1, 1)
    big = oracle.fill_uniform(512, 512, seed=1, matrix=0)
    assert abs(big.mean()) < 0.01 and abs(big.var() - 1.0 / 3.0) < 0.01


def _bf16_torch(x):
    return torch.from_numpy(np.asarray(x, np.float32)).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)


def test_bf16_rne_matches_torch_and_ties(oracle):
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32),
                        (rng.standard_normal(2000) * 1e-30).astype(np.float32),
                        (rng.standard_normal(2000) * 1e30).astype(np.float32),
                        np.array([0.0, -0.0, 1.0, -1.0, 3.4e38, -3.4e38, 1e-45], np.float32)])
    # exact ties: 1 + 2^-8 (even -> 1.0) and 1 + 3*2^-8 (odd -> up)
    x = np.concatenate([x, np.array([1 + 2 ** -8, 1 + 3 * 2 ** -8, -(1 + 2 ** -8)], np.float32)])
    hi, _ = oracle.split_bf16(x)
    np.testing.assert_array_equal(hi, _bf16_torch(x))
    assert hi[-3] == 0x3F80 and hi[-2] == 0x3F82 and hi[-1] == 0xBF80


def test_split_bf16_residual_bound(oracle):
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, 50000).astype(np.float32)
    hi, lo = oracle.split_bf16(x)
    h = oracle.bf16_to_f32(hi).astype(np.float64)
    l = oracle.bf16_to_f32(lo).astype(np.float64)
    xd = x.astype(np.float64)
    assert np.all(np.abs(xd - h) <= 2.0 ** -8 * np.abs(xd))
    assert np.all(np.abs(xd - h - l) <= 2.0 ** -16 * np.abs(xd))
    # lo is the RNE rounding of the exact residual
    np.testing.assert_array_equal(lo, _bf16_torch((xd - h).astype(np.float32)))


def _e4m3_table(oracle):
    codes = np.arange(0x7F, dtype=np.uint8)  # positive finite codes 0..126
    return codes, np.array([oracle.e4m3_to_f32(c) for c in codes])


def test_e4m3_decode_spec_values(oracle):
    f = oracle.e4m3_to_f32
    assert f(0x7E) == 448.0 and f(0xFE) == -448.0
    assert f(0x08) == 2.0 ** -6 and f(0x01) == 2.0 ** -9 and f(0x07) == 7 * 2.0 ** -9
    assert f(0x38) == 1.0 and f(0x3C) == 1.5 and f(0x40) == 2.0 and f(0x00) == 0.0
    assert math.isnan(f(0x7F)) and math.isnan(f(0xFF))
    _, vals = _e4m3_table(oracle)
    assert np.all(np.diff(vals) > 0)  # codes are monotone in value


def test_e4m3_encode_is_nearest_even(oracle):
    codes, vals = _e4m3_table(oracle)
    rng = np.random.default_rng(2)
    xs = np.concatenate([rng.uniform(0, 448, 3000), np.exp(rng.uniform(-14, 6, 3000)),
                         (vals[:-1] + vals[1:]) / 2, vals]).astype(np.float32)
    for x in xs:
        got = oracle.e4m3_rne_sat(float(x))
        d = np.abs(vals - np.float64(x))
        best = np.flatnonzero(d == d.min())
        want = best[0] if len(best) == 1 else [c for c in best if c % 2 == 0][0]
        assert got == want, (x, got, want)
        assert oracle.e4m3_rne_sat(-float(x)) == (want | 0x80)
    assert oracle.e4m3_rne_sat(449.0) == 0x7E and oracle.e4m3_rne_sat(1e30) == 0x7E
    assert oracle.e4m3_rne_sat(-1e30) == 0xFE
    assert oracle.e4m3_rne_sat(float("nan")) & 0x7F == 0x7F


def test_e4m3_matches_torch_in_range(oracle):
    rng = np.random.default_rng(3)
    x = (rng.standard_normal(5000) * 50).astype(np.float32)
    x = x[np.abs(x) <= 448]
    want = torch.from_numpy(x).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    got = np.array([oracle.e4m3_rne_sat(float(v)) for v in x], np.uint8)
    np.testing.assert_array_equal(got, want)


def test_fp8_block_exponent(oracle):
    assert oracle.fp8_block_exponent(0.0) == 0
    rng = np.random.default_rng(4)
    for amax in np.exp(rng.uniform(-60, 60, 2000)).astype(np.float32):
        e = oracle.fp8_block_exponent(float(amax))
        assert amax <= 448.0 * 2.0 ** e and amax > 448.0 * 2.0 ** (e - 1)
    assert oracle.fp8_block_exponent(448.0) == 0 and oracle.fp8_block_exponent(448.0001) == 1


def test_quantize_fp8_block_roundtrip(oracle):
    rng = np.random.default_rng(5)
    x = (np.sign(rng.standard_normal((200, 300))) * np.exp(rng.standard_normal((200, 300)))).astype(np.float32)
    q, e = oracle.quantize_fp8_block(x, 128)
    assert e.shape == (2, 3)
    dq = np.vectorize(oracle.e4m3_to_f32)(q) * np.repeat(np.repeat(2.0 ** e, 128, 0), 128, 1)[:200, :300]
    # block maxima land in [224, 448] after scaling; normals keep 3 mantissa bits
    rel = np.abs(dq - x) / np.abs(x)
    scaled = np.abs(x) / np.repeat(np.repeat(2.0 ** e, 128, 0), 128, 1)[:200, :300]
    assert np.all(rel[scaled >= 2.0 ** -6] <= 2.0 ** -4 + 1e-12)
    for bi in range(2):
        for bj in range(3):
            blk = np.abs(x[bi * 128:(bi + 1) * 128, bj * 128:(bj + 1) * 128]).max()
            assert 224.0 < blk / 2.0 ** e[bi, bj] <= 448.0


def test_cpu_baseline_computes_reference_gemm(oracle):
    n = 80
    A, B, C, want = oracle.gemm_reference(n)
    got = C.copy()
    oracle.cpu_baseline().cpu_gemm_rows(A.ravel(), B.ravel(), got.ravel(), n, 32412.0, 2123.0, 0, n)
    assert oracle.percent_diff_count(want, got) == 0
    A2, C2, want2 = oracle.syrk_reference(n)
    got2 = C2.copy()
    oracle.cpu_baseline().cpu_syrk_rows(A2.ravel(), got2.ravel(), n, 123.0, 14512.0, 0, n)
    assert oracle.percent_diff_count(want2, got2) == 0


def test_cpu_baseline_syr2k_rows(oracle):
    """cpu_syr2k_rows follows polybench/syr2k.cpp:78-87: the reference's own
    inputs pass its verify rule against the restated serial loop, and seeded
    non-symmetric inputs with alpha != 1, beta != 1 match the FP64 form
    alpha (A B^T + B A^T) + beta C on a partial row range (a dropped term,
    a transposed operand or a missing alpha / beta fails)."""
    base = oracle.cpu_baseline()
    n = 80
    A, B, C, want = oracle.syr2k_reference(n)
    got = C.copy()
    base.cpu_syr2k_rows(A.ravel(), B.ravel(), got.ravel(), n, 1.0, 1.0, 0, n)
    assert oracle.percent_diff_count(want, got) == 0
    n = 96
    A, B, C = (oracle.fill_uniform(n, n, 5, i) for i in range(3))
    got = C.copy()
    r0, r1 = 17, 70
    base.cpu_syr2k_rows(A.ravel(), B.ravel(), got.ravel(), n, 1.5, 1.2, r0, r1)
    a, b, c = A.astype(np.float64), B.astype(np.float64), C.astype(np.float64)
    ref = 1.5 * (a @ b.T + b @ a.T) + 1.2 * c
    assert oracle.rel_frobenius(got[r0:r1], ref[r0:r1]) < 1e-6
    np.testing.assert_array_equal(got[:r0], C[:r0])  # rows outside the range untouched
    np.testing.assert_array_equal(got[r1:], C[r1:])


def test_cpu_baseline_matmul_rows(oracle):
    """cpu_matmul_rows is runtime/matmulchain.cpp:20-28's c = a b (with a
    float accumulator): exact on the reference's identity chain, and equal
    to the FP64 product of seeded inputs on a partial row range."""
    base = oracle.cpu_baseline()
    n = 64
    eye = oracle.matmulchain_identity(n)
    got = np.full((n, n), 7.0, np.float32)
    base.cpu_matmul_rows(eye.ravel(), eye.ravel(), got.ravel(), n, 0, n)
    np.testing.assert_array_equal(got, oracle.matmulchain_multiply_int(eye, eye))
    n = 96
    A, B = oracle.fill_uniform(n, n, 9, 0), oracle.fill_uniform(n, n, 9, 1)
    got = np.full((n, n), np.nan, np.float32)
    base.cpu_matmul_rows(A.ravel(), B.ravel(), got.ravel(), n, 8, 40)
    ref = A.astype(np.float64) @ B.astype(np.float64)
    assert oracle.rel_frobenius(got[8:40], ref[8:40]) < 1e-6
    assert np.isnan(got[:8]).all() and np.isnan(got[40:]).all()


@pytest.mark.parametrize("m,n,k,beta", [(70, 45, 33, 1.2), (17, 130, 100, 0.0), (64, 16, 1, 2.0)])
def test_cpu_baseline_gemm_mnk_rows(oracle, m, n, k, beta):
    """cpu_gemm_mnk_rows: polybench/gemm.cpp:82-91's body on an m x n x k
    problem; beta == 0 writes C without reading it (a NaN-filled C stays
    out of the result)."""
    base = oracle.cpu_baseline()
    A, B = oracle.fill_uniform(m, k, 4, 0), oracle.fill_uniform(k, n, 4, 1)
    C = oracle.fill_uniform(m, n, 4, 2) if beta else np.full((m, n), np.nan, np.float32)
    got = C.copy()
    r0, r1 = 3, m - 2
    base.cpu_gemm_mnk_rows(A.ravel(), B.ravel(), got.ravel(), m, n, k, 1.5, beta, r0, r1)
    ref = 1.5 * (A.astype(np.float64) @ B.astype(np.float64)) + (beta * C.astype(np.float64) if beta else 0.0)
    assert oracle.rel_frobenius(got[r0:r1], ref[r0:r1]) < 1e-6
    np.testing.assert_array_equal(got[:r0], C[:r0])
    np.testing.assert_array_equal(got[r1:], C[r1:])


def test_lognormal_fill_distribution(oracle):
    """The FP8 workload's sign * lognormal(mu, sigma) inputs: deterministic
    per (seed, matrix), distinct across them, signs balanced, and ln|x|
    normal with the requested mean and deviation (bounds of about five
    standard errors for 2^18 samples; a wrong sigma, mu or sign rule fails)."""
    rows, cols = 512, 512
    for mu, sigma in ((0.0, 2.0), (1.5, 0.5)):
        x = oracle.fill_lognormal(rows, cols, 7, 0, mu, sigma).astype(np.float64)
        np.testing.assert_array_equal(x, oracle.fill_lognormal(rows, cols, 7, 0, mu, sigma))
        assert not np.array_equal(x, oracle.fill_lognormal(rows, cols, 7, 1, mu, sigma))
        assert not np.array_equal(x, oracle.fill_lognormal(rows, cols, 8, 0, mu, sigma))
        n = x.size
        assert abs((x < 0).mean() - 0.5) < 5 * 0.5 / np.sqrt(n)
        z = np.log(np.abs(x))
        assert abs(z.mean() - mu) < 5 * sigma / np.sqrt(n)
        assert abs(z.std() - sigma) < 5 * sigma / np.sqrt(2 * n)
        # Box-Muller normals: about 68.3 % within one deviation, 95.4 % within two
        t = np.abs(z - mu) / sigma
        assert abs((t < 1).mean() - 0.6827) < 0.005 and abs((t < 2).mean() - 0.9545) < 0.003


# ---- covariance / correlation (polybench/covariance.cpp, polybench/correlation.cpp) ----

def _stats_f64(x, n_const, correlation, eps=0.005):
    """FP64 statistics of the 1-based interior x (n x m): covariance, or
    correlation with a unit diagonal."""
    mean = x.sum(axis=0) / n_const
    c = x - mean
    if correlation:
        sd = np.sqrt((c * c).sum(axis=0) / n_const)
        sd = np.where(sd <= eps, 1.0, sd)
        c = c / np.sqrt(n_const) / sd
    s = c.T @ c
    if correlation:
        np.fill_diagonal(s, 1.0)
    return s


def test_covariance_inputs_follow_the_reference_layout(oracle):
    n = 7
    d = oracle.covariance_inputs(n)
    for i in range(n + 1):
        for j in range(n + 1):
            want = np.float32(np.float32(i * j) / n) if i < n and j < n else 0.0
            assert d[i, j] == want


def test_correlation_inputs_use_row_stride_n(oracle):
    n = 6
    flat = oracle.correlation_inputs(n).reshape(-1)
    want = np.zeros((n + 1) * (n + 1), np.float32)
    for i in range(n + 1):
        for j in range(n + 1):
            want[i * n + j] = np.float32(np.float32(i * j) / (n + 1))
    np.testing.assert_array_equal(flat, want)


@pytest.mark.parametrize("n", [16, 33, 96])
def test_covariance_oracle_matches_fp64(oracle, n):
    d, s = oracle.covariance_reference(n)
    want = _stats_f64(d[1:, 1:].astype(np.float64), float(oracle.COV_FLOAT_N), False)
    got = s[1:, 1:]
    assert np.array_equal(got, got.T)
    assert oracle.percent_diff_count(want, got) == 0


@pytest.mark.parametrize("n", [16, 33, 300])
def test_correlation_oracle_matches_fp64(oracle, n):
    d, s = oracle.correlation_reference(n)
    want = _stats_f64(d[1:, 1:].astype(np.float64), oracle.CORR_FLOAT_N, True)
    got = s[1:, 1:]
    assert np.array_equal(got, got.T)
    assert np.all(np.diag(got) == 1.0)
    assert oracle.percent_diff_count(want, got) == 0
    # at n = 300 the deviations exceed eps and the correlations are O(1): the
    # check above is then not vacuous (percentDiff skips values below 0.01)
    if n == 300:
        assert np.abs(got).max() > 0.5


def test_covariance_oracle_catches_the_textbook_mean(oracle):
    # The reference divides the column sums by the constant 3214212.01, not by
    # the number of observations; an oracle that used the textbook mean would
    # fail the reference's own rule.
    n = 64
    d, s = oracle.covariance_reference(n)
    textbook = _stats_f64(d[1:, 1:].astype(np.float64), float(n), False)
    assert oracle.percent_diff_count(textbook, s[1:, 1:]) > 0
