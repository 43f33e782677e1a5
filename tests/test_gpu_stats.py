"""Covariance / correlation (mfx_covariance, mfx_correlation) on the GPU.

Bars: on the reference's own inputs, the reference's verify rule (percentDiff
<= 0.05 % against the oracle's restatement of the reference CPU loops, every
element of the 1-based N x N block); on seeded, ragged inputs, relative
Frobenius error against FP64 with the precision's tolerance; the input left
untouched and repeated calls bit-identical; a unit correlation diagonal."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda"
TOL = {"fp32": 2e-6, "split-bf16": 1e-5, "bf16": 1e-2}


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def run_cov(handle, data, prec, n_const):
    n, m = data.shape
    mean = torch.empty(m, device=DEV)
    cen = torch.empty((n, m), device=DEV)
    sym = torch.full((m, m), float("nan"), device=DEV)
    handle.covariance(data, mean, cen, sym, n_const, prec=prec)
    return sym, mean, cen


def run_corr(handle, data, prec, n_const, eps):
    n, m = data.shape
    mean = torch.empty(m, device=DEV)
    sd = torch.empty(m, device=DEV)
    cen = torch.empty((n, m), device=DEV)
    sym = torch.full((m, m), float("nan"), device=DEV)
    handle.correlation(data, mean, sd, cen, sym, n_const, eps, prec=prec)
    return sym, mean, sd


@pytest.mark.parametrize("prec", ["fp32", "split-bf16"])
@pytest.mark.parametrize("n", [64, 130, 257, 520])
def test_covariance_reference_inputs(handle, oracle, prec, n):
    d, want = oracle.covariance_reference(n)
    full = to_dev(d)
    data = full[1:, 1:]  # the 1-based interior, ld = n + 1
    sym, _, _ = run_cov(handle, data, prec, float(oracle.COV_FLOAT_N))
    got = sym.cpu().numpy()
    assert oracle.percent_diff_count(want[1:, 1:], got) == 0
    assert torch.equal(full, to_dev(d))  # input untouched


@pytest.mark.parametrize("prec", ["fp32", "split-bf16"])
@pytest.mark.parametrize("n", [64, 300, 513])
def test_correlation_reference_inputs(handle, oracle, prec, n):
    d, want = oracle.correlation_reference(n)
    data = to_dev(d)[1:, 1:]
    sym, _, _ = run_corr(handle, data, prec, oracle.CORR_FLOAT_N, oracle.CORR_EPS)
    got = sym.cpu().numpy()
    assert np.all(np.diag(got) == 1.0)
    assert oracle.percent_diff_count(want[1:, 1:], got) == 0


def f64_stats(x, n_const, correlation, eps=0.005):
    x = x.double()
    mean = x.sum(0) / n_const
    c = x - mean
    if correlation:
        sd = torch.sqrt((c * c).sum(0) / n_const)
        sd = torch.where(sd <= eps, torch.ones_like(sd), sd)
        c = c / np.sqrt(n_const) / sd
    s = c.T.contiguous() @ c
    if correlation:
        s.fill_diagonal_(1.0)
    return s


@pytest.mark.parametrize("prec", ["fp32", "split-bf16", "bf16"])
@pytest.mark.parametrize("n,m", [(300, 200), (129, 257), (1024, 512)])
def test_seeded_ragged_against_fp64(handle, prec, n, m):
    data = torch.empty((n, m), device=DEV)
    handle.fill_uniform(data, 21, 0)
    data += 0.25  # a non-zero mean to subtract
    for correlation in (False, True):
        n_const = float(n) if correlation else 1000.0
        if correlation:
            sym, mean, sd = run_corr(handle, data, prec, n_const, 0.005)
        else:
            sym, mean, _ = run_cov(handle, data, prec, n_const)
        ref = f64_stats(data, n_const, correlation)
        err = (torch.linalg.norm(sym.double() - ref) / torch.linalg.norm(ref)).item()
        assert err < TOL[prec] * 4, (correlation, err)
        assert torch.allclose(mean.double(), data.double().sum(0) / n_const, rtol=1e-5, atol=1e-7)


def test_repeated_calls_are_identical(handle):
    data = torch.empty((400, 300), device=DEV)
    handle.fill_uniform(data, 5, 1)
    a, _, _ = run_corr(handle, data, "split-bf16", 400.0, 0.005)
    b, _, _ = run_corr(handle, data, "split-bf16", 400.0, 0.005)
    assert torch.equal(a, b)
    c, _, _ = run_cov(handle, data, "fp32", 400.0)
    e, _, _ = run_cov(handle, data, "fp32", 400.0)
    assert torch.equal(c, e)


def test_small_deviation_is_replaced_by_one(handle):
    # every column constant but one: their deviations are 0 <= eps -> 1
    data = torch.ones((50, 40), device=DEV)
    data[:, 3] = torch.arange(50, device=DEV, dtype=torch.float32)
    sym, mean, sd = run_corr(handle, data, "fp32", 50.0, 0.005)
    sdh = sd.cpu().numpy()
    assert np.all(sdh[np.arange(40) != 3] == 1.0) and sdh[3] > 1.0
    assert torch.isfinite(sym).all()


def test_invalid_arguments(handle, mfx):
    data = torch.empty((10, 8), device=DEV)
    mean = torch.empty(8, device=DEV)
    cen = torch.empty((10, 8), device=DEV)
    sym = torch.empty((8, 8), device=DEV)
    with pytest.raises(mfx.MfxError):
        handle.covariance(data, mean, cen, sym, 10.0, prec="fp8-block")
    small = torch.empty((8, 4), device=DEV)
    with pytest.raises(mfx.MfxError):
        handle.covariance(data, mean, cen, small[:, :3], 10.0)
