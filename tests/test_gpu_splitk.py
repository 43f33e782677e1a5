"""FP32 split-K parity (GPU, through the C ABI).

A single FP32 product whose grid would leave the CUs short of work runs
each tile as several k slices (one kernel writes the partial tiles, a second
sums them in slice order and applies alpha / beta).  Checks, against the FP64
oracle product at the FP32 bar (relative Frobenius error 2e-6 on U[-1, 1)
inputs), with the 64x64 four-wave tile (configuration 2) unless named:

- the split path is the one that ran (its kernel name in the timing records),
- every operand layout, ragged edges, padded leading dimensions and the
  scalar-load path (odd K),
- beta = 0 never reads C (C starts as NaN),
- repeated calls are bit-identical (the reduction order is fixed),
- back-to-back calls of different shapes on the shared partial buffer,
- a slice count forced through mfx_force_splits, and 1 turning split-K off,
- the k-grouped tile configuration (7: four groups of four waves summed
  through LDS) unsplit, on full and triangular schedules, and as the
  automatic choice for a grid of at most one 64x64 tile per CU.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = "cuda"
TOL = 2e-6  # relative Frobenius error vs FP64, FP32 method


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def padded(a, pad):
    """Device copy of `a` with leading dimension cols + pad."""
    rows, cols = a.shape
    base = torch.full((rows, cols + pad), float("nan"), device=DEV)
    view = base[:, :cols]
    view.copy_(to_dev(a))
    return view


def run(handle, A, B, C, alpha, beta, ta, tb):
    handle.timing_begin()
    handle.gemm(A, B, C, alpha=alpha, beta=beta, prec="fp32", transA=ta, transB=tb)
    rec = handle.timing_end()
    torch.cuda.synchronize()
    return rec


def case(oracle, m, n, k, ta, tb, seed, beta):
    rng = np.random.default_rng(seed)
    A = rng.uniform(-1, 1, (k, m) if ta else (m, k)).astype(np.float32)
    B = rng.uniform(-1, 1, (n, k) if tb else (k, n)).astype(np.float32)
    C = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    want = oracle.gemm_f64(A.astype(np.float64), B.astype(np.float64), C.astype(np.float64) if beta else None,
                           1.5, beta, ta=ta, tb=tb)
    return A, B, C, want


@pytest.mark.parametrize("m,n,k,ta,tb,pad", [
    (1024, 1024, 1024, False, False, 0),   # the gemm workload's shape
    (1024, 1024, 1024, True, True, 4),
    (300, 517, 1000, False, True, 4),      # ragged tiles, padded rows
    (517, 300, 2048, True, False, 0),
    (640, 320, 1001, False, False, 3),     # odd K: scalar loads
    (64, 64, 4096, True, True, 0),         # one tile, eight slices
])
@pytest.mark.parametrize("beta", [1.2, 0.0])
def test_splitk_matches_fp64(handle, oracle, m, n, k, ta, tb, pad, beta):
    A, B, C, want = case(oracle, m, n, k, ta, tb, m + 3 * n + 7 * k, beta)
    print(f"split-K m={m} n={n} k={k} ta={ta} tb={tb} pad={pad} beta={beta}", flush=True)
    Ad, Bd = padded(A, pad), padded(B, pad)
    Cd = padded(C if beta else np.full_like(C, np.nan), pad)
    handle.force_config("fp32", 2)
    try:
        rec = run(handle, Ad, Bd, Cd, 1.5, beta, ta, tb)
    finally:
        handle.force_config("fp32", -1)
    assert "gemm_f32_splitk" in rec, rec
    got = Cd.cpu().numpy()
    assert np.isfinite(got).all()
    err = oracle.rel_frobenius(got, want)
    print(f"  rel Frobenius error {err:.3e}", flush=True)
    assert err < TOL


@pytest.mark.parametrize("config", [1, 2, 6, 7])
def test_splitk_each_tile_config(handle, oracle, config):
    m, n, k = 512, 768, 1536
    A, B, C, want = case(oracle, m, n, k, False, False, 100 + config, 1.2)
    print(f"split-K forced config {config}", flush=True)
    handle.force_config("fp32", config)
    try:
        Cd = to_dev(C)
        rec = run(handle, to_dev(A), to_dev(B), Cd, 1.5, 1.2, False, False)
    finally:
        handle.force_config("fp32", -1)
    assert "gemm_f32_splitk" in rec, rec
    assert oracle.rel_frobenius(Cd.cpu().numpy(), want) < TOL


def test_splitk_repeat_bit_identical(handle, oracle):
    """Same call twice gives the same bits; different shapes in between and
    after still verify."""
    shapes = [(1024, 1024, 1024), (384, 640, 2048), (1024, 1024, 1024), (128, 192, 777)]
    outs = []
    for i, (m, n, k) in enumerate(shapes):
        A, B, C, want = case(oracle, m, n, k, False, False, 7, 1.2)
        print(f"split-K sequence call {i}: {m}x{n}x{k}", flush=True)
        Cd = to_dev(C)
        handle.force_config("fp32", 2)
        try:
            rec = run(handle, to_dev(A), to_dev(B), Cd, 1.5, 1.2, False, False)
        finally:
            handle.force_config("fp32", -1)
        assert "gemm_f32_splitk" in rec, rec
        got = Cd.cpu().numpy()
        assert oracle.rel_frobenius(got, want) < TOL
        outs.append(got)
    np.testing.assert_array_equal(outs[0], outs[2])


def test_splitk_large_grid(handle, oracle):
    """64x64 tiles, 2 x CUs of them: two slices each."""
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    m, n, k = 1024, 2 * cus * 64 * 64 // 1024, 1024
    A, B, C, want = case(oracle, m, n, k, False, False, 5, 1.2)
    print(f"split-K large grid {m}x{n}x{k} ({2 * cus} tiles)", flush=True)
    handle.force_config("fp32", 2)
    try:
        Cd = to_dev(C)
        rec = run(handle, to_dev(A), to_dev(B), Cd, 1.5, 1.2, False, False)
    finally:
        handle.force_config("fp32", -1)
    assert "gemm_f32_splitk" in rec, rec
    assert oracle.rel_frobenius(Cd.cpu().numpy(), want) < TOL


@pytest.mark.parametrize("splits", [1, 2, 3, 5])
def test_forced_splits(handle, mfx, oracle, splits):
    m, n, k = 700, 900, 1100
    A, B, C, want = case(oracle, m, n, k, True, False, 40 + splits, 1.2)
    print(f"forced splits {splits}", flush=True)
    handle.force_config("fp32", 1)
    handle.force_splits(splits)
    try:
        Cd = to_dev(C)
        rec = run(handle, to_dev(A), to_dev(B), Cd, 1.5, 1.2, True, False)
    finally:
        handle.force_config("fp32", -1)
        handle.force_splits(-1)
    assert ("gemm_f32_splitk" in rec) == (splits > 1), rec
    assert ("gemm_f32_mfma" in rec) == (splits == 1), rec
    assert oracle.rel_frobenius(Cd.cpu().numpy(), want) < TOL
    for bad in (0, -2, 65):
        with pytest.raises(mfx.MfxError):
            handle.force_splits(bad)


@pytest.mark.parametrize("config", [7])
@pytest.mark.parametrize("m,n,k,ta,tb", [(1024, 1024, 1024, False, False), (333, 270, 517, True, False),
                                         (200, 130, 1001, False, True), (64, 64, 40, True, True)])
def test_kgroup_tiles_unsplit(handle, oracle, config, m, n, k, ta, tb):
    A, B, C, want = case(oracle, m, n, k, ta, tb, config + m + k, 1.2)
    print(f"k-grouped config {config} {m}x{n}x{k} ta={ta} tb={tb}", flush=True)
    handle.force_config("fp32", config)
    handle.force_splits(1)
    try:
        Cd = padded(C, 4)
        rec = run(handle, padded(A, 0), padded(B, 4), Cd, 1.5, 1.2, ta, tb)
    finally:
        handle.force_config("fp32", -1)
        handle.force_splits(-1)
    assert list(rec) == ["gemm_f32_mfma"], rec
    assert oracle.rel_frobenius(Cd.cpu().numpy(), want) < TOL


@pytest.mark.parametrize("config", [7])
def test_kgroup_tiles_syrk(handle, oracle, config):
    n, k = 450, 700
    rng = np.random.default_rng(config)
    A = rng.uniform(-1, 1, (n, k)).astype(np.float32)
    C = rng.uniform(-1, 1, (n, n)).astype(np.float32)
    C = ((C + C.T) / 2).astype(np.float32)
    print(f"k-grouped config {config} syrk n={n} k={k}", flush=True)
    handle.force_config("fp32", config)
    try:
        Cd = to_dev(C)
        handle.syrk(to_dev(A), Cd, alpha=1.5, beta=1.2, prec="fp32")
        torch.cuda.synchronize()
    finally:
        handle.force_config("fp32", -1)
    want = oracle.gemm_f64(A.astype(np.float64), A.astype(np.float64), C.astype(np.float64), 1.5, 1.2, tb=True)
    assert oracle.rel_frobenius(Cd.cpu().numpy(), want) < TOL


@pytest.mark.parametrize("m,n,k,splits", [(1024, 1024, 1024, 1), (512, 512, 512, 4), (512, 512, 8192, 4),
                                          (2048, 2048, 2048, None)])
def test_automatic_choice(handle, oracle, m, n, k, splits):
    """The heuristic's picks: the k-grouped tile while its 64x64 tiles fit one
    per CU (with k slices when there are few), the four-wave tiles beyond."""
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    A, B, C, want = case(oracle, m, n, k, False, False, 9, 1.2)
    print(f"automatic choice {m}x{n}x{k}", flush=True)
    Cd = to_dev(C)
    rec = run(handle, to_dev(A), to_dev(B), Cd, 1.5, 1.2, False, False)
    assert oracle.rel_frobenius(Cd.cpu().numpy(), want) < TOL
    if cus == 256 and splits is not None:
        assert ("gemm_f32_splitk" in rec) == (splits > 1), rec
