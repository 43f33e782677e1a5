"""The C-ABI library loads and exports exactly what include/mfx.h declares (CPU only)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mfx.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mfx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("mfx_gemm", "mfx_syrk", "mfx_syr2k", "mfx_gemm_grouped", "mfx_chain", "mfx_gemm_f64",
                     "mfx_rel_frobenius", "mfx_autotune", "mfx_status_string", "mfx_timing_begin", "mfx_timing_end"):
        assert required in names


def test_library_exports_every_declared_symbol(mfx):
    lib = mfx.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert missing == []


def test_binding_covers_every_declared_symbol(mfx):
    assert sorted(mfx.SIGNATURES) == declared_functions()


def test_status_strings(mfx):
    lib = mfx.load()
    for code, name in enumerate(mfx.STATUS):
        assert lib.mfx_status_string(code).decode() == name
    assert lib.mfx_status_string(99).decode() == "MFX_ERROR_UNKNOWN"


def test_config_tables(mfx):
    for prec in ("fp32", "split-bf16", "bf16"):
        n = mfx.config_count(prec)
        assert n >= 3
        for c in range(n):
            bm, bn, bk, waves = mfx.config_info(prec, c)
            assert bm % 32 == 0 and bn % 32 == 0 and bk % 2 == 0 and waves >= 1
    with pytest.raises(mfx.MfxError):
        mfx.config_info("fp32", 99)


def test_workspace_size_is_positive_for_packed_precisions(mfx):
    lib = mfx.load()
    # FP32 needs workspace only for split-K partial tiles: small grids, at most 32 MiB
    assert 0 < lib.mfx_gemm_workspace_size(mfx.FP32, 1024, 1024, 1024) <= 32 << 20
    assert lib.mfx_gemm_workspace_size(mfx.FP32, 8192, 8192, 8192) == 0
    ws = lib.mfx_gemm_workspace_size(mfx.SPLIT_BF16, 1000, 1000, 1000)
    # hi + lo BF16 copies of both padded operands
    assert ws >= 2 * 2 * 2 * 1000 * 1000
    assert lib.mfx_gemm_workspace_size(mfx.BF16, 1000, 1000, 1000) < ws


def test_create_without_gpu_reports_no_device(mfx):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = mfx.load()
    h = ctypes.c_void_p()
    assert lib.mfx_create(ctypes.byref(h), 0) == 5  # MFX_ERROR_NO_DEVICE
    assert lib.mfx_last_error()


def test_null_handle_is_rejected(mfx):
    lib = mfx.load()
    st = lib.mfx_gemm(None, 0, 0, 0, 4, 4, 4, 1.0, None, 4, None, 4, 0.0, None, 4, None)
    assert st == 1  # MFX_ERROR_INVALID_VALUE
    assert b"null handle" in lib.mfx_last_error()


def test_gemm_check_is_host_only_and_bounds_k(mfx):
    """The pre-flight check needs no device; the packed BF16 paths refuse K at
    the limit of their 32-bit panel offsets (8 * kbs * planes * 64 < 2^31)."""
    lib = mfx.load()
    N, T = mfx.OP_N, mfx.OP_T
    assert lib.mfx_gemm_check(mfx.SPLIT_BF16, N, N, 64, 64, 64, 64, 64, 64) == 0
    assert lib.mfx_gemm_check(mfx.FP32, N, N, 64, 64, 64, 63, 64, 64) == 1  # lda < k
    assert lib.mfx_gemm_check(mfx.FP32, T, N, 64, 64, 100, 64, 64, 64) == 0  # A is k x m: lda >= m
    assert lib.mfx_gemm_check(7, N, N, 64, 64, 64, 64, 64, 64) == 1
    assert lib.mfx_gemm_check(mfx.FP32, N, N, -1, 64, 64, 64, 64, 64) == 1
    # split BF16: k blocks of 16 padded to the 32-wide stage; 8 * kbs * 2 * 64 < 2^31  <=>  K <= 2^25 - 32
    k_max = (1 << 25) - 32
    assert lib.mfx_gemm_check(mfx.SPLIT_BF16, N, T, 32, 32, k_max, k_max, k_max, 32) == 0
    assert lib.mfx_gemm_check(mfx.SPLIT_BF16, N, T, 32, 32, k_max + 1, k_max + 1, k_max + 1, 32) == 2
    assert b"limit" in lib.mfx_last_error()
    # single BF16: 64-wide stage, one plane  <=>  K <= 2^26 - 64
    k1 = (1 << 26) - 64
    assert lib.mfx_gemm_check(mfx.BF16, N, T, 32, 32, k1, k1, k1, 32) == 0
    assert lib.mfx_gemm_check(mfx.BF16, N, T, 32, 32, k1 + 1, k1 + 1, k1 + 1, 32) == 2
    # FP32 and FP8 index with 64-bit offsets: no such limit
    assert lib.mfx_gemm_check(mfx.FP32, N, T, 32, 32, 1 << 30, 1 << 30, 1 << 30, 32) == 0
    # empty outputs need no contraction, so no K limit applies
    assert lib.mfx_gemm_check(mfx.SPLIT_BF16, N, T, 0, 32, 1 << 30, 1 << 30, 1 << 30, 32) == 0
