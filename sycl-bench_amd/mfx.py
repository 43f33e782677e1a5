"""Python view of libmfx.so (the C ABI declared in include/mfx.h).

Host-side plumbing for tests and bench.py: device memory comes from torch
tensors, the work runs in the library's HIP kernels.  Loading fails loudly
when libmfx.so has not been built -- there is no fallback path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmfx.so")

FP32, SPLIT_BF16, BF16, FP8_BLOCK, FP8_BLOCK2 = 0, 1, 2, 3, 4
PRECISIONS = {"fp32": FP32, "split-bf16": SPLIT_BF16, "bf16": BF16, "fp8-block": FP8_BLOCK, "fp8-block2": FP8_BLOCK2}
OP_N, OP_T = 0, 1
LOWER, UPPER, FULL = 0, 1, 2
STATUS = ["MFX_SUCCESS", "MFX_ERROR_INVALID_VALUE", "MFX_ERROR_NOT_SUPPORTED", "MFX_ERROR_HIP", "MFX_ERROR_ALLOC",
          "MFX_ERROR_NO_DEVICE", "MFX_ERROR_TIMEOUT", "MFX_ERROR_IO"]


class MfxError(RuntimeError):
    def __init__(self, status, fn, detail):
        self.status = status
        name = STATUS[status] if 0 <= status < len(STATUS) else str(status)
        super().__init__(f"{fn}: {name}: {detail}")


class Problem(ctypes.Structure):
    _fields_ = [("m", ctypes.c_int64), ("n", ctypes.c_int64), ("k", ctypes.c_int64),
                ("alpha", ctypes.c_float), ("beta", ctypes.c_float),
                ("A", ctypes.c_void_p), ("lda", ctypes.c_int64), ("transA", ctypes.c_int32),
                ("B", ctypes.c_void_p), ("ldb", ctypes.c_int64), ("transB", ctypes.c_int32),
                ("C", ctypes.c_void_p), ("ldc", ctypes.c_int64)]


class Dep(ctypes.Structure):
    _fields_ = [("consumer", ctypes.c_int32), ("producer", ctypes.c_int32)]


_V, _I64, _I32, _F, _D = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_float, ctypes.c_double
_ST = ctypes.c_int

# name -> (restype, argtypes); mirrors include/mfx.h one to one.
SIGNATURES = {
    "mfx_create": (_ST, [ctypes.POINTER(_V), ctypes.c_int]),
    "mfx_destroy": (_ST, [_V]),
    "mfx_reserve_workspace": (_ST, [_V, ctypes.c_size_t]),
    "mfx_gemm_workspace_size": (ctypes.c_size_t, [_ST, _I64, _I64, _I64]),
    "mfx_gemm_check": (_ST, [_ST, _ST, _ST, _I64, _I64, _I64, _I64, _I64, _I64]),
    "mfx_status_string": (ctypes.c_char_p, [_ST]),
    "mfx_last_error": (ctypes.c_char_p, []),
    "mfx_device_name": (_ST, [_V, ctypes.c_char_p, ctypes.c_size_t]),
    "mfx_gemm": (_ST, [_V, _ST, _ST, _ST, _I64, _I64, _I64, _F, _V, _I64, _V, _I64, _F, _V, _I64, _V]),
    "mfx_gemm_noalloc": (_ST, [_V, _ST, _ST, _ST, _I64, _I64, _I64, _F, _V, _I64, _V, _I64, _F, _V, _I64, _V]),
    "mfx_syrk": (_ST, [_V, _ST, _ST, ctypes.c_int, _I64, _I64, _F, _V, _I64, _F, _V, _I64, _V]),
    "mfx_syr2k": (_ST, [_V, _ST, _ST, ctypes.c_int, _I64, _I64, _F, _V, _I64, _V, _I64, _F, _V, _I64, _V]),
    "mfx_gemm_grouped": (_ST, [_V, _ST, ctypes.POINTER(Problem), _I32, ctypes.POINTER(Dep), _I32, _V]),
    "mfx_chain": (_ST, [_V, _ST, _I32, ctypes.POINTER(_V), _I64, ctypes.POINTER(_V), _V, _V]),
    "mfx_grouped_graph_create": (_ST, [_V, _ST, ctypes.POINTER(Problem), _I32, ctypes.POINTER(Dep), _I32, _V,
                                       ctypes.POINTER(_V)]),
    "mfx_graph_launch": (_ST, [_V, _V]),
    "mfx_graph_status": (_ST, [_V, _V]),
    "mfx_graph_destroy": (_ST, [_V]),
    "mfx_grouped_status": (_ST, [_V, _V]),
    "mfx_autotune_grouped": (_ST, [_V, _ST, ctypes.POINTER(Problem), _I32, ctypes.POINTER(Dep), _I32,
                                   ctypes.POINTER(_I32), _V]),
    "mfx_grouped_policy_count": (_I32, []),
    "mfx_grouped_policy_info": (_ST, [_I32, ctypes.POINTER(_I32), ctypes.POINTER(_I32), ctypes.POINTER(_D),
                                      ctypes.POINTER(_I32)]),
    "mfx_force_grouped_policy": (_ST, [_V, _I32]),
    "mfx_covariance": (_ST, [_V, _ST, _I64, _I64, _V, _I64, _F, _V, _V, _I64, _V, _I64, _V]),
    "mfx_correlation": (_ST, [_V, _ST, _I64, _I64, _V, _I64, _D, _D, _V, _V, _V, _I64, _V, _I64, _V]),
    "mfx_mfma_burn_waves": (_I64, [_V]),
    "mfx_mfma_burn": (_ST, [_V, _ST, _I64, _I32, _V, _I64, ctypes.POINTER(_D), ctypes.POINTER(_F), _V]),
    "mfx_hbm_copy": (_ST, [_V, _V, _V, _I64, _V]),
    "mfx_gemm_f64": (_ST, [_V, _ST, _ST, _I64, _I64, _I64, _D, _V, _I64, _V, _I64, _D, _V, _I64, _V]),
    "mfx_convert_f32_f64": (_ST, [_V, _I64, _I64, _V, _I64, _V, _I64, _V]),
    "mfx_rel_frobenius": (_ST, [_V, _ST, _I64, _I64, _V, _I64, _V, _I64, ctypes.POINTER(_D), _V]),
    "mfx_percent_diff_count": (_ST, [_V, _I64, _I64, _V, _I64, _V, _I64, _D, ctypes.POINTER(_I64), _V]),
    "mfx_fill_uniform": (_ST, [_V, _I64, _I64, _V, _I64, ctypes.c_uint64, ctypes.c_uint32, _V]),
    "mfx_fill_lognormal": (_ST, [_V, _I64, _I64, _V, _I64, ctypes.c_uint64, ctypes.c_uint32, _F, _F, _V]),
    "mfx_split_bf16": (_ST, [_V, _I64, _V, _V, _V, _V]),
    "mfx_quantize_fp8_block": (_ST, [_V, _I64, _I64, _V, _I64, _I32, _V, _V, _V]),
    "mfx_timing_begin": (_ST, [_V]),
    "mfx_timing_end": (_ST, [_V, _I32, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(_I64), ctypes.POINTER(_D),
                              ctypes.POINTER(_I32)]),
    "mfx_autotune": (_ST, [_V, _ST, _I64, _I64, _I64, ctypes.POINTER(_I32), _V]),
    "mfx_autotune_save": (_ST, [ctypes.c_char_p]),
    "mfx_autotune_load": (_ST, [ctypes.c_char_p]),
    "mfx_config_count": (_I32, [_ST]),
    "mfx_config_info": (_ST, [_ST, _I32, ctypes.POINTER(_I32), ctypes.POINTER(_I32), ctypes.POINTER(_I32),
                              ctypes.POINTER(_I32)]),
    "mfx_force_config": (_ST, [_V, _ST, _I32]),
    "mfx_force_splits": (_ST, [_V, _I32]),
}

_lib = None


def load():
    """Loads libmfx.so and declares every prototype; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() or `make -C sycl-bench_amd`")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def _check(status, fn):
    if status != 0:
        raise MfxError(status, fn, load().mfx_last_error().decode(errors="replace"))


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _ld(t):
    """Row stride of a 2-D tensor whose last dimension is contiguous."""
    assert t.dim() == 2 and (t.stride(1) == 1 or t.shape[1] <= 1), "last dimension must be contiguous"
    return max(t.stride(0), t.shape[1], 1)


def _stream(stream):
    if stream is None:
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return ctypes.c_void_p(stream)
    return ctypes.c_void_p(stream.cuda_stream)


def prec_id(p):
    return PRECISIONS[p] if isinstance(p, str) else int(p)


class Handle:
    """One mfx handle on a device; every method enqueues on `stream` (torch's current stream by default)."""

    def __init__(self, device=0):
        lib = load()
        h = ctypes.c_void_p()
        _check(lib.mfx_create(ctypes.byref(h), device), "mfx_create")
        self.h = h
        self.lib = lib

    def close(self):
        if self.h:
            self.lib.mfx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_name(self):
        buf = ctypes.create_string_buffer(256)
        _check(self.lib.mfx_device_name(self.h, buf, 256), "mfx_device_name")
        return buf.value.decode()

    def gemm(self, A, B, C, alpha=1.0, beta=0.0, prec="fp32", transA=False, transB=False, stream=None):
        m, n = C.shape
        k = A.shape[0] if transA else A.shape[1]
        _check(self.lib.mfx_gemm(self.h, prec_id(prec), int(transA), int(transB), m, n, k, alpha, _ptr(A), _ld(A),
                                 _ptr(B), _ld(B), beta, _ptr(C), _ld(C), _stream(stream)), "mfx_gemm")

    def gemm_noalloc(self, A, B, C, alpha=1.0, beta=0.0, prec="split-bf16", transA=False, transB=False,
                     stream=None):
        """mfx_gemm_noalloc: mfx_gemm with its packed operands in the library's fixed per-device array."""
        m, n = C.shape
        k = A.shape[0] if transA else A.shape[1]
        _check(self.lib.mfx_gemm_noalloc(self.h, prec_id(prec), int(transA), int(transB), m, n, k, alpha, _ptr(A),
                                         _ld(A), _ptr(B), _ld(B), beta, _ptr(C), _ld(C), _stream(stream)),
               "mfx_gemm_noalloc")

    def syrk(self, A, C, alpha=1.0, beta=0.0, prec="fp32", uplo=LOWER, mirror=True, stream=None):
        n, k = A.shape
        _check(self.lib.mfx_syrk(self.h, prec_id(prec), uplo, int(mirror), n, k, alpha, _ptr(A), _ld(A), beta,
                                 _ptr(C), _ld(C), _stream(stream)), "mfx_syrk")

    def syr2k(self, A, B, C, alpha=1.0, beta=0.0, prec="fp32", uplo=LOWER, mirror=True, stream=None):
        n, k = A.shape
        _check(self.lib.mfx_syr2k(self.h, prec_id(prec), uplo, int(mirror), n, k, alpha, _ptr(A), _ld(A), _ptr(B),
                                  _ld(B), beta, _ptr(C), _ld(C), _stream(stream)), "mfx_syr2k")

    def covariance(self, data, mean, centered, symmat, n_const, prec="fp32", stream=None):
        """data: n x m observations; mean (m), centered (n x m), symmat (m x m) are outputs."""
        n, m = data.shape
        _check(self.lib.mfx_covariance(self.h, prec_id(prec), m, n, _ptr(data), _ld(data), n_const, _ptr(mean),
                                       _ptr(centered), _ld(centered), _ptr(symmat), _ld(symmat), _stream(stream)),
               "mfx_covariance")

    def correlation(self, data, mean, stddev, centered, symmat, n_const, eps, prec="fp32", stream=None):
        n, m = data.shape
        _check(self.lib.mfx_correlation(self.h, prec_id(prec), m, n, _ptr(data), _ld(data), n_const, eps, _ptr(mean),
                                        _ptr(stddev), _ptr(centered), _ld(centered), _ptr(symmat), _ld(symmat),
                                        _stream(stream)), "mfx_correlation")

    def mfma_burn(self, out, iterations, prec="split-bf16", random_operands=False, stream=None):
        """Register-only MFMA loop; returns (flop, expected per-wave checksum)."""
        flop, expected = ctypes.c_double(), ctypes.c_float()
        _check(self.lib.mfx_mfma_burn(self.h, prec_id(prec), iterations, int(bool(random_operands)), _ptr(out),
                                      out.numel(), ctypes.byref(flop), ctypes.byref(expected), _stream(stream)),
               "mfx_mfma_burn")
        return flop.value, expected.value

    def mfma_burn_waves(self):
        return self.lib.mfx_mfma_burn_waves(self.h)

    def hbm_copy(self, src, dst, stream=None):
        nbytes = src.numel() * src.element_size()
        assert dst.numel() * dst.element_size() >= nbytes
        _check(self.lib.mfx_hbm_copy(self.h, _ptr(src), _ptr(dst), nbytes, _stream(stream)), "mfx_hbm_copy")

    def gemm_f64(self, A, B, C, alpha=1.0, beta=0.0, transA=False, transB=False, stream=None):
        m, n = C.shape
        k = A.shape[0] if transA else A.shape[1]
        _check(self.lib.mfx_gemm_f64(self.h, int(transA), int(transB), m, n, k, alpha, _ptr(A), _ld(A), _ptr(B),
                                     _ld(B), beta, _ptr(C), _ld(C), _stream(stream)), "mfx_gemm_f64")

    def convert_f64(self, src, dst, stream=None):
        r, c = src.shape
        _check(self.lib.mfx_convert_f32_f64(self.h, r, c, _ptr(src), _ld(src), _ptr(dst), _ld(dst),
                                            _stream(stream)), "mfx_convert_f32_f64")

    def rel_frobenius(self, X, R, uplo=FULL, stream=None):
        out = ctypes.c_double()
        r, c = X.shape
        _check(self.lib.mfx_rel_frobenius(self.h, uplo, r, c, _ptr(X), _ld(X), _ptr(R), _ld(R), ctypes.byref(out),
                                          _stream(stream)), "mfx_rel_frobenius")
        return out.value

    def percent_diff_count(self, X, R, threshold=0.05, stream=None):
        out = ctypes.c_int64()
        r, c = X.shape
        _check(self.lib.mfx_percent_diff_count(self.h, r, c, _ptr(X), _ld(X), _ptr(R), _ld(R), threshold,
                                               ctypes.byref(out), _stream(stream)), "mfx_percent_diff_count")
        return out.value

    def fill_uniform(self, X, seed, matrix_id, stream=None):
        r, c = X.shape
        _check(self.lib.mfx_fill_uniform(self.h, r, c, _ptr(X), _ld(X), seed, matrix_id, _stream(stream)),
               "mfx_fill_uniform")

    def fill_lognormal(self, X, seed, matrix_id, mu=0.0, sigma=2.0, stream=None):
        r, c = X.shape
        _check(self.lib.mfx_fill_lognormal(self.h, r, c, _ptr(X), _ld(X), seed, matrix_id, mu, sigma,
                                           _stream(stream)), "mfx_fill_lognormal")

    def split_bf16(self, x, hi, lo, stream=None):
        _check(self.lib.mfx_split_bf16(self.h, x.numel(), _ptr(x), _ptr(hi), _ptr(lo), _stream(stream)),
               "mfx_split_bf16")

    def quantize_fp8_block(self, x, q, exps, block=128, stream=None):
        r, c = x.shape
        _check(self.lib.mfx_quantize_fp8_block(self.h, r, c, _ptr(x), _ld(x), block, _ptr(q), _ptr(exps),
                                               _stream(stream)), "mfx_quantize_fp8_block")

    def reserve_workspace(self, nbytes):
        _check(self.lib.mfx_reserve_workspace(self.h, nbytes), "mfx_reserve_workspace")

    def force_config(self, prec, config):
        _check(self.lib.mfx_force_config(self.h, prec_id(prec), config), "mfx_force_config")

    def force_splits(self, splits):
        """FP32 split-K override: -1 automatic, 1 never, >= 2 that many k slices."""
        _check(self.lib.mfx_force_splits(self.h, splits), "mfx_force_splits")

    def force_grouped_policy(self, policy):
        _check(self.lib.mfx_force_grouped_policy(self.h, policy), "mfx_force_grouped_policy")

    def timing_begin(self):
        _check(self.lib.mfx_timing_begin(self.h), "mfx_timing_begin")

    def timing_end(self):
        cap = 64
        names = (ctypes.c_char_p * cap)()
        launches = (ctypes.c_int64 * cap)()
        total = (ctypes.c_double * cap)()
        n = ctypes.c_int32()
        _check(self.lib.mfx_timing_end(self.h, cap, names, launches, total, ctypes.byref(n)), "mfx_timing_end")
        return {names[i].decode(): (launches[i], total[i]) for i in range(min(n.value, cap))}

    def grouped(self, problems, deps=(), prec="split-bf16", stream=None):
        arr, darr = _problem_arrays(problems, deps)
        _check(self.lib.mfx_gemm_grouped(self.h, prec_id(prec), arr, len(problems), darr, len(deps),
                                         _stream(stream)), "mfx_gemm_grouped")

    def grouped_status(self, stream=None):
        """Blocking; raises MfxError(MFX_ERROR_TIMEOUT) if the last grouped launch skipped work."""
        _check(self.lib.mfx_grouped_status(self.h, _stream(stream)), "mfx_grouped_status")

    def autotune_grouped(self, problems, deps=(), prec="split-bf16", stream=None):
        arr, darr = _problem_arrays(problems, deps)
        best = ctypes.c_int32()
        _check(self.lib.mfx_autotune_grouped(self.h, prec_id(prec), arr, len(problems), darr, len(deps),
                                             ctypes.byref(best), _stream(stream)), "mfx_autotune_grouped")
        return best.value

    def chain(self, mats, scratch, out, prec="split-bf16", stream=None):
        n = out.shape[0]
        mp = (ctypes.c_void_p * len(mats))(*[m.data_ptr() for m in mats])
        sp = (ctypes.c_void_p * max(1, len(scratch)))(*[s.data_ptr() for s in scratch])
        _check(self.lib.mfx_chain(self.h, prec_id(prec), len(mats), mp, n, sp, _ptr(out), _stream(stream)),
               "mfx_chain")

    def graph_create(self, problems, deps=(), prec="split-bf16", stream=None):
        arr, darr = _problem_arrays(problems, deps)
        g = ctypes.c_void_p()
        _check(self.lib.mfx_grouped_graph_create(self.h, prec_id(prec), arr, len(problems), darr, len(deps),
                                                 _stream(stream), ctypes.byref(g)), "mfx_grouped_graph_create")
        return Graph(self.lib, g)

    def autotune(self, prec, m, n, k, stream=None):
        best = ctypes.c_int32()
        _check(self.lib.mfx_autotune(self.h, prec_id(prec), m, n, k, ctypes.byref(best), _stream(stream)),
               "mfx_autotune")
        return best.value


class Graph:
    def __init__(self, lib, g):
        self.lib, self.g = lib, g

    def launch(self, stream=None):
        _check(self.lib.mfx_graph_launch(self.g, _stream(stream)), "mfx_graph_launch")

    def status(self, stream=None):
        _check(self.lib.mfx_graph_status(self.g, _stream(stream)), "mfx_graph_status")

    def close(self):
        if self.g:
            self.lib.mfx_graph_destroy(self.g)
            self.g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def problem(A, B, C, alpha=1.0, beta=0.0, transA=False, transB=False):
    m, n = C.shape
    k = A.shape[0] if transA else A.shape[1]
    return Problem(m, n, k, alpha, beta, A.data_ptr(), _ld(A), int(transA), B.data_ptr(), _ld(B), int(transB),
                   C.data_ptr(), _ld(C))


def _problem_arrays(problems, deps):
    arr = (Problem * max(1, len(problems)))(*problems)
    darr = (Dep * max(1, len(deps)))(*[Dep(c, p) for c, p in deps])
    return arr, darr


def config_count(prec):
    return load().mfx_config_count(prec_id(prec))


def config_info(prec, c):
    v = [ctypes.c_int32() for _ in range(4)]
    _check(load().mfx_config_info(prec_id(prec), c, *[ctypes.byref(x) for x in v]), "mfx_config_info")
    return tuple(x.value for x in v)


def grouped_policy_count():
    return load().mfx_grouped_policy_count()


def grouped_policy_info(policy):
    """(large config, small config, waste bound, one launch per level) of a grouped tile policy."""
    big, small, bound, lv = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_double(), ctypes.c_int32()
    _check(load().mfx_grouped_policy_info(policy, ctypes.byref(big), ctypes.byref(small), ctypes.byref(bound),
                                          ctypes.byref(lv)), "mfx_grouped_policy_info")
    return big.value, small.value, bound.value, bool(lv.value)
