"""ctypes view of the CPU oracle (oracle.c) and CPU baseline (cpu_baseline.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, to compare against.  The library and the benchmark
executables never load it.

Every wrapper takes and returns numpy arrays; each C function cites the
reference lines it restates.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_F32P = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_F64P = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_U16P = np.ctypeslib.ndpointer(dtype=np.uint16, flags="C_CONTIGUOUS")
_U8P = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_I32P = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_SZ = ctypes.c_size_t
_I64 = ctypes.c_int64


def _load(name):
    path = os.path.join(_HERE, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run `make -C oracle` (or __graft_entry__.build())")
    return ctypes.CDLL(path)


_lib = _load("liboracle.so")


def _sig(fn, res, *args):
    f = getattr(_lib, fn)
    f.restype = res
    f.argtypes = list(args)
    return f


_sig("gemm_init", None, _F32P, _F32P, _F32P, _SZ)
_sig("gemm_cpu", None, _F32P, _F32P, _F32P, _SZ)
_sig("mm2_init", None, _F32P, _F32P, _F32P, _F32P, _SZ)
_sig("mm2_cpu", None, _F32P, _F32P, _F32P, _F32P, _F32P, _SZ)
_sig("mm3_init", None, _F32P, _F32P, _F32P, _F32P, _SZ)
_sig("mm3_cpu", None, _F32P, _F32P, _F32P, _F32P, _F32P, _F32P, _F32P, _SZ)
_sig("syrk_init", None, _F32P, _F32P, _SZ)
_sig("covariance_init", None, _F32P, _SZ)
_sig("covariance_cpu", None, _F32P, _F32P, _F32P, _SZ)
_sig("correlation_init", None, _F32P, _SZ)
_sig("correlation_cpu", None, _F32P, _F32P, _F32P, _F32P, _SZ)
_sig("syrk_cpu", None, _F32P, _F32P, _SZ)
_sig("syr2k_init", None, _F32P, _F32P, _F32P, _SZ)
_sig("syr2k_cpu", None, _F32P, _F32P, _F32P, _SZ)
_sig("matmulchain_identity", None, _F32P, _SZ)
_sig("matmulchain_multiply_int", None, _F32P, _F32P, _F32P, _SZ)
_sig("percent_diff", ctypes.c_float, ctypes.c_double, ctypes.c_double)
_sig("percent_diff_count", _I64, _F32P, _F32P, _SZ, ctypes.c_double)
for _n in ("gemm_closed", "mm2_closed", "mm3_closed", "syrk_closed", "syr2k_closed"):
    _sig(_n, None, _F64P, _SZ)
_sig("gemm_f64", None, ctypes.c_int, ctypes.c_int, _I64, _I64, _I64, ctypes.c_double, _F64P, _I64, _F64P, _I64,
     ctypes.c_double, _F64P, _I64)
_sig("keyed_hash", ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64)
_sig("splitmix64_mix", ctypes.c_uint64, ctypes.c_uint64)
_sig("fill_uniform", None, _F32P, _I64, _I64, _I64, ctypes.c_uint64, ctypes.c_uint32)
_sig("fill_lognormal", None, _F32P, _I64, _I64, _I64, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_double,
     ctypes.c_double)
_sig("bf16_rne", ctypes.c_uint16, ctypes.c_float)
_sig("split_bf16", None, _F32P, _U16P, _U16P, _I64)
_sig("e4m3_rne_sat", ctypes.c_uint8, ctypes.c_float)
_sig("e4m3_to_f32", ctypes.c_float, ctypes.c_uint8)
_sig("fp8_block_exponent", ctypes.c_int, ctypes.c_float)
_sig("quantize_fp8_block", None, _F32P, _I64, _I64, _I64, ctypes.c_int, _U8P, _I32P)


def _z(n, dtype=np.float32):
    return np.zeros(n, dtype=dtype)


# ---- reference benchmarks: inputs and serial answers ----------------------
def gemm_reference(n):
    """(A, B, C_in, C_out) of polybench/gemm.cpp:18-56 at size n."""
    A, B, C = _z(n * n), _z(n * n), _z(n * n)
    _lib.gemm_init(A, B, C, n)
    out = C.copy()
    _lib.gemm_cpu(A, B, out, n)
    return A.reshape(n, n), B.reshape(n, n), C.reshape(n, n), out.reshape(n, n)


def gemm_inputs(n):
    A, B, C = _z(n * n), _z(n * n), _z(n * n)
    _lib.gemm_init(A, B, C, n)
    return A.reshape(n, n), B.reshape(n, n), C.reshape(n, n)


def mm2_reference(n):
    """(A, B, C_in, D, E) of polybench/2mm.cpp:16-69; E is the CPU answer."""
    A, B, C, D, E = (_z(n * n) for _ in range(5))
    _lib.mm2_init(A, B, C, D, n)
    Cin = C.copy()
    _lib.mm2_cpu(A, B, C, D, E, n)
    return tuple(x.reshape(n, n) for x in (A, B, Cin, D, E))


def mm2_inputs(n):
    A, B, C, D = (_z(n * n) for _ in range(4))
    _lib.mm2_init(A, B, C, D, n)
    return tuple(x.reshape(n, n) for x in (A, B, C, D))


def mm3_reference(n):
    """(A, B, C, D, G) of polybench/3mm.cpp:17-86; G is the CPU answer."""
    A, B, C, D, E, F, G = (_z(n * n) for _ in range(7))
    _lib.mm3_init(A, B, C, D, n)
    _lib.mm3_cpu(A, B, C, D, E, F, G, n)
    return tuple(x.reshape(n, n) for x in (A, B, C, D, G))


def mm3_inputs(n):
    A, B, C, D = (_z(n * n) for _ in range(4))
    _lib.mm3_init(A, B, C, D, n)
    return tuple(x.reshape(n, n) for x in (A, B, C, D))


def syrk_reference(n):
    """(A, C_in, C_out) of polybench/syrk.cpp:18-51."""
    A, C = _z(n * n), _z(n * n)
    _lib.syrk_init(A, C, n)
    out = C.copy()
    _lib.syrk_cpu(A, out, n)
    return A.reshape(n, n), C.reshape(n, n), out.reshape(n, n)


def syrk_inputs(n):
    A, C = _z(n * n), _z(n * n)
    _lib.syrk_init(A, C, n)
    return A.reshape(n, n), C.reshape(n, n)


def syr2k_reference(n):
    """(A, B, C_in, C_out) of polybench/syr2k.cpp:18-52."""
    A, B, C = _z(n * n), _z(n * n), _z(n * n)
    _lib.syr2k_init(A, B, C, n)
    out = C.copy()
    _lib.syr2k_cpu(A, B, out, n)
    return A.reshape(n, n), B.reshape(n, n), C.reshape(n, n), out.reshape(n, n)


def syr2k_inputs(n):
    A, B, C = _z(n * n), _z(n * n), _z(n * n)
    _lib.syr2k_init(A, B, C, n)
    return A.reshape(n, n), B.reshape(n, n), C.reshape(n, n)


COV_FLOAT_N = np.float32(3214212.01)  # polybench/covariance.cpp:18
CORR_FLOAT_N = 3214212.01              # polybench/correlation.cpp:8
CORR_EPS = 0.005                       # polybench/correlation.cpp:9


def covariance_inputs(n):
    """(n+1) x (n+1) data of polybench/covariance.cpp:20-29 (used 1-based)."""
    data = _z((n + 1) * (n + 1))
    _lib.covariance_init(data, n)
    return data.reshape(n + 1, n + 1)


def covariance_reference(n):
    """(data_in, symmat) of polybench/covariance.cpp:20-61, (n+1) x (n+1), 1-based."""
    data = covariance_inputs(n).copy().reshape(-1)
    d_in = data.copy().reshape(n + 1, n + 1)
    symmat, mean = _z((n + 1) * (n + 1)), _z(n + 1)
    _lib.covariance_cpu(data, symmat, mean, n)
    return d_in, symmat.reshape(n + 1, n + 1)


def correlation_inputs(n):
    """(n+1) x (n+1) data of polybench/correlation.cpp:17-25 (row stride n in the flat array)."""
    data = _z((n + 1) * (n + 1))
    _lib.correlation_init(data, n)
    return data.reshape(n + 1, n + 1)


def correlation_reference(n):
    """(data_in, symmat) of polybench/correlation.cpp:17-70, (n+1) x (n+1), 1-based."""
    data = correlation_inputs(n).copy().reshape(-1)
    d_in = data.copy().reshape(n + 1, n + 1)
    mean, stddev, symmat = _z(n + 1), _z(n + 1), _z((n + 1) * (n + 1))
    _lib.correlation_cpu(data, mean, stddev, symmat, n)
    return d_in, symmat.reshape(n + 1, n + 1)


def matmulchain_identity(n):
    M = _z(n * n)
    _lib.matmulchain_identity(M, n)
    return M.reshape(n, n)


def matmulchain_multiply_int(a, b):
    """runtime/matmulchain.cpp:12-30 with its int accumulator."""
    n = a.shape[0]
    c = _z(n * n)
    _lib.matmulchain_multiply_int(np.ascontiguousarray(a, np.float32).ravel(),
                                  np.ascontiguousarray(b, np.float32).ravel(), c, n)
    return c.reshape(n, n)


def percent_diff(v1, v2):
    return float(_lib.percent_diff(float(v1), float(v2)))


def percent_diff_count(ref, x, threshold=0.05):
    ref = np.ascontiguousarray(ref, np.float32).ravel()
    x = np.ascontiguousarray(x, np.float32).ravel()
    assert ref.size == x.size
    return int(_lib.percent_diff_count(ref, x, ref.size, threshold))


_CLOSED = {"gemm": "gemm_closed", "2mm": "mm2_closed", "3mm": "mm3_closed", "syrk": "syrk_closed",
           "syr2k": "syr2k_closed"}


def closed_form(name, n):
    """FP64 closed form of the reference output for its own inputs at size n."""
    out = _z(n * n, np.float64)
    getattr(_lib, _CLOSED[name])(out, n)
    return out.reshape(n, n)


# ---- FP64 brute force -------------------------------------------------------
def gemm_f64(A, B, C=None, alpha=1.0, beta=0.0, ta=False, tb=False):
    """alpha * op(A) @ op(B) + beta * C in float64 with a sequential k loop."""
    A = np.ascontiguousarray(A, np.float64)
    B = np.ascontiguousarray(B, np.float64)
    m = A.shape[1] if ta else A.shape[0]
    k = A.shape[0] if ta else A.shape[1]
    n = B.shape[0] if tb else B.shape[1]
    out = np.zeros((m, n), np.float64) if C is None else np.array(C, np.float64, copy=True, order="C")
    if C is None:
        beta = 0.0
    _lib.gemm_f64(int(ta), int(tb), m, n, k, alpha, A.ravel(), A.shape[1], B.ravel(), B.shape[1], beta,
                  out.ravel(), n)
    return out


# ---- seeded inputs ----------------------------------------------------------
def keyed_hash(seed, matrix, index):
    return int(_lib.keyed_hash(seed, matrix, index))


def splitmix64_mix(z):
    return int(_lib.splitmix64_mix(z))


def fill_uniform(rows, cols, seed, matrix):
    out = _z(rows * cols)
    _lib.fill_uniform(out, rows, cols, cols, seed, matrix)
    return out.reshape(rows, cols)


def fill_lognormal(rows, cols, seed, matrix, mu=0.0, sigma=2.0):
    out = _z(rows * cols)
    _lib.fill_lognormal(out, rows, cols, cols, seed, matrix, mu, sigma)
    return out.reshape(rows, cols)


# ---- number formats -----------------------------------------------------------
def split_bf16(x):
    x = np.ascontiguousarray(x, np.float32).ravel()
    hi = np.zeros(x.size, np.uint16)
    lo = np.zeros(x.size, np.uint16)
    _lib.split_bf16(x, hi, lo, x.size)
    return hi, lo


def bf16_to_f32(h):
    return (np.asarray(h, np.uint32) << 16).view(np.float32)


def e4m3_rne_sat(x):
    return int(_lib.e4m3_rne_sat(float(x)))


def e4m3_to_f32(q):
    return float(_lib.e4m3_to_f32(int(q)))


def fp8_block_exponent(amax):
    return int(_lib.fp8_block_exponent(float(amax)))


def quantize_fp8_block(x, block=128):
    x = np.ascontiguousarray(x, np.float32)
    rows, cols = x.shape
    q = np.zeros(rows * cols, np.uint8)
    tr, tc = -(-rows // block), -(-cols // block)
    exps = np.zeros(tr * tc, np.int32)
    _lib.quantize_fp8_block(x.ravel(), rows, cols, cols, block, q, exps)
    return q.reshape(rows, cols), exps.reshape(tr, tc)


def rel_frobenius(x, ref):
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.linalg.norm(x - ref) / np.linalg.norm(ref))


# ---- CPU baseline (OpenMP restatement, -O3 -ffast-math) ----------------------
_base = None


def cpu_baseline():
    """The OpenMP restatement library, loaded on first use."""
    global _base
    if _base is None:
        _base = _load("libcpubaseline.so")
        _base.cpu_threads.restype = ctypes.c_int
        _base.cpu_gemm_rows.argtypes = [_F32P, _F32P, _F32P, _I64, ctypes.c_float, ctypes.c_float, _I64, _I64]
        _base.cpu_gemm_rows.restype = None
        _base.cpu_syrk_rows.argtypes = [_F32P, _F32P, _I64, ctypes.c_float, ctypes.c_float, _I64, _I64]
        _base.cpu_syrk_rows.restype = None
        _base.cpu_syr2k_rows.argtypes = [_F32P, _F32P, _F32P, _I64, ctypes.c_float, ctypes.c_float, _I64, _I64]
        _base.cpu_syr2k_rows.restype = None
        _base.cpu_matmul_rows.argtypes = [_F32P, _F32P, _F32P, _I64, _I64, _I64]
        _base.cpu_matmul_rows.restype = None
        _base.cpu_gemm_mnk_rows.argtypes = [_F32P, _F32P, _F32P, _I64, _I64, _I64, ctypes.c_float, ctypes.c_float,
                                            _I64, _I64]
        _base.cpu_gemm_mnk_rows.restype = None
    return _base
