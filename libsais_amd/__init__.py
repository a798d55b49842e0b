"""MI355X suffix-array engine: Python mirror of the libsais 2.10.4 C API.

The product is the C-ABI shared library ``libsais_amd/lib/libsais.so`` (HIP kernels
for gfx950 behind the libsais prototypes in ``include/libsais.h`` and
``include/libsais64.h``, plus the device-level ``include/sagpu.h``).  This module
binds it with ctypes so Python callers and tests use the same names, argument order
and return codes as C callers:

    rc = libsais(T, SA, n, fs, freq)          # reference include/libsais.h:84
    rc = libsais_gsa(T, SA, n, fs, freq)      # reference include/libsais.h:95
    idx = libsais_bwt(T, U, A, n, fs, freq)   # reference include/libsais.h:182
    rc = libsais_plcp(T, SA, PLCP, n)         # reference include/libsais.h:368
    rc = libsais_plcp_gsa(T, SA, PLCP, n)     # reference include/libsais.h:378
    rc = libsais_lcp(PLCP, SA, LCP, n)        # reference include/libsais.h:398
    libsais64*, same shapes with int64        # reference include/libsais64.h

Arrays are numpy arrays (uint8 text, int32/int64 index arrays) or None, which is
passed as NULL.  There is no CPU fallback: if the library is missing the import of
any entry point raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
# LIBSAIS_AMD_LIB selects another build of the same library (the host-sanitizer
# builds under lib/san and lib/tsan, used by tests/test_sanitizers.py)
LIB_PATH = os.environ.get("LIBSAIS_AMD_LIB") or os.path.join(LIB_DIR, "libsais.so")
GEN_PATH = os.path.join(LIB_DIR, "libsagen.so")

SAGPU_FLAG_GSA = 1

_lib = None


class EngineNotBuilt(RuntimeError):
    pass


def load() -> ctypes.CDLL:
    """Load the engine library (RTLD_LOCAL, so a reference libsais can coexist)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise EngineNotBuilt(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "or `make -C libsais_amd/csrc`")
        lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_LOCAL)
        _declare(lib)
        _lib = lib
    return _lib


class SagpuStats(ctypes.Structure):
    _fields_ = [
        ("rounds", ctypes.c_int32),
        ("initial_symbols", ctypes.c_int32),
        ("bits_per_symbol", ctypes.c_int32),
        ("sort_passes", ctypes.c_int32),
        ("active", ctypes.c_int64 * 64),
        ("sorted_elems", ctypes.c_int64),
        ("sorted_bytes", ctypes.c_int64),
        ("big_rounds", ctypes.c_int32),
        ("big_elems", ctypes.c_int64),
        ("ms_total", ctypes.c_double),
        ("ms_onesweep", ctypes.c_double),
        ("onesweep_launches", ctypes.c_int32),
        ("ms_initial", ctypes.c_double),
        ("ms_rounds", ctypes.c_double),
        ("ms_bwt", ctypes.c_double),
        ("ms_plcp", ctypes.c_double),
        ("ms_lcp", ctypes.c_double),
        ("ms_onesweep_u32", ctypes.c_double),
        ("onesweep_launches_u32", ctypes.c_int32),
        ("sorted_bytes_u32", ctypes.c_int64),
        ("ms_plcp_pairs_pass", ctypes.c_double),
        ("plcp_pairs_launches", ctypes.c_int32),
        ("device_allocs", ctypes.c_int64),
        ("peak_bytes", ctypes.c_int64),
        ("kernel_ms", ctypes.c_double * 16),
        ("kernel_launches", ctypes.c_int32 * 16),
        ("kernel_bytes", ctypes.c_int64 * 16),
        ("round_graphs", ctypes.c_int32),
        ("round_graph_launches", ctypes.c_int32),
        ("big_pivot_rounds", ctypes.c_int32),
        ("big_pivot_elems", ctypes.c_int64),
    ]


_vp = ctypes.c_void_p
_i32, _i64 = ctypes.c_int32, ctypes.c_int64

# (name, restype, argtypes) of every symbol the C headers declare.
ABI_32 = [
    ("libsais", _i32, [_vp, _vp, _i32, _i32, _vp]),
    ("libsais_gsa", _i32, [_vp, _vp, _i32, _i32, _vp]),
    ("libsais_bwt", _i32, [_vp, _vp, _vp, _i32, _i32, _vp]),
    ("libsais_bwt_aux", _i32, [_vp, _vp, _vp, _i32, _i32, _vp, _i32, _vp]),
    ("libsais_unbwt", _i32, [_vp, _vp, _vp, _i32, _vp, _i32]),
    ("libsais_unbwt_aux", _i32, [_vp, _vp, _vp, _i32, _vp, _i32, _vp]),
    ("libsais_plcp", _i32, [_vp, _vp, _vp, _i32]),
    ("libsais_plcp_gsa", _i32, [_vp, _vp, _vp, _i32]),
    ("libsais_lcp", _i32, [_vp, _vp, _vp, _i32]),
]
ABI_32 += [
    ("libsais_int", _i32, [_vp, _vp, _i32, _i32, _i32]),
    ("libsais_plcp_int", _i32, [_vp, _vp, _vp, _i32]),
    ("libsais_create_ctx", _vp, []),
    ("libsais_free_ctx", None, [_vp]),
    ("libsais_ctx", _i32, [_vp, _vp, _vp, _i32, _i32, _vp]),
    ("libsais_gsa_ctx", _i32, [_vp, _vp, _vp, _i32, _i32, _vp]),
    ("libsais_bwt_ctx", _i32, [_vp, _vp, _vp, _vp, _i32, _i32, _vp]),
    ("libsais_bwt_aux_ctx", _i32, [_vp, _vp, _vp, _vp, _i32, _i32, _vp, _i32, _vp]),
    ("libsais_unbwt_create_ctx", _vp, []),
    ("libsais_unbwt_free_ctx", None, [_vp]),
    ("libsais_unbwt_ctx", _i32, [_vp, _vp, _vp, _vp, _i32, _vp, _i32]),
    ("libsais_unbwt_aux_ctx", _i32, [_vp, _vp, _vp, _vp, _i32, _vp, _i32, _vp]),
    ("libsais_create_ctx_omp", _vp, [_i32]),
    ("libsais_omp", _i32, [_vp, _vp, _i32, _i32, _vp, _i32]),
    ("libsais_gsa_omp", _i32, [_vp, _vp, _i32, _i32, _vp, _i32]),
    ("libsais_int_omp", _i32, [_vp, _vp, _i32, _i32, _i32, _i32]),
    ("libsais_bwt_omp", _i32, [_vp, _vp, _vp, _i32, _i32, _vp, _i32]),
    ("libsais_bwt_aux_omp", _i32, [_vp, _vp, _vp, _i32, _i32, _vp, _i32, _vp, _i32]),
    ("libsais_unbwt_create_ctx_omp", _vp, [_i32]),
    ("libsais_unbwt_omp", _i32, [_vp, _vp, _vp, _i32, _vp, _i32, _i32]),
    ("libsais_unbwt_aux_omp", _i32, [_vp, _vp, _vp, _i32, _vp, _i32, _vp, _i32]),
    ("libsais_plcp_omp", _i32, [_vp, _vp, _vp, _i32, _i32]),
    ("libsais_plcp_gsa_omp", _i32, [_vp, _vp, _vp, _i32, _i32]),
    ("libsais_plcp_int_omp", _i32, [_vp, _vp, _vp, _i32, _i32]),
    ("libsais_lcp_omp", _i32, [_vp, _vp, _vp, _i32, _i32]),
]
ABI_64 = [
    ("libsais64", _i64, [_vp, _vp, _i64, _i64, _vp]),
    ("libsais64_gsa", _i64, [_vp, _vp, _i64, _i64, _vp]),
    ("libsais64_bwt", _i64, [_vp, _vp, _vp, _i64, _i64, _vp]),
    ("libsais64_bwt_aux", _i64, [_vp, _vp, _vp, _i64, _i64, _vp, _i64, _vp]),
    ("libsais64_unbwt", _i64, [_vp, _vp, _vp, _i64, _vp, _i64]),
    ("libsais64_unbwt_aux", _i64, [_vp, _vp, _vp, _i64, _vp, _i64, _vp]),
    ("libsais64_plcp", _i64, [_vp, _vp, _vp, _i64]),
    ("libsais64_plcp_gsa", _i64, [_vp, _vp, _vp, _i64]),
    ("libsais64_lcp", _i64, [_vp, _vp, _vp, _i64]),
]
ABI_64 += [
    ("libsais64_long", _i64, [_vp, _vp, _i64, _i64, _i64]),
    ("libsais64_omp", _i64, [_vp, _vp, _i64, _i64, _vp, _i64]),
    ("libsais64_gsa_omp", _i64, [_vp, _vp, _i64, _i64, _vp, _i64]),
    ("libsais64_long_omp", _i64, [_vp, _vp, _i64, _i64, _i64, _i64]),
    ("libsais64_bwt_omp", _i64, [_vp, _vp, _vp, _i64, _i64, _vp, _i64]),
    ("libsais64_bwt_aux_omp", _i64, [_vp, _vp, _vp, _i64, _i64, _vp, _i64, _vp, _i64]),
    ("libsais64_unbwt_omp", _i64, [_vp, _vp, _vp, _i64, _vp, _i64, _i64]),
    ("libsais64_unbwt_aux_omp", _i64, [_vp, _vp, _vp, _i64, _vp, _i64, _vp, _i64]),
    ("libsais64_plcp_omp", _i64, [_vp, _vp, _vp, _i64, _i64]),
    ("libsais64_plcp_gsa_omp", _i64, [_vp, _vp, _vp, _i64, _i64]),
    ("libsais64_lcp_omp", _i64, [_vp, _vp, _vp, _i64, _i64]),
]
# 16-bit alphabet (libsais16.h / libsais16x64.h)
ABI_16 = [
    ("libsais16_create_ctx", _vp, []),
    ("libsais16_free_ctx", None, [_vp]),
    ("libsais16", _i32, [_vp, _vp, _i32, _i32, _vp]),
    ("libsais16_gsa", _i32, [_vp, _vp, _i32, _i32, _vp]),
    ("libsais16_int", _i32, [_vp, _vp, _i32, _i32, _i32]),
    ("libsais16_ctx", _i32, [_vp, _vp, _vp, _i32, _i32, _vp]),
    ("libsais16_gsa_ctx", _i32, [_vp, _vp, _vp, _i32, _i32, _vp]),
    ("libsais16_bwt", _i32, [_vp, _vp, _vp, _i32, _i32, _vp]),
    ("libsais16_bwt_aux", _i32, [_vp, _vp, _vp, _i32, _i32, _vp, _i32, _vp]),
    ("libsais16_bwt_ctx", _i32, [_vp, _vp, _vp, _vp, _i32, _i32, _vp]),
    ("libsais16_bwt_aux_ctx", _i32, [_vp, _vp, _vp, _vp, _i32, _i32, _vp, _i32, _vp]),
    ("libsais16_unbwt_create_ctx", _vp, []),
    ("libsais16_unbwt_free_ctx", None, [_vp]),
    ("libsais16_unbwt", _i32, [_vp, _vp, _vp, _i32, _vp, _i32]),
    ("libsais16_unbwt_ctx", _i32, [_vp, _vp, _vp, _vp, _i32, _vp, _i32]),
    ("libsais16_unbwt_aux", _i32, [_vp, _vp, _vp, _i32, _vp, _i32, _vp]),
    ("libsais16_unbwt_aux_ctx", _i32, [_vp, _vp, _vp, _vp, _i32, _vp, _i32, _vp]),
    ("libsais16_plcp", _i32, [_vp, _vp, _vp, _i32]),
    ("libsais16_plcp_gsa", _i32, [_vp, _vp, _vp, _i32]),
    ("libsais16_lcp", _i32, [_vp, _vp, _vp, _i32]),
    ("libsais16_create_ctx_omp", _vp, [_i32]),
    ("libsais16_omp", _i32, [_vp, _vp, _i32, _i32, _vp, _i32]),
    ("libsais16_gsa_omp", _i32, [_vp, _vp, _i32, _i32, _vp, _i32]),
    ("libsais16_int_omp", _i32, [_vp, _vp, _i32, _i32, _i32, _i32]),
    ("libsais16_bwt_omp", _i32, [_vp, _vp, _vp, _i32, _i32, _vp, _i32]),
    ("libsais16_bwt_aux_omp", _i32, [_vp, _vp, _vp, _i32, _i32, _vp, _i32, _vp, _i32]),
    ("libsais16_unbwt_create_ctx_omp", _vp, [_i32]),
    ("libsais16_unbwt_omp", _i32, [_vp, _vp, _vp, _i32, _vp, _i32, _i32]),
    ("libsais16_unbwt_aux_omp", _i32, [_vp, _vp, _vp, _i32, _vp, _i32, _vp, _i32]),
    ("libsais16_plcp_omp", _i32, [_vp, _vp, _vp, _i32, _i32]),
    ("libsais16_plcp_gsa_omp", _i32, [_vp, _vp, _vp, _i32, _i32]),
    ("libsais16_lcp_omp", _i32, [_vp, _vp, _vp, _i32, _i32]),
]
ABI_16X64 = [
    ("libsais16x64", _i64, [_vp, _vp, _i64, _i64, _vp]),
    ("libsais16x64_gsa", _i64, [_vp, _vp, _i64, _i64, _vp]),
    ("libsais16x64_long", _i64, [_vp, _vp, _i64, _i64, _i64]),
    ("libsais16x64_bwt", _i64, [_vp, _vp, _vp, _i64, _i64, _vp]),
    ("libsais16x64_bwt_aux", _i64, [_vp, _vp, _vp, _i64, _i64, _vp, _i64, _vp]),
    ("libsais16x64_unbwt", _i64, [_vp, _vp, _vp, _i64, _vp, _i64]),
    ("libsais16x64_unbwt_aux", _i64, [_vp, _vp, _vp, _i64, _vp, _i64, _vp]),
    ("libsais16x64_plcp", _i64, [_vp, _vp, _vp, _i64]),
    ("libsais16x64_plcp_gsa", _i64, [_vp, _vp, _vp, _i64]),
    ("libsais16x64_lcp", _i64, [_vp, _vp, _vp, _i64]),
    ("libsais16x64_omp", _i64, [_vp, _vp, _i64, _i64, _vp, _i64]),
    ("libsais16x64_gsa_omp", _i64, [_vp, _vp, _i64, _i64, _vp, _i64]),
    ("libsais16x64_long_omp", _i64, [_vp, _vp, _i64, _i64, _i64, _i64]),
    ("libsais16x64_bwt_omp", _i64, [_vp, _vp, _vp, _i64, _i64, _vp, _i64]),
    ("libsais16x64_bwt_aux_omp", _i64, [_vp, _vp, _vp, _i64, _i64, _vp, _i64, _vp, _i64]),
    ("libsais16x64_unbwt_omp", _i64, [_vp, _vp, _vp, _i64, _vp, _i64, _i64]),
    ("libsais16x64_unbwt_aux_omp", _i64, [_vp, _vp, _vp, _i64, _vp, _i64, _vp, _i64]),
    ("libsais16x64_plcp_omp", _i64, [_vp, _vp, _vp, _i64, _i64]),
    ("libsais16x64_plcp_gsa_omp", _i64, [_vp, _vp, _vp, _i64, _i64]),
    ("libsais16x64_lcp_omp", _i64, [_vp, _vp, _vp, _i64, _i64]),
]
ABI_SAGPU = [
    ("sagpu_init", _i32, []),
    ("sagpu_shutdown", None, []),
    ("sagpu_last_error", ctypes.c_char_p, []),
    ("sagpu_malloc", _vp, [ctypes.c_size_t]),
    ("sagpu_free", None, [_vp]),
    ("sagpu_memcpy_h2d", _i32, [_vp, _vp, ctypes.c_size_t]),
    ("sagpu_memcpy_d2h", _i32, [_vp, _vp, ctypes.c_size_t]),
    ("sagpu_host_memcpy", _i32, [_vp, _vp, ctypes.c_size_t, _i32]),
    ("sagpu_memset", _i32, [_vp, ctypes.c_int, ctypes.c_size_t]),
    ("sagpu_memcpy_d2d", _i32, [_vp, _vp, ctypes.c_size_t]),
    ("sagpu_synchronize", _i32, []),
    ("sagpu_set_memory_limit", None, [ctypes.c_uint64]),
    ("sagpu_trim", None, []),
    ("sagpu_build32", _i64, [_vp, _i64, _i32, _vp, _vp, _vp, _vp, _vp]),
    ("sagpu_build32_noalloc", _i64, [_vp, _i64, _i32, _vp, _vp, _vp, _vp, _vp]),
    ("sagpu_gsa_batch32", _i32, [_vp, _i64, _vp, _i32, _vp, _vp]),
    ("sagpu_run_job", _i32, [_vp]),
    ("sagpu_server_create", _vp, [_i32, ctypes.c_uint64, _i64]),
    ("sagpu_server_submit", _i32, [_vp, _vp]),
    ("sagpu_server_wait", _i32, [_vp, _vp]),
    ("sagpu_server_drain", _i32, [_vp]),
    ("sagpu_server_destroy", None, [_vp]),
    ("sagpu_clock", ctypes.c_double, []),
    ("sagpu_plcp32", _i32, [_vp, _vp, _vp, _i64, _i32]),
    ("sagpu_lcp32", _i32, [_vp, _vp, _vp, _i64]),
    ("sagpu_sort_pairs_u32", _i32, [_vp, _vp, _i64, _i32, _i32]),
    ("sagpu_sort_pairs_u64", _i32, [_vp, _vp, _i64, _i32, _i32]),
    ("sagpu_set_option", _i32, [ctypes.c_char_p, _i64]),
    ("sagpu_sort_bench", _i32, [_i64, _i32, _i32, _i32, ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_double)]),
    ("sagpu_set_timing", None, [_i32]),
    ("sagpu_get_stats", _i32, [ctypes.POINTER(SagpuStats)]),
]
ABI = ABI_32 + ABI_64 + ABI_16 + ABI_16X64 + ABI_SAGPU


def _declare(lib):
    for name, res, args in ABI:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("arrays must be C-contiguous")
        return a.ctypes.data
    if isinstance(a, (bytes, bytearray)):
        raise TypeError("pass text as a numpy uint8 array")
    return a  # raw device/host address (int)


def _mk(name):
    def call(*args):
        lib = load()
        conv = [_ptr(a) if (a is None or isinstance(a, np.ndarray)) else a for a in args]
        return getattr(lib, name)(*conv)
    call.__name__ = name
    call.__doc__ = f"ctypes binding of `{name}` (same arguments and return code as the C function)."
    return call


libsais = _mk("libsais")
libsais_gsa = _mk("libsais_gsa")
libsais_bwt = _mk("libsais_bwt")
libsais_bwt_aux = _mk("libsais_bwt_aux")
libsais_unbwt = _mk("libsais_unbwt")
libsais_unbwt_aux = _mk("libsais_unbwt_aux")
libsais_plcp = _mk("libsais_plcp")
libsais_plcp_gsa = _mk("libsais_plcp_gsa")
libsais_lcp = _mk("libsais_lcp")
libsais64 = _mk("libsais64")
libsais64_gsa = _mk("libsais64_gsa")
libsais64_bwt = _mk("libsais64_bwt")
libsais64_bwt_aux = _mk("libsais64_bwt_aux")
libsais64_unbwt = _mk("libsais64_unbwt")
libsais64_unbwt_aux = _mk("libsais64_unbwt_aux")
libsais64_plcp = _mk("libsais64_plcp")
libsais64_plcp_gsa = _mk("libsais64_plcp_gsa")
libsais64_lcp = _mk("libsais64_lcp")

# every other exported entry point under its C name (contexts, OpenMP forms,
# integer alphabets, 16-bit alphabet)
for _name, _res, _args in ABI_32 + ABI_64 + ABI_16 + ABI_16X64:
    if _name not in globals():
        globals()[_name] = _mk(_name)


def last_error() -> str:
    e = load().sagpu_last_error()
    return e.decode() if e else ""


# ------------------------------------------------------------ convenience layer

def suffix_array(T: np.ndarray, gsa: bool = False, freq: bool = False):
    """SA of T (int32).  Returns SA, or (SA, freq) when freq=True."""
    T = np.ascontiguousarray(T, dtype=np.uint8)
    n = len(T)
    SA = np.zeros(max(n, 1), dtype=np.int32)
    fr = np.zeros(256, dtype=np.int32) if freq else None
    rc = (libsais_gsa if gsa else libsais)(T, SA, n, 0, fr)
    if rc != 0:
        raise RuntimeError(f"libsais{'_gsa' if gsa else ''} returned {rc}: {last_error()}")
    return (SA[:n], fr) if freq else SA[:n]


def bwt(T: np.ndarray):
    """(U, primary) exactly as libsais_bwt."""
    T = np.ascontiguousarray(T, dtype=np.uint8)
    n = len(T)
    U = np.zeros(max(n, 1), dtype=np.uint8)
    A = np.zeros(max(n, 1), dtype=np.int32)
    idx = libsais_bwt(T, U, A, n, 0, None)
    if idx < 0:
        raise RuntimeError(f"libsais_bwt returned {idx}: {last_error()}")
    return U[:n], idx


def bwt_aux(T: np.ndarray, r: int):
    """(U, I) exactly as libsais_bwt_aux."""
    T = np.ascontiguousarray(T, dtype=np.uint8)
    n = len(T)
    U = np.zeros(max(n, 1), dtype=np.uint8)
    A = np.zeros(1, dtype=np.int32)
    I = np.zeros((n - 1) // r + 1 if n > 0 else 1, dtype=np.int32)
    rc = libsais_bwt_aux(T, U, A, n, 0, None, r, I)
    if rc != 0:
        raise RuntimeError(f"libsais_bwt_aux returned {rc}: {last_error()}")
    return U[:n], I


def unbwt(U: np.ndarray, primary: int):
    """Original string from a BWT and its primary index (libsais_unbwt)."""
    U = np.ascontiguousarray(U, dtype=np.uint8)
    n = len(U)
    out = np.zeros(max(n, 1), dtype=np.uint8)
    A = np.zeros(1, dtype=np.int32)
    rc = libsais_unbwt(U, out, A, n, None, primary)
    if rc != 0:
        raise RuntimeError(f"libsais_unbwt returned {rc}: {last_error()}")
    return out[:n]


def plcp(T: np.ndarray, SA: np.ndarray, gsa: bool = False):
    T = np.ascontiguousarray(T, dtype=np.uint8)
    SA = np.ascontiguousarray(SA, dtype=np.int32)
    n = len(T)
    P = np.zeros(max(n, 1), dtype=np.int32)
    rc = (libsais_plcp_gsa if gsa else libsais_plcp)(T, SA, P, n)
    if rc != 0:
        raise RuntimeError(f"libsais_plcp returned {rc}: {last_error()}")
    return P[:n]


def lcp(PLCP: np.ndarray, SA: np.ndarray):
    PLCP = np.ascontiguousarray(PLCP, dtype=np.int32)
    SA = np.ascontiguousarray(SA, dtype=np.int32)
    n = len(SA)
    L = np.zeros(max(n, 1), dtype=np.int32)
    rc = libsais_lcp(PLCP, SA, L, n)
    if rc != 0:
        raise RuntimeError(f"libsais_lcp returned {rc}: {last_error()}")
    return L[:n]


def stats() -> SagpuStats:
    s = SagpuStats()
    load().sagpu_get_stats(ctypes.byref(s))
    return s
