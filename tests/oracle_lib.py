"""ctypes access to the checkers under oracle/ (test infrastructure only).

liboracle.so is the C restatement (oracle/oracle.c); libsais_ref.so is the
reference libsais 2.10.4 compiled from its own sources by oracle/Makefile.  The
product library never loads either.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.environ.get("LIBSAIS_ORACLE_LIB") or os.path.join(ROOT, "oracle", "lib", "liboracle.so")
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "libsais_ref.so")

_o = None
_r = None
_vp, _i64, _i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32


def oracle():
    global _o
    if _o is None:
        o = ctypes.CDLL(ORACLE_PATH, mode=ctypes.RTLD_LOCAL)
        o.oracle_sa32.argtypes = [_vp, _vp, _i64, ctypes.c_int]
        o.oracle_sa32.restype = ctypes.c_int
        o.oracle_sa.argtypes = [_vp, _vp, _i64, ctypes.c_int]
        o.oracle_sa.restype = ctypes.c_int
        o.oracle_check_sa32.argtypes = [_vp, _vp, _i64, ctypes.c_int]
        o.oracle_check_sa32.restype = ctypes.c_int
        o.oracle_check_sa.argtypes = [_vp, _vp, _i64, ctypes.c_int]
        o.oracle_check_sa.restype = ctypes.c_int
        o.oracle_bwt32.argtypes = [_vp, _vp, _vp, _i64]
        o.oracle_bwt32.restype = _i64
        o.oracle_plcp32.argtypes = [_vp, _vp, _vp, _i64, ctypes.c_int]
        o.oracle_plcp32.restype = None
        o.oracle_lcp32.argtypes = [_vp, _vp, _vp, _i64]
        o.oracle_lcp32.restype = None
        o.oracle_freq.argtypes = [_vp, _i64, _vp]
        o.oracle_freq.restype = None
        o.oracle_gsa_valid.argtypes = [_vp, _i64]
        o.oracle_gsa_valid.restype = ctypes.c_int
        _o = o
    return _o


def have_ref():
    return os.path.exists(REF_PATH)


def ref():
    global _r
    if _r is None:
        r = ctypes.CDLL(REF_PATH, mode=ctypes.RTLD_LOCAL)
        for name in ("libsais", "libsais_gsa"):
            getattr(r, name).argtypes = [_vp, _vp, _i32, _i32, _vp]
            getattr(r, name).restype = _i32
        r.libsais_bwt.argtypes = [_vp, _vp, _vp, _i32, _i32, _vp]
        r.libsais_bwt.restype = _i32
        for name in ("libsais_plcp", "libsais_plcp_gsa", "libsais_lcp"):
            getattr(r, name).argtypes = [_vp, _vp, _vp, _i32]
            getattr(r, name).restype = _i32
        for name in ("libsais_omp", "libsais_gsa_omp"):
            getattr(r, name).argtypes = [_vp, _vp, _i32, _i32, _vp, _i32]
            getattr(r, name).restype = _i32
        r.libsais_bwt_omp.argtypes = [_vp, _vp, _vp, _i32, _i32, _vp, _i32]
        r.libsais_bwt_omp.restype = _i32
        for name in ("libsais_plcp_omp", "libsais_plcp_gsa_omp", "libsais_lcp_omp"):
            getattr(r, name).argtypes = [_vp, _vp, _vp, _i32, _i32]
            getattr(r, name).restype = _i32
        for name in ("libsais64", "libsais64_gsa"):
            getattr(r, name).argtypes = [_vp, _vp, _i64, _i64, _vp]
            getattr(r, name).restype = _i64
        r.libsais64_bwt.argtypes = [_vp, _vp, _vp, _i64, _i64, _vp]
        r.libsais64_bwt.restype = _i64
        for name in ("libsais64_plcp", "libsais64_plcp_gsa", "libsais64_lcp"):
            getattr(r, name).argtypes = [_vp, _vp, _vp, _i64]
            getattr(r, name).restype = _i64
        r.libsais_bwt_aux.argtypes = [_vp, _vp, _vp, _i32, _i32, _vp, _i32, _vp]
        r.libsais_bwt_aux.restype = _i32
        r.libsais_unbwt.argtypes = [_vp, _vp, _vp, _i32, _vp, _i32]
        r.libsais_unbwt.restype = _i32
        r.libsais_unbwt_aux.argtypes = [_vp, _vp, _vp, _i32, _vp, _i32, _vp]
        r.libsais_unbwt_aux.restype = _i32
        r.libsais_int.argtypes = [_vp, _vp, _i32, _i32, _i32]
        r.libsais_int.restype = _i32
        r.libsais_plcp_int.argtypes = [_vp, _vp, _vp, _i32]
        r.libsais_plcp_int.restype = _i32
        r.libsais64_long.argtypes = [_vp, _vp, _i64, _i64, _i64]
        r.libsais64_long.restype = _i64
        for name in ("libsais16", "libsais16_gsa"):
            getattr(r, name).argtypes = [_vp, _vp, _i32, _i32, _vp]
            getattr(r, name).restype = _i32
        r.libsais16_bwt.argtypes = [_vp, _vp, _vp, _i32, _i32, _vp]
        r.libsais16_bwt.restype = _i32
        r.libsais16_bwt_aux.argtypes = [_vp, _vp, _vp, _i32, _i32, _vp, _i32, _vp]
        r.libsais16_bwt_aux.restype = _i32
        r.libsais16_unbwt.argtypes = [_vp, _vp, _vp, _i32, _vp, _i32]
        r.libsais16_unbwt.restype = _i32
        for name in ("libsais16_plcp", "libsais16_plcp_gsa"):
            getattr(r, name).argtypes = [_vp, _vp, _vp, _i32]
            getattr(r, name).restype = _i32
        r.libsais64_omp.argtypes = [_vp, _vp, _i64, _i64, _vp, _i64]
        r.libsais64_omp.restype = _i64
        for name in ("libsais64_plcp_omp", "libsais64_lcp_omp"):
            getattr(r, name).argtypes = [_vp, _vp, _vp, _i64, _i64]
            getattr(r, name).restype = _i64
        _r = r
    return _r


def _p(a):
    return a.ctypes.data


# ---------------------------------------------------------------- oracle
def sa(T, gsa=False):
    n = len(T)
    SA = np.zeros(max(n, 1), np.int32)
    rc = oracle().oracle_sa32(_p(T), _p(SA), n, int(gsa))
    return rc, SA[:n]


def check_sa(T, SA, gsa=False):
    SA = np.ascontiguousarray(SA, dtype=np.int32)
    return bool(oracle().oracle_check_sa32(_p(T), _p(SA), len(T), int(gsa)))


def bwt(T, SA):
    n = len(T)
    U = np.zeros(max(n, 1), np.uint8)
    SA = np.ascontiguousarray(SA, dtype=np.int32)
    idx = oracle().oracle_bwt32(_p(T), _p(SA), _p(U), n)
    return U[:n], idx


def plcp(T, SA, gsa=False):
    n = len(T)
    P = np.zeros(max(n, 1), np.int32)
    SA = np.ascontiguousarray(SA, dtype=np.int32)
    oracle().oracle_plcp32(_p(T), _p(SA), _p(P), n, int(gsa))
    return P[:n]


def lcp(P, SA):
    n = len(SA)
    L = np.zeros(max(n, 1), np.int32)
    P = np.ascontiguousarray(P, np.int32)   # converted copies must outlive the call
    SA = np.ascontiguousarray(SA, np.int32)
    oracle().oracle_lcp32(_p(P), _p(SA), _p(L), n)
    return L[:n]


def freq(T):
    f = np.zeros(256, np.int64)
    oracle().oracle_freq(_p(T), len(T), _p(f))
    return f


def gsa_valid(T):
    return bool(oracle().oracle_gsa_valid(_p(T), len(T)))


def gsa_valid16(T):
    o = oracle()
    o.oracle_gsa_valid16.argtypes = [_vp, _i64]
    o.oracle_gsa_valid16.restype = ctypes.c_int
    T = np.ascontiguousarray(T, np.uint16)
    return bool(o.oracle_gsa_valid16(_p(T), len(T)))


# ---------------------------------------------------------------- reference
def ref_sa(T, gsa=False, threads=0):
    n = len(T)
    SA = np.zeros(max(n, 1), np.int32)
    f = np.zeros(256, np.int32)
    r = ref()
    if threads:
        rc = (r.libsais_gsa_omp if gsa else r.libsais_omp)(_p(T), _p(SA), n, 0, _p(f), threads)
    else:
        rc = (r.libsais_gsa if gsa else r.libsais)(_p(T), _p(SA), n, 0, _p(f))
    return rc, SA[:n], f


def ref_bwt(T, threads=0):
    n = len(T)
    U = np.zeros(max(n, 1), np.uint8)
    A = np.zeros(max(n, 1), np.int32)
    r = ref()
    if threads:
        idx = r.libsais_bwt_omp(_p(T), _p(U), _p(A), n, 0, None, threads)
    else:
        idx = r.libsais_bwt(_p(T), _p(U), _p(A), n, 0, None)
    return U[:n], idx


def ref_plcp(T, SA, gsa=False, threads=0):
    n = len(T)
    P = np.zeros(max(n, 1), np.int32)
    SA = np.ascontiguousarray(SA, dtype=np.int32)
    r = ref()
    if threads:
        (r.libsais_plcp_gsa_omp if gsa else r.libsais_plcp_omp)(_p(T), _p(SA), _p(P), n, threads)
    else:
        (r.libsais_plcp_gsa if gsa else r.libsais_plcp)(_p(T), _p(SA), _p(P), n)
    return P[:n]


def ref_lcp(P, SA, threads=0):
    n = len(SA)
    L = np.zeros(max(n, 1), np.int32)
    r = ref()
    if threads:
        r.libsais_lcp_omp(_p(P), _p(SA), _p(L), n, threads)
    else:
        r.libsais_lcp(_p(P), _p(SA), _p(L), n)
    return L[:n]


# ------------------------------------------------------- 64-bit (n > INT32_MAX)
def bwt64(T, SA):
    """Oracle BWT from an int64 SA: (U, primary)."""
    n = len(T)
    U = np.zeros(max(n, 1), np.uint8)
    o = oracle()
    o.oracle_bwt.argtypes = [_vp, _vp, _vp, _i64]
    o.oracle_bwt.restype = _i64
    idx = o.oracle_bwt(_p(T), _p(SA), _p(U), n)
    return U[:n], idx


def ref64_sa(T, threads):
    n = len(T)
    SA = np.empty(max(n, 1), np.int64)
    f = np.zeros(256, np.int64)
    rc = ref().libsais64_omp(_p(T), _p(SA), n, 0, _p(f), threads)
    return rc, SA[:n], f


def ref64_plcp(T, SA, threads):
    n = len(T)
    P = np.empty(max(n, 1), np.int64)
    rc = ref().libsais64_plcp_omp(_p(T), _p(SA), _p(P), n, threads)
    return rc, P[:n]


def ref64_lcp(P, SA, threads):
    n = len(SA)
    L = np.empty(max(n, 1), np.int64)
    rc = ref().libsais64_lcp_omp(_p(P), _p(SA), _p(L), n, threads)
    return rc, L[:n]


def chunked_equal(a, b, chunk=1 << 26):
    """Array equality without an n-sized temporary."""
    if len(a) != len(b):
        return False
    for o in range(0, len(a), chunk):
        if not np.array_equal(a[o:o + chunk], b[o:o + chunk]):
            return False
    return True


# ------------------------------------------------------------ BWT aux / inverse
def bwt_aux(T, SA, r):
    """Oracle (U, I) of libsais_bwt_aux from a suffix array."""
    n = len(T)
    o = oracle()
    o.oracle_bwt_aux.argtypes = [_vp, _vp, _vp, _i64, _i64, _vp]
    o.oracle_bwt_aux.restype = ctypes.c_int
    U = np.zeros(max(n, 1), np.uint8)
    I = np.zeros((n - 1) // r + 1 if n > 0 else 1, np.int64)
    S64 = np.ascontiguousarray(SA, np.int64)
    o.oracle_bwt_aux(_p(T), _p(S64), _p(U), n, r, _p(I))
    return U[:n], I


def unbwt(U, p):
    """Oracle inverse BWT of (U, primary index)."""
    n = len(U)
    o = oracle()
    o.oracle_unbwt.argtypes = [_vp, _vp, _i64, _i64]
    o.oracle_unbwt.restype = ctypes.c_int
    out = np.zeros(max(n, 1), np.uint8)
    assert o.oracle_unbwt(_p(U), _p(out), n, p) == 0
    return out[:n]


def ref_bwt_aux(T, r):
    n = len(T)
    U = np.zeros(max(n, 1), np.uint8)
    A = np.zeros(max(n, 1), np.int32)
    I = np.zeros((n - 1) // r + 1 if n > 0 else 1, np.int32)
    rc = ref().libsais_bwt_aux(_p(T), _p(U), _p(A), n, 0, None, r, _p(I))
    return rc, U[:n], I


def ref_unbwt(U, p):
    n = len(U)
    out = np.zeros(max(n, 1), np.uint8)
    A = np.zeros(n + 1, np.int32)
    rc = ref().libsais_unbwt(_p(U), _p(out), _p(A), n, None, p)
    return rc, out[:n]


def ref_unbwt_aux(U, r, I):
    n = len(U)
    out = np.zeros(max(n, 1), np.uint8)
    A = np.zeros(n + 1, np.int32)
    I32 = np.ascontiguousarray(I, np.int32)
    rc = ref().libsais_unbwt_aux(_p(U), _p(out), _p(A), n, None, r, _p(I32))
    return rc, out[:n]


# ------------------------------------------------------------ integer alphabet
def sa_int(T):
    """Oracle SA of an integer string (int64 result)."""
    T64 = np.ascontiguousarray(T, np.int64)
    n = len(T64)
    SA = np.zeros(max(n, 1), np.int64)
    o = oracle()
    o.oracle_sa_int.argtypes = [_vp, _vp, _i64]
    o.oracle_sa_int.restype = ctypes.c_int
    assert o.oracle_sa_int(_p(T64), _p(SA), n) == 0
    return SA[:n]


def plcp_int(T, SA):
    T64 = np.ascontiguousarray(T, np.int64)
    S64 = np.ascontiguousarray(SA, np.int64)
    n = len(T64)
    P = np.zeros(max(n, 1), np.int64)
    o = oracle()
    o.oracle_plcp_int.argtypes = [_vp, _vp, _vp, _i64]
    o.oracle_plcp_int.restype = None
    o.oracle_plcp_int(_p(T64), _p(S64), _p(P), n)
    return P[:n]


def ref_sa_int(T, k):
    T32 = np.ascontiguousarray(T, np.int32).copy()
    n = len(T32)
    SA = np.zeros(max(n, 1), np.int32)
    rc = ref().libsais_int(_p(T32), _p(SA), n, k, 0)
    return rc, SA[:n], T32


def ref_sa_long(T, k):
    T64 = np.ascontiguousarray(T, np.int64).copy()
    n = len(T64)
    SA = np.zeros(max(n, 1), np.int64)
    rc = ref().libsais64_long(_p(T64), _p(SA), n, k, 0)
    return rc, SA[:n], T64


def ref_plcp_int(T, SA):
    T32 = np.ascontiguousarray(T, np.int32)
    S32 = np.ascontiguousarray(SA, np.int32)
    n = len(T32)
    P = np.zeros(max(n, 1), np.int32)
    rc = ref().libsais_plcp_int(_p(T32), _p(S32), _p(P), n)
    return rc, P[:n]


# ------------------------------------------------------------ 16-bit alphabet
def _o16():
    o = oracle()
    o.oracle_sa16.argtypes = [_vp, _vp, _i64, ctypes.c_int]
    o.oracle_sa16.restype = ctypes.c_int
    o.oracle_bwt16.argtypes = [_vp, _vp, _vp, _i64]
    o.oracle_bwt16.restype = _i64
    o.oracle_plcp16.argtypes = [_vp, _vp, _vp, _i64, ctypes.c_int]
    o.oracle_plcp16.restype = None
    o.oracle_unbwt16.argtypes = [_vp, _vp, _i64, _i64]
    o.oracle_unbwt16.restype = ctypes.c_int
    return o


def sa16(T, gsa=False):
    T = np.ascontiguousarray(T, np.uint16)
    n = len(T)
    SA = np.zeros(max(n, 1), np.int64)
    rc = _o16().oracle_sa16(_p(T), _p(SA), n, int(gsa))
    return rc, SA[:n]


def bwt16(T, SA):
    T = np.ascontiguousarray(T, np.uint16)
    n = len(T)
    U = np.zeros(max(n, 1), np.uint16)
    SA = np.ascontiguousarray(SA, np.int64)
    idx = _o16().oracle_bwt16(_p(T), _p(SA), _p(U), n)
    return U[:n], idx


def plcp16(T, SA, gsa=False):
    T = np.ascontiguousarray(T, np.uint16)
    n = len(T)
    P = np.zeros(max(n, 1), np.int64)
    SA = np.ascontiguousarray(SA, np.int64)
    _o16().oracle_plcp16(_p(T), _p(SA), _p(P), n, int(gsa))
    return P[:n]


def unbwt16(U, p):
    U = np.ascontiguousarray(U, np.uint16)
    n = len(U)
    out = np.zeros(max(n, 1), np.uint16)
    assert _o16().oracle_unbwt16(_p(U), _p(out), n, p) == 0
    return out[:n]


def ref_sa16(T, gsa=False):
    T = np.ascontiguousarray(T, np.uint16)
    n = len(T)
    SA = np.zeros(max(n, 1), np.int32)
    f = np.zeros(65536, np.int32)
    rc = (ref().libsais16_gsa if gsa else ref().libsais16)(_p(T), _p(SA), n, 0, _p(f))
    return rc, SA[:n], f


def ref_bwt16(T):
    T = np.ascontiguousarray(T, np.uint16)
    n = len(T)
    U = np.zeros(max(n, 1), np.uint16)
    A = np.zeros(max(n, 1), np.int32)
    idx = ref().libsais16_bwt(_p(T), _p(U), _p(A), n, 0, None)
    return U[:n], idx


def ref_plcp16(T, SA, gsa=False):
    T = np.ascontiguousarray(T, np.uint16)
    n = len(T)
    P = np.zeros(max(n, 1), np.int32)
    SA = np.ascontiguousarray(SA, np.int32)
    (ref().libsais16_plcp_gsa if gsa else ref().libsais16_plcp)(_p(T), _p(SA), _p(P), n)
    return P[:n]


def ref_unbwt16(U, p):
    U = np.ascontiguousarray(U, np.uint16)
    n = len(U)
    out = np.zeros(max(n, 1), np.uint16)
    A = np.zeros(n + 1, np.int32)
    rc = ref().libsais16_unbwt(_p(U), _p(out), _p(A), n, None, p)
    return rc, out[:n]


def aux_samples(SA, r):
    """I[t] = 1 + row of suffix t*r: the auxiliary samples of the bwt_aux forms, which
    depend on the suffix array only (any symbol width)."""
    isa = np.empty(len(SA), np.int64)
    isa[np.asarray(SA, np.int64)] = np.arange(len(SA))
    return isa[0::r] + 1


def gsa_set16(rng, count, maxlen, k):
    """A valid 16-bit GSA input: `count` non-empty strings over 1..k-1, each followed by 0."""
    parts = []
    for _ in range(count):
        m = int(rng.integers(1, maxlen + 1))
        parts.append(rng.integers(1, k, m).astype(np.uint16))
        parts.append(np.zeros(1, np.uint16))
    return np.concatenate(parts)
