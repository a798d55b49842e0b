"""The allocation-failure contract: -2 when working memory cannot be obtained.

The reference returns -2 when its allocations fail (src/libsais.c:6927-6932,
src/libsais64.c:6982-6995).  Here the engine's device working memory is capped with
sagpu_set_memory_limit below what a call needs: every entry point must then return
-2 with an error string, leave no state behind that breaks the next call, and, with
the cap lifted, the next calls (on the same context and on a new one) must return
results identical to the oracle's.
"""
import numpy as np
import pytest

import libsais_amd as la
from libsais_amd import workloads
from tests import oracle_lib as ol

pytestmark = pytest.mark.gpu

N = 1 << 20


@pytest.fixture(scope="module")
def text():
    T = workloads.dna_like(N, seed=31)
    rc, SA = ol.sa(T)
    assert rc == 0
    return T, SA, ol.plcp(T, SA)


def _calls(T, SA, P):
    """(name, thunk) for every entry point under test; each thunk returns the code."""
    n = len(T)
    i32, i64 = np.int32, np.int64
    G = T.copy()
    G[-1] = 0
    ctx = la.libsais_create_ctx()
    out = [
        ("libsais", lambda: la.libsais(T, np.zeros(n, i32), n, 0, None)),
        ("libsais_gsa", lambda: la.libsais_gsa(G, np.zeros(n, i32), n, 0, None)),
        ("libsais_bwt", lambda: la.libsais_bwt(T, np.zeros(n, np.uint8), np.zeros(n, i32), n, 0, None)),
        ("libsais_plcp", lambda: la.libsais_plcp(T, SA, np.zeros(n, i32), n)),
        ("libsais_lcp", lambda: la.libsais_lcp(P, SA, np.zeros(n, i32), n)),
        ("libsais64", lambda: la.libsais64(T, np.zeros(n, i64), n, 0, None)),
        ("libsais64_bwt", lambda: la.libsais64_bwt(T, np.zeros(n, np.uint8), np.zeros(n, i64), n, 0, None)),
        ("libsais64_plcp", lambda: la.libsais64_plcp(T, SA.astype(i64), np.zeros(n, i64), n)),
        ("libsais_ctx", lambda: la.libsais_ctx(ctx, T, np.zeros(n, i32), n, 0, None)),
    ]
    return out, ctx


@pytest.mark.parametrize("limit_mib", [1, 8, 20])
def test_capped_memory_returns_minus_two_then_recovers(text, limit_mib):
    T, SA, P = text
    n = len(T)
    lib = la.load()
    calls, ctx = _calls(T, SA, P)
    try:
        lib.sagpu_set_memory_limit(limit_mib << 20)
        for name, call in calls:
            if limit_mib >= 20 and name == "libsais_lcp":
                continue   # needs 12 n bytes = 12 MiB, which fits under this cap
            print(f"[-2] cap {limit_mib} MiB: {name}", flush=True)
            rc = call()
            assert rc == -2, (name, rc)
            assert la.last_error(), name
    finally:
        lib.sagpu_set_memory_limit(0)
    # the same contexts work again
    SA1 = np.zeros(n, np.int32)
    f = np.zeros(256, np.int32)
    assert la.libsais(T, SA1, n, 0, f) == 0, la.last_error()
    assert np.array_equal(SA1, SA) and np.array_equal(f, ol.freq(T).astype(np.int32))
    SA2 = np.zeros(n, np.int32)
    assert la.libsais_ctx(ctx, T, SA2, n, 0, None) == 0 and np.array_equal(SA2, SA)
    U = np.zeros(n, np.uint8)
    idx = la.libsais_bwt(T, U, np.zeros(n, np.int32), n, 0, None)
    U_r, idx_r = ol.bwt(T, SA)
    assert idx == idx_r and np.array_equal(U, U_r)
    P1 = np.zeros(n, np.int32)
    assert la.libsais_plcp(T, SA, P1, n) == 0 and np.array_equal(P1, P)
    L = np.zeros(n, np.int32)
    assert la.libsais_lcp(P, SA, L, n) == 0 and np.array_equal(L, ol.lcp(P, SA))
    S64 = np.zeros(n, np.int64)
    assert la.libsais64(T, S64, n, 0, None) == 0 and np.array_equal(S64, SA.astype(np.int64))
    la.libsais_free_ctx(ctx)


def test_capped_memory_on_device_entry_points(text):
    """The device-level build reports -2 the same way."""
    T, SA, _ = text
    n = len(T)
    lib = la.load()
    dT = lib.sagpu_malloc(n)
    dSA = lib.sagpu_malloc(4 * n)
    try:
        assert lib.sagpu_memcpy_h2d(dT, T.ctypes.data, n) == 0
        lib.sagpu_set_memory_limit(8 << 20)
        try:
            assert lib.sagpu_build32(dT, n, 0, dSA, None, None, None, None) == -2
            assert "limit" in la.last_error()
        finally:
            lib.sagpu_set_memory_limit(0)
        assert lib.sagpu_build32(dT, n, 0, dSA, None, None, None, None) == 0, la.last_error()
        out = np.empty(n, np.int32)
        assert lib.sagpu_memcpy_d2h(out.ctypes.data, dSA, 4 * n) == 0
        assert np.array_equal(out, SA)
    finally:
        lib.sagpu_free(dT)
        lib.sagpu_free(dSA)
