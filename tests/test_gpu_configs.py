"""Benchmark-sized inputs on the GPU, compared bit-exactly with the reference.

* 1 MiB i.i.d. ACGT (seed 1): digests of SA, BWT, PLCP, LCP made by the reference
  build and committed in tests/golden/golden.json.
* 64 MiB degenerate strings (all 'A', period 17 with one mutation, Fibonacci word)
  and the 256 MiB DNA-like string: the reference libsais (oracle/_ref, OpenMP) runs
  on the host in the same test; without it the O(n) suffix-array checker and the
  closed forms are used.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import libsais_amd as la
from libsais_amd import workloads
from tests import oracle_lib as ol

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "golden.json")
THREADS = min(16, os.cpu_count() or 1)


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _gpu_all(T):
    n = len(T)
    SA = np.empty(n, np.int32)
    freq = np.empty(256, np.int32)
    assert la.libsais(T, SA, n, 0, freq) == 0, la.last_error()
    U = np.empty(n, np.uint8)
    A = np.empty(1, np.int32)
    idx = la.libsais_bwt(T, U, A, n, 0, None)
    assert idx > 0, la.last_error()
    P = np.empty(n, np.int32)
    assert la.libsais_plcp(T, SA, P, n) == 0
    L = np.empty(n, np.int32)
    assert la.libsais_lcp(P, SA, L, n) == 0
    return SA, freq, U, idx, P, L


def test_acgt_1mib_golden():
    g = json.load(open(GOLDEN))["acgt_1MiB_seed1"]
    T = workloads.iid_acgt(g["n"], seed=1)
    assert _sha(T) == g["T_sha256"]
    SA, freq, U, idx, P, L = _gpu_all(T)
    assert _sha(SA) == g["SA_int32_sha256"]
    assert _sha(U) == g["U_sha256"] and idx == g["primary"]
    assert _sha(P) == g["PLCP_int32_sha256"]
    assert _sha(L) == g["LCP_int32_sha256"]
    assert freq.tolist() == g["freq"]
    SA64 = np.empty(len(T), np.int64)
    assert la.libsais64(T, SA64, len(T), 0, None) == 0
    assert _sha(SA64) == g["SA_int64_sha256"]


def _compare_with_reference(T):
    SA, freq, U, idx, P, L = _gpu_all(T)
    if ol.have_ref():
        rc, SA_r, f_r = ol.ref_sa(T, threads=THREADS)
        assert rc == 0
        assert np.array_equal(SA, SA_r)
        assert np.array_equal(freq, f_r)
        U_r, idx_r = ol.ref_bwt(T, threads=THREADS)
        assert idx == idx_r and np.array_equal(U, U_r)
        P_r = ol.ref_plcp(T, SA_r, threads=THREADS)
        assert np.array_equal(P, P_r)
        assert np.array_equal(L, ol.ref_lcp(P_r, SA_r, threads=THREADS))
    else:
        assert ol.check_sa(T, SA)
        U_o, idx_o = ol.bwt(T, SA)
        assert idx == idx_o and np.array_equal(U, U_o)
        P_o = ol.plcp(T, SA)
        assert np.array_equal(P, P_o)
        assert np.array_equal(L, P_o[SA])
    return SA, P, L


@pytest.mark.slow
def test_all_a_64mib():
    n = 1 << 26
    T = workloads.all_a(n)
    SA, P, L = _compare_with_reference(T)
    assert SA[0] == n - 1 and SA[-1] == 0
    assert L[-1] == n - 1


@pytest.mark.slow
def test_period17_64mib():
    _compare_with_reference(workloads.periodic(1 << 26, 17, 3))


@pytest.mark.slow
def test_fibonacci_64mib():
    _compare_with_reference(workloads.fibonacci(1 << 26))


@pytest.mark.slow
def test_dna_like_256mib():
    _compare_with_reference(workloads.dna_like(1 << 28, seed=2))


def test_fused_build_matches_separate_calls():
    """sagpu_build32 (device buffers, one pass) == the four host API calls."""
    lib = la.load()
    T = workloads.dna_like(1 << 20, seed=12)
    n = len(T)
    SA, freq, U, idx, P, L = _gpu_all(T)
    dT = lib.sagpu_malloc(n)
    bufs = [lib.sagpu_malloc(4 * n), lib.sagpu_malloc(n), lib.sagpu_malloc(4 * n), lib.sagpu_malloc(4 * n)]
    assert lib.sagpu_memcpy_h2d(dT, T.ctypes.data, n) == 0
    f64 = np.zeros(256, np.int64)
    r = lib.sagpu_build32(dT, n, 0, bufs[0], bufs[1], bufs[2], bufs[3], f64.ctypes.data)
    assert r == idx, la.last_error()
    outs = [np.empty(n, np.int32), np.empty(n, np.uint8), np.empty(n, np.int32), np.empty(n, np.int32)]
    for o, b in zip(outs, bufs):
        lib.sagpu_memcpy_d2h(o.ctypes.data, b, o.nbytes)
        lib.sagpu_free(b)
    lib.sagpu_free(dT)
    assert np.array_equal(outs[0], SA)
    assert np.array_equal(outs[1], U)
    assert np.array_equal(outs[2], P)
    assert np.array_equal(outs[3], L)
    assert np.array_equal(f64, freq.astype(np.int64))
