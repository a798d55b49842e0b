"""Copies between caller host buffers and the device through pinned staging.

The libsais calls must give the same outputs (against the oracle) with every form
of the copies: direct hipMemcpy (`host_staging` 0), 32 MiB pinned slots for copies
of 4 MiB and more (1, the default), and 64 KiB slots for every copy (2: many chunks
per array and a ragged last one), with one, three or the default number of copy
threads.  Outputs that alias inputs (U over T, LCP over SA) must still see the
whole input, which the upload has read before any output is written.
"""
import numpy as np
import pytest

import libsais_amd as la
from libsais_amd import workloads
from tests import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _set(mode, threads):
    lib = la.load()
    assert lib.sagpu_set_option(b"host_staging", mode) == 0
    assert lib.sagpu_set_option(b"copy_threads", threads) == 0


@pytest.fixture
def reset_staging():
    yield
    _set(1, 0)


def _check_api(T):
    n = len(T)
    rc, SA_ref = ol.sa(T)
    assert rc == 0
    U_ref, idx_ref = ol.bwt(T, SA_ref)
    P_ref = ol.plcp(T, SA_ref)
    L_ref = ol.lcp(P_ref, SA_ref)

    SA = np.empty(n, np.int32)
    assert la.libsais(T, SA, n, 0, None) == 0
    assert np.array_equal(SA, SA_ref)
    SA64 = np.empty(n, np.int64)
    assert la.libsais64(T, SA64, n, 0, None) == 0
    assert np.array_equal(SA64, SA_ref.astype(np.int64))

    U = np.empty(n, np.uint8)
    A = np.empty(n, np.int32)
    assert la.libsais_bwt(T, U, A, n, 0, None) == idx_ref
    assert np.array_equal(U, U_ref)
    buf = T.copy()
    assert la.libsais_bwt(buf, buf, A, n, 0, None) == idx_ref
    assert np.array_equal(buf, U_ref)

    P = np.empty(n, np.int32)
    assert la.libsais_plcp(T, SA_ref, P, n) == 0
    assert np.array_equal(P, P_ref)
    P64 = np.empty(n, np.int64)
    assert la.libsais64_plcp(T, SA64, P64, n) == 0
    assert np.array_equal(P64, P_ref.astype(np.int64))

    S = SA_ref.copy()
    assert la.libsais_lcp(P_ref, S, S, n) == 0   # LCP written over SA
    assert np.array_equal(S, L_ref)
    L64 = np.empty(n, np.int64)
    assert la.libsais64_lcp(P64, SA64, L64, n) == 0
    assert np.array_equal(L64, L_ref.astype(np.int64))

    R = np.empty(n, np.uint8)
    assert la.libsais_unbwt(U_ref, R, A, n, None, idx_ref) == 0
    assert np.array_equal(R, T)


@pytest.mark.parametrize("mode,threads", [(0, 0), (1, 0), (2, 1), (2, 3), (2, 0)])
def test_api_small(mode, threads, reset_staging):
    """300,007 bytes: with 64 KiB slots the SA is 19 chunks, the int64 SA 37, the
    text 5; each ends ragged."""
    _set(mode, threads)
    _check_api(workloads.dna_like(300_007, seed=31))


@pytest.mark.parametrize("mode", [1, 0])
def test_api_large(mode, reset_staging):
    """9,000,011 bytes: with the default 32 MiB slots the SA (36 MB) takes two chunks
    and the int64 SA (72 MB) three, so slots are reused within one copy."""
    _set(mode, 0)
    _check_api(workloads.dna_like(9_000_011, seed=32))


def test_16bit_and_int(reset_staging):
    _set(2, 0)
    rng = np.random.default_rng(77)
    T16 = rng.integers(0, 300, 100_003).astype(np.uint16)
    n = len(T16)
    SA = np.empty(n, np.int32)
    assert la.libsais16(T16, SA, n, 0, None) == 0
    assert np.array_equal(SA, ol.sa16(T16)[1])
    U = np.empty(n, np.uint16)
    A = np.empty(n, np.int32)
    U_ref, idx_ref = ol.bwt16(T16, SA)
    assert la.libsais16_bwt(T16, U, A, n, 0, None) == idx_ref
    assert np.array_equal(U, U_ref)
    R = np.empty(n, np.uint16)
    assert la.libsais16_unbwt(U_ref, R, A, n, None, idx_ref) == 0
    assert np.array_equal(R, T16)
    TI = rng.integers(0, 5000, 100_003).astype(np.int32)
    SAI = np.empty(n, np.int32)
    assert la.libsais_int(TI.copy(), SAI, n, 5000, 0) == 0
    assert np.array_equal(SAI, ol.sa_int(TI))
