"""BWT with auxiliary indexes and the inverse BWT on the GPU, through the C ABI.

libsais_bwt_aux must equal the oracle's (U, I) bit for bit (and the compiled
reference's where it is built); libsais_unbwt / libsais_unbwt_aux must return the
original text for every BWT the reference layout can produce.  Both decoding paths
are covered: independent chunk walks (many auxiliary samples) and list ranking
(plain unbwt, or few samples)."""
import numpy as np
import pytest

import libsais_amd as la
from libsais_amd import workloads
from tests import oracle_lib as ol
from tests.test_gpu_parity import CASES

pytestmark = pytest.mark.gpu

SHAPES = [c for c in CASES if len(c[1]) >= 2]


@pytest.mark.parametrize("name,T", SHAPES, ids=[c[0] for c in SHAPES])
def test_unbwt_round_trip(name, T):
    print(f"unbwt case {name} n={len(T)}", flush=True)
    n = len(T)
    rc, SA = ol.sa(T)
    U_ref, p = ol.bwt(T, SA)
    out = np.zeros(n, np.uint8)
    A = np.zeros(n + 1, np.int32)
    assert la.libsais_unbwt(U_ref, out, A, n, None, p) == 0, la.last_error()
    assert np.array_equal(out, T)
    for r in (2, 64, 4096):
        U, I = ol.bwt_aux(T, SA, r)
        I32 = I.astype(np.int32)
        GU = np.zeros(n, np.uint8)
        GI = np.zeros(len(I), np.int32)
        assert la.libsais_bwt_aux(T, GU, A, n, 0, None, r, GI) == 0, la.last_error()
        assert np.array_equal(GU, U) and np.array_equal(GI, I32), r
        out[:] = 0
        assert la.libsais_unbwt_aux(U, out, A, n, None, r, I32) == 0
        assert np.array_equal(out, T), r


def test_unbwt_in_place_and_64bit():
    T = workloads.dna_like(1 << 20, seed=11)
    n = len(T)
    rc, SA = ol.sa(T)
    U, p = ol.bwt(T, SA)
    buf = U.copy()
    A64 = np.zeros(1, np.int64)
    assert la.libsais64_unbwt(buf, buf, A64, n, None, p) == 0  # U aliases T
    assert np.array_equal(buf, T)
    for r in (16, 1 << 12):
        Ue, I = ol.bwt_aux(T, SA, r)
        GU = np.zeros(n, np.uint8)
        GI = np.zeros(len(I), np.int64)
        assert la.libsais64_bwt_aux(T, GU, A64, n, 0, None, r, GI) == 0
        assert np.array_equal(GU, Ue) and np.array_equal(GI, I)
        out = np.zeros(n, np.uint8)
        assert la.libsais64_unbwt_aux(Ue, out, A64, n, None, r, I) == 0
        assert np.array_equal(out, T)


def test_unbwt_large_list_ranking_and_chunk_walks():
    T = workloads.iid(24 << 20, seed=12, alphabet=b"ACGT")
    n = len(T)
    U = np.zeros(n, np.uint8)
    A = np.zeros(1, np.int32)
    p = la.libsais_bwt(T, U, A, n, 0, None)
    assert p > 0
    out = np.zeros(n, np.uint8)
    assert la.libsais_unbwt(U, out, A, n, None, p) == 0          # list ranking
    assert np.array_equal(out, T)
    GU = np.zeros(n, np.uint8)
    I = np.zeros((n - 1) // 64 + 1, np.int32)
    assert la.libsais_bwt_aux(T, GU, A, n, 0, None, 64, I) == 0
    assert np.array_equal(GU, U) and I[0] == p
    out[:] = 0
    assert la.libsais_unbwt_aux(U, out, A, n, None, 64, I) == 0  # 393216 chunk walks
    assert np.array_equal(out, T)


@pytest.mark.skipif(not ol.have_ref(), reason="reference build oracle/_ref absent")
def test_bwt_aux_matches_reference():
    for T in (workloads.dna_like(1 << 20, seed=13), workloads.fibonacci(200000), workloads.periodic(300000, 17, 3)):
        n = len(T)
        for r in (2, 256, 1 << 16):
            rc, U_ref, I_ref = ol.ref_bwt_aux(T, r)
            assert rc == 0
            U, I = la.bwt_aux(T, r)
            assert np.array_equal(U, U_ref) and np.array_equal(I, I_ref), (n, r)
            rc, out = ol.ref_unbwt_aux(U, r, I)
            assert rc == 0 and np.array_equal(out, T)
            assert np.array_equal(la.unbwt(U, int(I[0])), T)


def test_inconsistent_primary_index_stays_bounded():
    # A primary index that does not belong to this BWT decodes to some string (the
    # reference does not detect it either); the call must still return normally.
    T = workloads.iid(100000, seed=14, alphabet=b"ACGT")
    rc, SA = ol.sa(T)
    U, p = ol.bwt(T, SA)
    out = np.zeros(len(T), np.uint8)
    A = np.zeros(1, np.int32)
    wrong = p % len(T) + 1
    assert la.libsais_unbwt(U, out, A, len(T), None, wrong) == 0
