"""Integer alphabets on the GPU (libsais_int, libsais64_long, libsais_plcp_int)
through the C ABI, against the oracle (and the compiled reference where it is
built).  The context and OpenMP forms are compared with the reference's in
test_gpu_forms.py."""
import numpy as np
import pytest

import libsais_amd as la
from libsais_amd import workloads
from tests import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _int_cases():
    rng = np.random.default_rng(404)
    cases = []
    for k in (1, 2, 4, 100, 1 << 16, (1 << 16) + 1, 1 << 20, (1 << 31) - 1):
        for n in (2, 3, 1000, 70001):
            cases.append((f"k{k}_n{n}", rng.integers(0, k, n).astype(np.int32), k))
    # repetitive integer strings (long groups, many rounds)
    base = rng.integers(0, 50000, 37).astype(np.int32)
    cases.append(("periodic37_k50000", np.tile(base, 3000), 50000))
    cases.append(("all_equal_big_symbol", np.full(40000, 123456789, np.int32), 123456790))
    cases.append(("dna_as_ints", workloads.dna_like(1 << 18, seed=9).astype(np.int32), 256))
    return cases


CASES = _int_cases()


@pytest.mark.parametrize("name,T,k", CASES, ids=[c[0] for c in CASES])
def test_libsais_int(name, T, k):
    print(f"int case {name}", flush=True)
    n = len(T)
    SA_ref = ol.sa_int(T)
    Tin = T.copy()
    SA = np.zeros(n, np.int32)
    assert la.libsais_int(Tin, SA, n, k, 0) == 0, la.last_error()
    assert np.array_equal(Tin, T)
    assert np.array_equal(SA, SA_ref.astype(np.int32))
    P = np.zeros(n, np.int32)
    assert la.libsais_plcp_int(T, SA, P, n) == 0
    assert np.array_equal(P, ol.plcp_int(T, SA_ref).astype(np.int32))


def test_libsais64_long_wide_values():
    rng = np.random.default_rng(5)
    for n in (2, 1000, 50000):
        T = rng.integers(0, 1 << 40, n).astype(np.int64)
        T[::7] = T[0]  # repeats
        SA = np.zeros(n, np.int64)
        assert la.libsais64_long(T.copy(), SA, n, 1 << 40, 0) == 0, la.last_error()
        assert np.array_equal(SA, ol.sa_int(T))
    T = rng.integers(0, 3, 100000).astype(np.int64)
    SA = np.zeros(len(T), np.int64)
    assert la.libsais64_long(T, SA, len(T), 3, 0) == 0
    assert np.array_equal(SA, ol.sa_int(T))


@pytest.mark.skipif(not ol.have_ref(), reason="reference build oracle/_ref absent")
def test_int_against_reference():
    rng = np.random.default_rng(6)
    for k in (2, 300, 1 << 20):
        T = rng.integers(0, k, 200000).astype(np.int32)
        rc, SA_ref, _ = ol.ref_sa_int(T, k)
        assert rc == 0
        SA = np.zeros(len(T), np.int32)
        assert la.libsais_int(T.copy(), SA, len(T), k, 0) == 0
        assert np.array_equal(SA, SA_ref)
        rc, P_ref = ol.ref_plcp_int(T, SA_ref)
        P = np.zeros(len(T), np.int32)
        assert la.libsais_plcp_int(T, SA, P, len(T)) == 0
        assert np.array_equal(P, P_ref)
