"""The 64-bit-key mode that texts longer than 2^31 - 1 take (libsais64 beyond
INT32_MAX), forced on for small inputs with the "wide_from" option so its kernels
(u64 rank keys, u64 PLCP scan values, 33-bit composite big-group keys) are checked
against the oracle at sizes it finishes quickly.  The full 4 GiB case lives in
test_gpu_c3d.py."""
import numpy as np
import pytest

import libsais_amd as la
from libsais_amd import batch, workloads
from tests import oracle_lib as ol
from tests.test_gpu_parity import fused

pytestmark = pytest.mark.gpu

INT32_MAX = 0x7FFFFFFF


@pytest.fixture(autouse=True)
def wide_mode():
    lib = la.load()
    assert lib.sagpu_set_option(b"wide_from", 0) == 0
    yield
    assert lib.sagpu_set_option(b"wide_from", INT32_MAX) == 0


def _cases():
    rng = np.random.default_rng(77)
    cases = [("acgt65536", workloads.iid(65536, seed=1, alphabet=b"ACGT")),
             ("acgt200003", workloads.iid(200003, seed=2, alphabet=b"ACGT")),
             ("bytes_k2", rng.integers(0, 2, 50000).astype(np.uint8)),
             ("bytes_k256", rng.integers(0, 256, 50000).astype(np.uint8)),
             ("all_a_70000", np.full(70000, 65, np.uint8)),
             ("periodic17", workloads.periodic(100000, 17, 3)),
             ("fibonacci", workloads.fibonacci(100000)),
             ("dna_like", workloads.dna_like(1 << 19, seed=5))]
    T = workloads.iid(20000, seed=257, alphabet=b"ACGT")
    T[5000:5257] = 65
    T[12000:12257] = 65
    cases.append(("runs257", T))
    cases.append(("copies300", np.tile(workloads.iid(3000, seed=99, alphabet=b"ACGT"), 300)))
    cases.append(("banana", np.frombuffer(b"banana", np.uint8).copy()))
    return cases


CASES = _cases()


@pytest.mark.parametrize("name,T", CASES, ids=[c[0] for c in CASES])
def test_wide_matches_oracle(name, T):
    print(f"wide case {name} n={len(T)}", flush=True)
    n = len(T)
    rc, SA_ref = ol.sa(T)
    assert rc == 0
    U_ref, idx_ref = ol.bwt(T, SA_ref)
    P_ref = ol.plcp(T, SA_ref)
    L_ref = ol.lcp(P_ref, SA_ref)

    SA = np.zeros(n, np.int64)
    f = np.zeros(256, np.int64)
    assert la.libsais64(T, SA, n, 0, f) == 0, la.last_error()
    assert np.array_equal(SA, SA_ref.astype(np.int64))
    assert np.array_equal(f, ol.freq(T))
    U = np.zeros(n, np.uint8)
    A = np.zeros(n, np.int64)
    assert la.libsais64_bwt(T, U, A, n, 0, None) == idx_ref
    assert np.array_equal(U, U_ref)
    P = np.zeros(n, np.int64)
    assert la.libsais64_plcp(T, SA, P, n) == 0
    assert np.array_equal(P, P_ref.astype(np.int64))
    L = np.zeros(n, np.int64)
    assert la.libsais64_lcp(P, SA, L, n) == 0
    assert np.array_equal(L, L_ref.astype(np.int64))

    fSA, fU, fidx, fP, fL, _ = fused(T)
    assert np.array_equal(fSA, SA_ref)
    assert fidx == idx_ref and np.array_equal(fU, U_ref)
    assert np.array_equal(fP, P_ref)
    assert np.array_equal(fL, L_ref)


def test_wide_gsa():
    rng = np.random.default_rng(12)
    parts = [workloads.iid(int(rng.integers(1, 3000)), seed=q, alphabet=b"ACGT").tobytes() for q in range(40)]
    parts += [b"A" * 900, b"AC" * 400]
    T = np.frombuffer(b"".join(p + b"\x00" for p in parts), np.uint8).copy()
    n = len(T)
    rc, SA_ref = ol.sa(T, gsa=True)
    assert rc == 0
    SA = np.zeros(n, np.int64)
    assert la.libsais64_gsa(T, SA, n, 0, None) == 0
    assert np.array_equal(SA, SA_ref.astype(np.int64))
    P = np.zeros(n, np.int64)
    assert la.libsais64_plcp_gsa(T, SA, P, n) == 0
    assert np.array_equal(P, ol.plcp(T, SA_ref, gsa=True).astype(np.int64))
    fSA, _, _, fP, fL, _ = fused(T, gsa=True)
    assert np.array_equal(fSA, SA_ref)
    assert np.array_equal(fL, ol.lcp(ol.plcp(T, SA_ref, gsa=True), SA_ref))


def test_wide_batched_gsa():
    T, offs = workloads.gsa_batch(count=200, seed=4, median=300, min_len=1, max_len=6000, target_total=60_000,
                                  shared_frac=0.2)
    SA, L = batch.batched(T, offs)
    for k in range(0, len(offs) - 1, 7):
        o, e = int(offs[k]), int(offs[k + 1])
        Tk = np.ascontiguousarray(T[o:e])
        rc, SAk = ol.sa(Tk, gsa=True)
        assert np.array_equal(SA[o:e], SAk), k
        assert np.array_equal(L[o:e], ol.lcp(ol.plcp(Tk, SAk, gsa=True), SAk)), k


def test_wide_from_rejects_out_of_range():
    lib = la.load()
    assert lib.sagpu_set_option(b"wide_from", -1) == -1
    assert lib.sagpu_set_option(b"wide_from", INT32_MAX + 1) == -1
