"""16-bit alphabet on the GPU (libsais16.h / libsais16x64.h) through the C ABI:
suffix arrays (plain and GSA, with the 65536-entry freq table), BWT (plain and
auxiliary), inverse BWT, PLCP / PLCP-GSA / LCP, the int / long forms and the
context and OpenMP forms, against the oracle (src/libsais16.c restated) and the
compiled reference where it is built."""
import numpy as np
import pytest

import libsais_amd as la
from tests import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _cases():
    rng = np.random.default_rng(1616)
    cases = []
    for k in (2, 4, 300, 1 << 12, 1 << 16):
        for n in (2, 3, 1000, 70001):
            cases.append((f"k{k}_n{n}", rng.integers(0, k, n).astype(np.uint16)))
    base = rng.integers(0, 65536, 29).astype(np.uint16)
    cases.append(("periodic29_full_range", np.tile(base, 4000)))
    cases.append(("all_65535", np.full(50000, 65535, np.uint16)))
    cases.append(("two_symbols_high", rng.choice(np.array([65534, 65535], np.uint16), 100000)))
    runs = np.repeat(rng.integers(0, 65536, 500).astype(np.uint16), rng.integers(1, 400, 500))
    cases.append(("runs", runs))
    return cases


CASES = _cases()
IDS = [c[0] for c in CASES]


@pytest.mark.parametrize("name,T", CASES, ids=IDS)
def test_libsais16_family(name, T):
    print(f"16-bit case {name} n={len(T)}", flush=True)
    n = len(T)
    rc, SA_o = ol.sa16(T)
    assert rc == 0
    SA = np.zeros(n, np.int32)
    freq = np.zeros(65536, np.int32)
    assert la.libsais16(T, SA, n, 0, freq) == 0, la.last_error()
    assert np.array_equal(SA, SA_o.astype(np.int32))
    assert np.array_equal(freq, np.bincount(T, minlength=65536).astype(np.int32))

    U_o, p_o = ol.bwt16(T, SA_o)
    U = np.zeros(n, np.uint16)
    A = np.zeros(n, np.int32)
    f2 = np.full(65536, 7, np.int32)
    p = la.libsais16_bwt(T, U, A, n, 0, f2)
    assert p == p_o and np.array_equal(U, U_o)
    assert np.array_equal(f2, freq)

    for r in (2, 64, 1 << 20):
        Ua = np.zeros(n, np.uint16)
        I = np.zeros((n - 1) // r + 1, np.int32)
        assert la.libsais16_bwt_aux(T, Ua, A, n, 0, None, r, I) == 0
        assert np.array_equal(Ua, U_o) and np.array_equal(I, ol.aux_samples(SA_o, r))
        back = np.zeros(n, np.uint16)
        assert la.libsais16_unbwt_aux(Ua, back, A, n, None, r, I) == 0
        assert np.array_equal(back, T)

    back = np.zeros(n, np.uint16)
    assert la.libsais16_unbwt(U, back, A, n, freq, p) == 0
    assert np.array_equal(back, T)

    P = np.zeros(n, np.int32)
    assert la.libsais16_plcp(T, SA, P, n) == 0
    P_o = ol.plcp16(T, SA_o)
    assert np.array_equal(P, P_o.astype(np.int32))
    L = np.zeros(n, np.int32)
    assert la.libsais16_lcp(P, SA, L, n) == 0
    assert np.array_equal(L, P[SA])


@pytest.mark.parametrize("name,T", CASES[::3], ids=IDS[::3])
def test_libsais16x64_family(name, T):
    print(f"16x64 case {name} n={len(T)}", flush=True)
    n = len(T)
    rc, SA_o = ol.sa16(T)
    SA = np.zeros(n, np.int64)
    freq = np.zeros(65536, np.int64)
    assert la.libsais16x64(T, SA, n, 0, freq) == 0, la.last_error()
    assert np.array_equal(SA, SA_o)
    assert np.array_equal(freq, np.bincount(T, minlength=65536))
    U_o, p_o = ol.bwt16(T, SA_o)
    U = np.zeros(n, np.uint16)
    A = np.zeros(n, np.int64)
    assert la.libsais16x64_bwt(T, U, A, n, 0, None) == p_o and np.array_equal(U, U_o)
    I = np.zeros((n - 1) // 16 + 1, np.int64)
    Ua = np.zeros(n, np.uint16)
    assert la.libsais16x64_bwt_aux(T, Ua, A, n, 0, None, 16, I) == 0
    assert np.array_equal(Ua, U_o) and np.array_equal(I, ol.aux_samples(SA_o, 16))
    back = np.zeros(n, np.uint16)
    assert la.libsais16x64_unbwt_aux(Ua, back, A, n, None, 16, I) == 0 and np.array_equal(back, T)
    back[:] = 0
    assert la.libsais16x64_unbwt(U, back, A, n, None, p_o) == 0 and np.array_equal(back, T)
    P = np.zeros(n, np.int64)
    assert la.libsais16x64_plcp(T, SA, P, n) == 0 and np.array_equal(P, ol.plcp16(T, SA_o))
    L = np.zeros(n, np.int64)
    assert la.libsais16x64_lcp(P, SA, L, n) == 0 and np.array_equal(L, P[SA])


@pytest.mark.parametrize("count,maxlen,k", [(1, 50, 3), (40, 200, 5), (300, 1000, 65536), (2000, 30, 2)])
def test_libsais16_gsa(count, maxlen, k):
    rng = np.random.default_rng(count * 31 + k)
    T = ol.gsa_set16(rng, count, maxlen, k)
    n = len(T)
    print(f"16-bit gsa count={count} n={n} k={k}", flush=True)
    rc, SA_o = ol.sa16(T, gsa=True)
    assert rc == 0
    SA = np.zeros(n, np.int32)
    freq = np.zeros(65536, np.int32)
    assert la.libsais16_gsa(T, SA, n, 0, freq) == 0, la.last_error()
    assert np.array_equal(SA, SA_o.astype(np.int32))
    assert np.array_equal(freq, np.bincount(T, minlength=65536).astype(np.int32))
    P = np.zeros(n, np.int32)
    assert la.libsais16_plcp_gsa(T, SA, P, n) == 0
    assert np.array_equal(P, ol.plcp16(T, SA_o, gsa=True).astype(np.int32))
    S64 = np.zeros(n, np.int64)
    assert la.libsais16x64_gsa(T, S64, n, 0, None) == 0 and np.array_equal(S64, SA_o)
    P64 = np.zeros(n, np.int64)
    assert la.libsais16x64_plcp_gsa(T, S64, P64, n) == 0 and np.array_equal(P64, P.astype(np.int64))


def test_gsa_requires_trailing_zero():
    T = np.array([3, 2, 1], np.uint16)
    SA = np.zeros(3, np.int32)
    assert la.libsais16_gsa(T, SA, 3, 0, None) == -1
    assert la.libsais16_plcp_gsa(T, SA, SA, 3) == -1


def _invalid16(rng):
    """16-bit GSA inputs with an empty string (a 0 at position 0, or two adjacent 0s),
    which the engine detects on the device while it counts symbols."""
    out = []
    for k, count, at in ((3, 40, "start"), (1000, 200, "middle"), (65536, 50, "end"), (5, 3000, "many")):
        G = ol.gsa_set16(rng, count, 60, k)
        if at == "start":
            G = np.concatenate([np.zeros(1, np.uint16), G])
        elif at == "middle":
            z = np.flatnonzero(G == 0)[count // 2]
            G = np.concatenate([G[:z + 1], np.zeros(1, np.uint16), G[z + 1:]])
        elif at == "end":
            G = np.concatenate([G, np.zeros(1, np.uint16)])
        else:
            z = np.flatnonzero(G == 0)
            G = np.insert(G, z[::7] + 1, 0)
        out.append((at, G))
    return out


def test_gsa16_rejects_empty_strings_like_the_oracle():
    """libsais16_gsa and libsais16x64_gsa return -1 for a set with an empty string
    (src/libsais16.c:6863), and 0 for the same set without it."""
    rng = np.random.default_rng(166)
    for at, G in _invalid16(rng):
        n = len(G)
        print(f"16-bit gsa invalid ({at}) n={n}", flush=True)
        assert not ol.gsa_valid16(G)
        SA = np.zeros(n, np.int32)
        f = np.zeros(65536, np.int32)
        assert la.libsais16_gsa(G, SA, n, 0, f) == -1, at
        S64 = np.zeros(n, np.int64)
        assert la.libsais16x64_gsa(G, S64, n, 0, None) == -1, at
        if ol.have_ref():
            rc, _, _ = ol.ref_sa16(G, gsa=True)
            assert rc == -1, at
    G = ol.gsa_set16(rng, 500, 60, 300)
    assert ol.gsa_valid16(G)
    SA = np.zeros(len(G), np.int32)
    assert la.libsais16_gsa(G, SA, len(G), 0, None) == 0
    assert np.array_equal(SA, ol.sa16(G, gsa=True)[1].astype(np.int32))


@pytest.mark.skipif(not ol.have_ref(), reason="reference build oracle/_ref absent")
@pytest.mark.parametrize("name,T", CASES[2::4], ids=IDS[2::4])
def test_against_reference(name, T):
    print(f"16-bit reference case {name}", flush=True)
    n = len(T)
    rc, SA_r, f_r = ol.ref_sa16(T)
    assert rc == 0
    SA = np.zeros(n, np.int32)
    f = np.zeros(65536, np.int32)
    assert la.libsais16(T, SA, n, 0, f) == 0
    assert np.array_equal(SA, SA_r) and np.array_equal(f, f_r)
    U_r, p_r = ol.ref_bwt16(T)
    U = np.zeros(n, np.uint16)
    A = np.zeros(n, np.int32)
    assert la.libsais16_bwt(T, U, A, n, 0, None) == p_r and np.array_equal(U, U_r)
    P = np.zeros(n, np.int32)
    assert la.libsais16_plcp(T, SA, P, n) == 0 and np.array_equal(P, ol.ref_plcp16(T, SA_r))
    rc, back = ol.ref_unbwt16(U_r, p_r)
    assert rc == 0 and np.array_equal(back, T)


def test_int_long_and_context_forms():
    rng = np.random.default_rng(77)
    T = rng.integers(0, 1000, 60000).astype(np.uint16)
    n = len(T)
    SA = np.zeros(n, np.int32)
    assert la.libsais16(T, SA, n, 0, None) == 0
    Ti = T.astype(np.int32)
    Si = np.zeros(n, np.int32)
    assert la.libsais16_int(Ti.copy(), Si, n, 1000, 0) == 0 and np.array_equal(Si, SA)
    assert la.libsais16_int_omp(Ti.copy(), Si, n, 1000, 0, 2) == 0 and np.array_equal(Si, SA)
    Sl = np.zeros(n, np.int64)
    assert la.libsais16x64_long(T.astype(np.int64), Sl, n, 1000, 0) == 0 and np.array_equal(Sl, SA)
    ctx = la.libsais16_create_ctx()
    try:
        S2 = np.zeros(n, np.int32)
        assert la.libsais16_ctx(ctx, T, S2, n, 0, None) == 0 and np.array_equal(S2, SA)
        U1, U2 = np.zeros(n, np.uint16), np.zeros(n, np.uint16)
        A = np.zeros(n, np.int32)
        p1 = la.libsais16_bwt(T, U1, A, n, 0, None)
        assert la.libsais16_bwt_ctx(ctx, T, U2, A, n, 0, None) == p1 and np.array_equal(U1, U2)
    finally:
        la.libsais16_free_ctx(ctx)
    S3 = np.zeros(n, np.int32)
    assert la.libsais16_omp(T, S3, n, 0, None, 4) == 0 and np.array_equal(S3, SA)
    uctx = la.libsais16_unbwt_create_ctx()
    try:
        out = np.zeros(n, np.uint16)
        assert la.libsais16_unbwt_ctx(uctx, U1, out, A, n, None, p1) == 0 and np.array_equal(out, T)
    finally:
        la.libsais16_unbwt_free_ctx(uctx)
    P1, P2 = np.zeros(n, np.int32), np.zeros(n, np.int32)
    assert la.libsais16_plcp(T, SA, P1, n) == 0
    assert la.libsais16_plcp_omp(T, SA, P2, n, 0) == 0 and np.array_equal(P1, P2)
    S64 = np.zeros(n, np.int64)
    assert la.libsais16x64_omp(T, S64, n, 0, None, 2) == 0 and np.array_equal(S64, SA.astype(np.int64))
