"""GPU parity: every libsais entry point through the C ABI against the oracle.

Inputs are seeded and small enough for the oracle to finish in seconds.  SA, BWT,
primary index, PLCP, LCP and freq must be bit-identical.
"""
import numpy as np
import pytest

import libsais_amd as la
from libsais_amd import workloads
from tests import oracle_lib as ol

pytestmark = pytest.mark.gpu


def fused(T, gsa=False):
    """sagpu_build32 on device buffers: (SA, U, primary, PLCP, LCP, freq)."""
    lib = la.load()
    n = len(T)
    dT = lib.sagpu_malloc(n)
    bufs = [lib.sagpu_malloc(4 * n), None if gsa else lib.sagpu_malloc(n), lib.sagpu_malloc(4 * n),
            lib.sagpu_malloc(4 * n)]
    assert lib.sagpu_memcpy_h2d(dT, T.ctypes.data, n) == 0
    f64 = np.zeros(256, np.int64)
    r = lib.sagpu_build32(dT, n, la.SAGPU_FLAG_GSA if gsa else 0, bufs[0], bufs[1], bufs[2], bufs[3],
                          f64.ctypes.data)
    assert r >= 0, la.last_error()
    outs = []
    for b, dt in zip(bufs, (np.int32, np.uint8, np.int32, np.int32)):
        if b is None:
            outs.append(None)
            continue
        o = np.empty(n, dt)
        lib.sagpu_memcpy_d2h(o.ctypes.data, b, o.nbytes)
        lib.sagpu_free(b)
        outs.append(o)
    lib.sagpu_free(dT)
    return outs[0], outs[1], r, outs[2], outs[3], f64


def _cases():
    rng = np.random.default_rng(1234)
    cases = []
    for n in (2, 3, 5, 17, 64, 1000, 4095, 4096, 4097, 65536, 200003):
        cases.append((f"acgt{n}", workloads.iid(n, seed=n, alphabet=b"ACGT")))
    for k in (1, 2, 3, 5, 20, 255, 256):
        n = 50000
        cases.append((f"bytes_k{k}", rng.integers(0, k, n).astype(np.uint8)))
    # every symbol width of the packed code stream (bits 4, 6, 7, 8) at radix tile
    # boundaries (8192- and 16384-key tiles)
    for k, n in ((16, 16383), (16, 16385), (40, 32768), (100, 16384), (129, 8193)):
        cases.append((f"bytes_k{k}_n{n}", rng.integers(0, k, n).astype(np.uint8)))
    for n in (16383, 16384, 16385, 49153):
        cases.append((f"acgt{n}", workloads.iid(n, seed=n, alphabet=b"ACGT")))
    cases.append(("bytes_with_zeros", rng.integers(0, 3, 30000).astype(np.uint8)))
    cases.append(("all_a_100", np.full(100, 65, np.uint8)))
    cases.append(("all_a_70000", np.full(70000, 65, np.uint8)))
    cases.append(("all_zero_5000", np.zeros(5000, np.uint8)))
    cases.append(("periodic17", workloads.periodic(100000, 17, 3)))
    cases.append(("fibonacci", workloads.fibonacci(100000)))
    cases.append(("dna_like_1m", workloads.dna_like(1 << 20, seed=2)))
    for run in (255, 256, 257, 300, 2100):
        T = workloads.iid(20000, seed=run, alphabet=b"ACGT")
        T[5000:5000 + run] = 65
        T[12000:12000 + run] = 65
        cases.append((f"runs{run}", T))
    rep = workloads.iid(3000, seed=99, alphabet=b"ACGT")
    cases.append(("copies300", np.tile(rep, 300)))
    # two copies of a 1.5 MiB block: lcps of up to 1.5 M symbols, past what one
    # workgroup of the long-match pass takes before the very-long pass splits them
    cases.append(("copies2_1536k", np.tile(workloads.iid(3 << 19, seed=98, alphabet=b"ACGT"), 2)))
    cases.append(("banana", np.frombuffer(b"banana", np.uint8).copy()))
    cases.append(("mississippi", np.frombuffer(b"mississippi", np.uint8).copy()))
    return cases


CASES = _cases()


@pytest.mark.parametrize("name,T", CASES, ids=[c[0] for c in CASES])
def test_sa_bwt_plcp_lcp(name, T):
    print(f"case {name} n={len(T)}", flush=True)
    n = len(T)
    rc, SA_ref = ol.sa(T)
    assert rc == 0
    SA = np.zeros(n, np.int32)
    freq = np.zeros(256, np.int32)
    assert la.libsais(T, SA, n, 0, freq) == 0, la.last_error()
    assert np.array_equal(SA, SA_ref)
    assert np.array_equal(freq.astype(np.int64), ol.freq(T))

    U_ref, idx_ref = ol.bwt(T, SA_ref)
    U = np.zeros(n, np.uint8)
    A = np.zeros(n, np.int32)
    idx = la.libsais_bwt(T, U, A, n, 0, None)
    assert idx == idx_ref
    assert np.array_equal(U, U_ref)

    P_ref = ol.plcp(T, SA_ref)
    P = np.zeros(n, np.int32)
    assert la.libsais_plcp(T, SA, P, n) == 0
    assert np.array_equal(P, P_ref)

    L = np.zeros(n, np.int32)
    assert la.libsais_lcp(P, SA, L, n) == 0
    L_ref = ol.lcp(P_ref, SA_ref)
    assert np.array_equal(L, L_ref)

    if n >= 2:
        fSA, fU, fidx, fP, fL, ff = fused(T)
        assert np.array_equal(fSA, SA_ref)
        assert fidx == idx_ref and np.array_equal(fU, U_ref)
        assert np.array_equal(fP, P_ref)
        assert np.array_equal(fL, L_ref)
        assert np.array_equal(ff, ol.freq(T))


def _gsa_set(rng, count, maxlen, alphabet=b"ACGT"):
    al = np.frombuffer(alphabet, np.uint8)
    parts = []
    for _ in range(count):
        L = int(rng.integers(1, maxlen + 1))
        parts.append(rng.choice(al, L))
        parts.append(np.zeros(1, np.uint8))
    return np.concatenate(parts)


GSA_CASES = [
    ("gsa_small", _gsa_set(np.random.default_rng(5), 10, 20)),
    ("gsa_many", _gsa_set(np.random.default_rng(6), 3000, 40)),
    ("gsa_binary", _gsa_set(np.random.default_rng(7), 500, 200, b"ab")),
    ("gsa_repeats", np.tile(np.frombuffer(b"ACGTACGT\x00", np.uint8), 2000)),
    ("gsa_single", np.frombuffer(b"banana\x00", np.uint8).copy()),
]


@pytest.mark.parametrize("name,T", GSA_CASES, ids=[c[0] for c in GSA_CASES])
def test_gsa(name, T):
    print(f"case {name} n={len(T)}", flush=True)
    n = len(T)
    rc, SA_ref = ol.sa(T, gsa=True)
    assert rc == 0
    SA = np.zeros(n, np.int32)
    assert la.libsais_gsa(T, SA, n, 0, None) == 0, la.last_error()
    assert np.array_equal(SA, SA_ref)
    P = np.zeros(n, np.int32)
    assert la.libsais_plcp_gsa(T, SA, P, n) == 0
    P_ref = ol.plcp(T, SA_ref, gsa=True)
    assert np.array_equal(P, P_ref)
    fSA, _, _, fP, fL, _ = fused(T, gsa=True)
    assert np.array_equal(fSA, SA_ref)
    assert np.array_equal(fP, P_ref)
    assert np.array_equal(fL, ol.lcp(P_ref, SA_ref))


def test_gsa_invalid_on_gpu():
    for bad in (b"\x00ab\x00", b"ab\x00\x00cd\x00", b"a\x00\x00"):
        T = np.frombuffer(bad, np.uint8).copy()
        SA = np.zeros(len(T), np.int32)
        assert la.libsais_gsa(T, SA, len(T), 0, None) == -1


def test_64bit_delegation():
    T = workloads.iid(30000, seed=9, alphabet=b"ACGT")
    n = len(T)
    rc, SA_ref = ol.sa(T)
    SA = np.zeros(n, np.int64)
    f = np.zeros(256, np.int64)
    assert la.libsais64(T, SA, n, 0, f) == 0
    assert np.array_equal(SA, SA_ref.astype(np.int64))
    assert np.array_equal(f, ol.freq(T))
    U = np.zeros(n, np.uint8)
    A = np.zeros(n, np.int64)
    U_ref, idx_ref = ol.bwt(T, SA_ref)
    assert la.libsais64_bwt(T, U, A, n, 0, None) == idx_ref
    assert np.array_equal(U, U_ref)
    P = np.zeros(n, np.int64)
    assert la.libsais64_plcp(T, SA, P, n) == 0
    P_ref = ol.plcp(T, SA_ref)
    assert np.array_equal(P, P_ref.astype(np.int64))
    L = np.zeros(n, np.int64)
    assert la.libsais64_lcp(P, SA, L, n) == 0
    assert np.array_equal(L, ol.lcp(P_ref, SA_ref).astype(np.int64))


def test_bwt_output_aliases_input():
    T = workloads.iid(10000, seed=3, alphabet=b"ACGT")
    U_ref, idx_ref = ol.bwt(T, ol.sa(T)[1])
    buf = T.copy()
    A = np.zeros(len(T), np.int32)
    assert la.libsais_bwt(buf, buf, A, len(T), 0, None) == idx_ref
    assert np.array_equal(buf, U_ref)


def test_lcp_aliases_sa():
    T = workloads.iid(10000, seed=4, alphabet=b"ACGT")
    SA = ol.sa(T)[1].copy()
    P = ol.plcp(T, SA)
    L_ref = ol.lcp(P, SA)
    assert la.libsais_lcp(P, SA, SA, len(T)) == 0
    assert np.array_equal(SA, L_ref)


def test_repeatable():
    T = workloads.dna_like(1 << 18, seed=77)
    a = la.suffix_array(T)
    b = la.suffix_array(T)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("variant", [1, 2])
@pytest.mark.parametrize("name", ["acgt200003", "bytes_k256", "all_a_70000", "dna_like_1m", "banana"])
def test_bwt_variants(name, variant):
    """Scatter (1) and plain gather (2) against the oracle; the default (0, seeded
    from the initial sort) is what every other BWT test runs."""
    T = dict(CASES)[name]
    lib = la.load()
    assert lib.sagpu_set_option(b"bwt_variant", variant) == 0
    try:
        U_ref, idx_ref = ol.bwt(T, ol.sa(T)[1])
        U = np.zeros(len(T), np.uint8)
        A = np.zeros(len(T), np.int32)
        assert la.libsais_bwt(T, U, A, len(T), 0, None) == idx_ref
        assert np.array_equal(U, U_ref)
        fSA, fU, fidx, _, _, _ = fused(T)
        assert fidx == idx_ref and np.array_equal(fU, U_ref)
    finally:
        assert lib.sagpu_set_option(b"bwt_variant", 0) == 0


@pytest.mark.parametrize("name", ["acgt200003", "bytes_k256", "all_a_70000", "dna_like_1m", "banana"])
def test_plcp_without_pair_words(name):
    """The fused PLCP pass with separate gathers (plcp_pairs = 0) against the oracle;
    the default pair-word form is what the other fused tests run."""
    T = dict(CASES)[name]
    lib = la.load()
    assert lib.sagpu_set_option(b"plcp_pairs", 0) == 0
    try:
        rc, SA_ref = ol.sa(T)
        P_ref = ol.plcp(T, SA_ref)
        fSA, fU, fidx, fP, fL, _ = fused(T)
        assert np.array_equal(fSA, SA_ref) and np.array_equal(fP, P_ref)
        assert np.array_equal(fL, ol.lcp(P_ref, SA_ref))
    finally:
        assert lib.sagpu_set_option(b"plcp_pairs", 1) == 0


@pytest.mark.parametrize("name", ["acgt200003", "bytes_k256", "all_a_70000", "dna_like_1m", "banana"])
def test_keys_written_before_the_sort(name):
    """key_stream = 0: the initial keys are written by their own kernel instead of
    being generated inside the first radix pass (the default every other test runs)."""
    T = dict(CASES)[name]
    lib = la.load()
    assert lib.sagpu_set_option(b"key_stream", 0) == 0
    try:
        rc, SA_ref = ol.sa(T)
        P_ref = ol.plcp(T, SA_ref)
        U_ref, idx_ref = ol.bwt(T, SA_ref)
        fSA, fU, fidx, fP, fL, _ = fused(T)
        assert np.array_equal(fSA, SA_ref) and np.array_equal(fP, P_ref)
        assert fidx == idx_ref and np.array_equal(fU, U_ref)
    finally:
        assert lib.sagpu_set_option(b"key_stream", 1) == 0


@pytest.mark.parametrize("name", ["acgt200003", "bytes_k256", "all_a_70000", "dna_like_1m", "banana", "acgt17"])
@pytest.mark.parametrize("option", [b"raw_round1", b"stream_compare", b"exact_round1", b"plcp_hint"])
def test_stream_shortcuts_off(name, option):
    """raw_round1 = 0: the first doubling round orders groups by rank (the default
    orders them by the packed key that follows, read from the code stream);
    stream_compare = 0: the PLCP pass compares bytes instead of stream symbols;
    exact_round1 = 0: round 1 marks its splits with lower bounds instead of the
    exact lcps its raw keys give; plcp_hint = 0: no PLCP hints in the ISA words,
    every position of the PLCP pass reads its pair word."""
    T = dict(CASES)[name]
    lib = la.load()
    assert lib.sagpu_set_option(option, 0) == 0
    try:
        rc, SA_ref = ol.sa(T)
        P_ref = ol.plcp(T, SA_ref)
        fSA, fU, fidx, fP, fL, _ = fused(T)
        assert np.array_equal(fSA, SA_ref) and np.array_equal(fP, P_ref)
        assert np.array_equal(fL, ol.lcp(P_ref, SA_ref))
    finally:
        assert lib.sagpu_set_option(option, 1) == 0


@pytest.mark.parametrize("bucketed", [1, 0])
def test_plcp_api_phi_forms(bucketed):
    """libsais_plcp computes Φ of the caller's SA with the bucketed two-pass scatter
    (default) or one pass; both match the oracle.  An SA with repeated and
    out-of-range entries has no defined PLCP: the call still returns 0 and touches
    nothing out of bounds (both kernels skip or drop such entries)."""
    lib = la.load()
    assert lib.sagpu_set_option(b"bucketed_phi", bucketed) == 0
    try:
        for name in ("acgt200003", "dna_like_1m", "all_a_70000", "bytes_k256"):
            T = dict(CASES)[name]
            rc, SA = ol.sa(T)
            P = np.zeros(len(T), np.int32)
            assert la.libsais_plcp(T, SA, P, len(T)) == 0
            assert np.array_equal(P, ol.plcp(T, SA))
        T = dict(CASES)["acgt200003"]
        bad = ol.sa(T)[1].copy()
        bad[100:5000] = bad[7]            # repeated targets
        bad[6000:6100] = len(T) + 12345   # out of range
        P = np.zeros(len(T), np.int32)
        assert la.libsais_plcp(T, bad, P, len(T)) == 0
    finally:
        assert lib.sagpu_set_option(b"bucketed_phi", 1) == 0


@pytest.mark.parametrize("n", [2, 3, 5, 4097, 100001])
def test_bwt_unaligned_output(n):
    """sagpu_build32 with a BWT output that is not 4-byte aligned."""
    T = workloads.iid(n, seed=n + 3, alphabet=b"ACGT")
    U_ref, idx_ref = ol.bwt(T, ol.sa(T)[1])
    lib = la.load()
    dT = lib.sagpu_malloc(n)
    dSA = lib.sagpu_malloc(4 * n)
    dUb = lib.sagpu_malloc(n + 8)
    try:
        assert lib.sagpu_memcpy_h2d(dT, T.ctypes.data, n) == 0
        assert lib.sagpu_memset(dUb, 0x5A, n + 8) == 0
        r = lib.sagpu_build32(dT, n, 0, dSA, dUb + 1, None, None, None)
        assert r == idx_ref, la.last_error()
        out = np.empty(n + 8, np.uint8)
        lib.sagpu_memcpy_d2h(out.ctypes.data, dUb, n + 8)
        assert np.array_equal(out[1:n + 1], U_ref)
        assert out[0] == 0x5A and np.all(out[n + 1:] == 0x5A)   # nothing written outside
    finally:
        for p in (dT, dSA, dUb):
            lib.sagpu_free(p)


@pytest.mark.parametrize("wlog", [12, 20, 32])
def test_isa_window_option(wlog):
    T = workloads.dna_like(3 << 20, seed=17)
    lib = la.load()
    assert lib.sagpu_set_option(b"isa_window_log2", wlog) == 0
    try:
        rc, SA_ref = ol.sa(T)
        assert np.array_equal(la.suffix_array(T), SA_ref)
    finally:
        assert lib.sagpu_set_option(b"isa_window_log2", 25) == 0


@pytest.mark.parametrize("name", ["dna_like_5M", "periodic_6M", "gsa_5M"])
def test_windowed_scatter_option(name):
    """Texts over 2^22 write ISA through windowed scatters (initial ranks and the
    rounds' rank updates); the fused build must equal the oracle with the option
    on and off."""
    if name == "dna_like_5M":
        T = workloads.dna_like(5 << 20, seed=19)
    elif name == "periodic_6M":
        T = workloads.periodic(6 << 20, 37, 5)
    else:
        rng = np.random.default_rng(20)
        parts = []
        while sum(len(p) for p in parts) < (5 << 20):
            parts.append(rng.integers(1, 5, int(rng.integers(1, 20000))).astype(np.uint8))
            parts.append(np.zeros(1, np.uint8))
        T = np.concatenate(parts)
    gsa = name.startswith("gsa")
    rc, SA_ref = ol.sa(T, gsa=gsa) if gsa else ol.sa(T)
    assert rc == 0
    P_ref = ol.plcp(T, SA_ref, gsa=gsa) if gsa else ol.plcp(T, SA_ref)
    L_ref = ol.lcp(P_ref, SA_ref)
    lib = la.load()
    try:
        # windowed on with the partition pass's closed-form histogram, on with a
        # counted one, and off
        for on, perm in ((1, 1), (1, 0), (0, 1)):
            print(f"[windowed] {name} n={len(T)} windowed_scatter={on} perm_hist={perm}", flush=True)
            assert lib.sagpu_set_option(b"windowed_scatter", on) == 0
            assert lib.sagpu_set_option(b"perm_hist", perm) == 0
            SA, U, r, P, L, f = fused(T, gsa=gsa)
            assert np.array_equal(SA, SA_ref), (name, on, perm)
            assert np.array_equal(P, P_ref) and np.array_equal(L, L_ref), (name, on, perm)
            if not gsa:
                U_ref, idx_ref = ol.bwt(T, SA_ref)
                assert r == idx_ref and np.array_equal(U, U_ref), (name, on, perm)
    finally:
        assert lib.sagpu_set_option(b"windowed_scatter", 1) == 0
        assert lib.sagpu_set_option(b"perm_hist", 1) == 0


@pytest.mark.parametrize("bits", [4, 11])
def test_radix_bits_option(bits):
    """4- and 11-bit radix digits (the initial sort, its generating first pass, the
    big-group sorts and the windowed ISA scatter's partition) give the same build."""
    lib = la.load()
    rng = np.random.default_rng(bits)
    cases = [("dna_like_5M", workloads.dna_like(5 << 20, seed=27), False),
             ("periodic_300k", workloads.periodic(300000, 19, 4), False),
             ("bytes_k200", rng.integers(0, 200, 100000).astype(np.uint8), False)]
    parts = []
    for _ in range(400):
        parts.append(rng.integers(1, 5, int(rng.integers(1, 3000))).astype(np.uint8))
        parts.append(np.zeros(1, np.uint8))
    cases.append(("gsa_set", np.concatenate(parts), True))
    try:
        assert lib.sagpu_set_option(b"radix_bits", bits) == 0
        for name, T, gsa in cases:
            print(f"[radix_bits] {bits}: {name} n={len(T)}", flush=True)
            rc, SA_r = ol.sa(T, gsa=gsa) if gsa else ol.sa(T)
            assert rc == 0
            P_r = ol.plcp(T, SA_r, gsa=gsa) if gsa else ol.plcp(T, SA_r)
            SA, U, r, P, L, f = fused(T, gsa=gsa)
            assert np.array_equal(SA, SA_r), name
            assert np.array_equal(P, P_r) and np.array_equal(L, ol.lcp(P_r, SA_r)), name
            if not gsa:
                U_r, idx_r = ol.bwt(T, SA_r)
                assert r == idx_r and np.array_equal(U, U_r), name
    finally:
        assert lib.sagpu_set_option(b"radix_bits", 8) == 0


def test_round_graphs_replay_across_builds():
    """Rounds from the second on replay captured graphs, updated in place from one
    build to the next (other pointers and list lengths) and between buffer parities.
    Consecutive builds of different texts, with graphs on and off, all equal the
    oracle; the big-group path runs between replays on the repeat-rich texts."""
    lib = la.load()
    texts = [("dna_like_1M", workloads.dna_like(1 << 20, seed=23)),
             ("acgt_300k", workloads.iid(300001, seed=24, alphabet=b"ACGT")),
             ("copies", np.tile(workloads.iid(7000, seed=25, alphabet=b"ACGT"), 40)),
             ("periodic_200k", workloads.periodic(200000, 23, 7)),
             ("dna_like_1M_again", workloads.dna_like(1 << 20, seed=23))]
    ref = {}
    for name, T in texts:
        rc, SA = ol.sa(T)
        assert rc == 0
        P = ol.plcp(T, SA)
        ref[name] = (SA, P, ol.lcp(P, SA), ol.bwt(T, SA))
    try:
        for on in (1, 0, 1):
            assert lib.sagpu_set_option(b"round_graphs", on) == 0
            for name, T in texts:
                SA, U, r, P, L, f = fused(T)
                SA_r, P_r, L_r, (U_r, idx_r) = ref[name]
                st = la.stats()
                print(f"[graphs] {name} round_graphs={on}: rounds {st.rounds}, graphs captured {st.round_graphs}, "
                      f"replays {st.round_graph_launches}", flush=True)
                assert st.round_graph_launches == (max(0, st.rounds - 1) if on else 0), (name, st.rounds)
                assert np.array_equal(SA, SA_r), (name, on)
                assert r == idx_r and np.array_equal(U, U_r), (name, on)
                assert np.array_equal(P, P_r) and np.array_equal(L, L_r), (name, on)
    finally:
        assert lib.sagpu_set_option(b"round_graphs", 1) == 0


def _pivot_cases():
    rng = np.random.default_rng(31)
    return [
        ("all_a_300k", np.full(300000, ord("A"), np.uint8), True),
        ("periodic17_1M", workloads.periodic(1 << 20, 17, 3), True),
        ("periodic3_500k", workloads.periodic(500000, 3, 9), True),
        ("fibonacci_1M", workloads.fibonacci(1 << 20), True),
        ("copies_280k", np.tile(workloads.iid(7000, seed=25, alphabet=b"ACGT"), 40), False),
        ("ab_runs_400k", np.repeat(rng.integers(65, 67, 4000).astype(np.uint8), 100), False),
        ("dna_like_1M", workloads.dna_like(1 << 20, seed=23), False),
        ("binary_iid_600k", workloads.iid(600000, seed=32, alphabet=b"01"), False),
    ]


@pytest.mark.parametrize("name,T,degenerate", _pivot_cases(), ids=[c[0] for c in _pivot_cases()])
@pytest.mark.parametrize("wide", [0, 1])
@pytest.mark.parametrize("flags", [1, 0])
def test_big_pivots(name, T, degenerate, wide, flags):
    """Big groups with few distinct keys: up to 4 keys per group frequent in a sample
    skip the sort (the pivot classes) and the rest is sorted and ranked around them.
    Forced on at every size (big_pivot_min = 1), in the 32-bit and the wide (u64
    key2) engine, with group words flagged for stored ranks (the default) and
    without, the fused build equals the oracle; on repeat-rich texts the path
    runs."""
    lib = la.load()
    rc, SA_ref = ol.sa(T)
    assert rc == 0
    P_ref = ol.plcp(T, SA_ref)
    L_ref = ol.lcp(P_ref, SA_ref)
    U_ref, idx_ref = ol.bwt(T, SA_ref)
    try:
        assert lib.sagpu_set_option(b"big_pivot_min", 1) == 0
        assert lib.sagpu_set_option(b"group_flags", flags) == 0
        if wide:
            assert lib.sagpu_set_option(b"wide_from", 0) == 0
        print(f"[pivots] {name} n={len(T)} wide={wide} flags={flags}", flush=True)
        SA, U, r, P, L, f = fused(T)
        st = la.stats()
        print(f"[pivots] {name}: rounds {st.rounds}, big rounds {st.big_rounds}, pivot rounds "
              f"{st.big_pivot_rounds}, pivot members {st.big_pivot_elems}", flush=True)
        assert np.array_equal(SA, SA_ref), name
        assert r == idx_ref and np.array_equal(U, U_ref), name
        assert np.array_equal(P, P_ref) and np.array_equal(L, L_ref), name
        if degenerate:
            assert st.big_pivot_rounds > 0 and st.big_pivot_elems > 0, name
    finally:
        assert lib.sagpu_set_option(b"big_pivot_min", 1 << 20) == 0
        assert lib.sagpu_set_option(b"group_flags", 1) == 0
        assert lib.sagpu_set_option(b"wide_from", 0x7FFFFFFF) == 0


@pytest.mark.parametrize("name", ["acgt200003", "bytes_k256", "all_a_70000", "dna_like_1m", "banana", "acgt17"])
def test_group_flags_off(name):
    """Every group ranked at its first slot (group_flags = 0): the fused build
    equals the oracle as with the default stored ranks of flagged groups."""
    lib = la.load()
    T = dict(CASES)[name]
    rc, SA_ref = ol.sa(T)
    assert rc == 0
    P_ref = ol.plcp(T, SA_ref)
    try:
        assert lib.sagpu_set_option(b"group_flags", 0) == 0
        SA, U, r, P, L, f = fused(T)
        assert np.array_equal(SA, SA_ref), name
        U_ref, idx_ref = ol.bwt(T, SA_ref)
        assert r == idx_ref and np.array_equal(U, U_ref), name
        assert np.array_equal(P, P_ref) and np.array_equal(L, ol.lcp(P_ref, SA_ref)), name
    finally:
        assert lib.sagpu_set_option(b"group_flags", 1) == 0


@pytest.mark.parametrize("name", ["acgt5", "acgt17", "acgt4097", "bytes_k2", "bytes_k5", "bytes_k16_n16385",
                                  "bytes_k256", "dna_like_1m", "all_a_70000"])
@pytest.mark.parametrize("option,value", [(b"hist_windows", 0), (b"init_rows", 16)])
def test_initial_sort_forms(name, option, value):
    """The radix histograms counted per pass instead of from the stream's 8-bit
    windows (hist_windows = 0), and the initial grouping in 4096-position tiles
    (init_rows = 16): the fused build equals the oracle.  (The closed-form partition
    histogram, perm_hist, runs above 2^22 suffixes: test_windowed_scatter_option.)"""
    lib = la.load()
    T = dict(CASES)[name]
    rc, SA_ref = ol.sa(T)
    assert rc == 0
    P_ref = ol.plcp(T, SA_ref)
    default = {b"hist_windows": 1, b"init_rows": 8}[option]
    try:
        assert lib.sagpu_set_option(option, value) == 0
        SA, U, r, P, L, f = fused(T)
        assert np.array_equal(SA, SA_ref), name
        U_ref, idx_ref = ol.bwt(T, SA_ref)
        assert r == idx_ref and np.array_equal(U, U_ref), name
        assert np.array_equal(P, P_ref) and np.array_equal(L, ol.lcp(P_ref, SA_ref)), name
    finally:
        assert lib.sagpu_set_option(option, default) == 0


def test_big_pivots_gsa():
    """The pivot path on a string set (separators end the repeats): equal to the
    oracle's GSA outputs."""
    lib = la.load()
    unit = workloads.iid(3000, seed=33, alphabet=b"ACGT")
    parts = []
    for k in range(60):
        parts.append(np.tile(unit[: 1000 + 17 * k], 3))
        parts.append(np.zeros(1, np.uint8))
    T = np.concatenate(parts)
    rc, SA_ref = ol.sa(T, gsa=True)
    assert rc == 0
    P_ref = ol.plcp(T, SA_ref, gsa=True)
    L_ref = ol.lcp(P_ref, SA_ref)
    try:
        assert lib.sagpu_set_option(b"big_pivot_min", 1) == 0
        SA, U, r, P, L, f = fused(T, gsa=True)
        st = la.stats()
        print(f"[pivots] gsa n={len(T)}: big rounds {st.big_rounds}, pivot rounds {st.big_pivot_rounds}", flush=True)
        assert np.array_equal(SA, SA_ref)
        assert np.array_equal(P, P_ref) and np.array_equal(L, L_ref)
    finally:
        assert lib.sagpu_set_option(b"big_pivot_min", 1 << 20) == 0


def _round1_cases():
    rng = np.random.default_rng(4242)

    def sym(alpha, k):
        return np.frombuffer(alpha, np.uint8)[rng.integers(0, len(alpha), k)]

    cases = []
    # a 16-symbol prefix repeated 3000 times with random continuations: one initial
    # group far above the LDS tile size, split in round 1 by the global sort
    base = sym(b"ACGT", 16)
    cases.append(("big_group", np.concatenate([np.concatenate([base, sym(b"ACGT", 24)]) for _ in range(3000)])))
    # the text ends inside a repeat of earlier text: the last suffixes share their
    # initial groups with suffixes whose round-1 keys are not padded
    R = sym(b"ACGT", 48)
    for cut in (48, 40, 33, 31, 20, 17, 16, 9):
        cases.append((f"tail_repeat_{cut}", np.concatenate([sym(b"ACGT", 5000), R, sym(b"ACGT", 3000), R[:cut]])))
    # alphabet widths (bits 1, 2, 3, 5, 8) with repeats so that round 1 splits groups
    for alpha in (b"ab", b"ACGT", b"ACGTN", b"abcdefghijklmnopqrstu", bytes(range(256))):
        unit = sym(alpha, 64)
        parts = []
        for _ in range(400):
            parts.append(unit[: int(rng.integers(20, 64))])
            parts.append(sym(alpha, int(rng.integers(1, 30))))
        cases.append((f"alpha{len(alpha)}", np.concatenate(parts)))
    return cases


@pytest.mark.parametrize("exact", [1, 0])
@pytest.mark.parametrize("name,T", _round1_cases(), ids=[c[0] for c in _round1_cases()])
def test_round1_exact_lcps(name, T, exact):
    """Round-1 splits on raw keys: exact lcps (default) and lower bounds both give
    the oracle's SA, BWT, PLCP and LCP, also for groups split by the global sort
    and for texts whose last suffixes repeat earlier text."""
    lib = la.load()
    assert lib.sagpu_set_option(b"exact_round1", exact) == 0
    try:
        rc, SA_ref = ol.sa(T)
        P_ref = ol.plcp(T, SA_ref)
        U_ref, idx_ref = ol.bwt(T, SA_ref)
        fSA, fU, fidx, fP, fL, _ = fused(T)
        assert np.array_equal(fSA, SA_ref)
        assert np.array_equal(fU, U_ref) and fidx == idx_ref
        assert np.array_equal(fP, P_ref)
        assert np.array_equal(fL, ol.lcp(P_ref, SA_ref))
    finally:
        assert lib.sagpu_set_option(b"exact_round1", 1) == 0
