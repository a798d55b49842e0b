"""The C-ABI library: exports, argument validation and n <= 1 cases (CPU only).

None of these calls reaches the GPU: every one returns from the host wrapper's
checks, exactly where the reference returns (src/libsais.c:7018-7121, 8363-8432;
src/libsais64.c:7058-7170, 8302-8355).  Where the reference build exists the same
calls are made on it and the results compared.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import libsais_amd as la
from tests import oracle_lib as ol

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    names = set()
    for h in ("libsais.h", "libsais64.h", "libsais16.h", "libsais16x64.h", "sagpu.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        names |= set(re.findall(r"\b((?:libsais|sagpu)\w*)\s*\(", src))
    return names


def test_exports_every_declared_symbol():
    lib = la.load()
    out = subprocess.run(["nm", "-D", "--defined-only", la.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    declared = header_functions()
    assert declared, "no declarations parsed"
    missing = declared - exported
    assert not missing, f"declared but not exported: {sorted(missing)}"
    for name in declared:
        getattr(lib, name)


def test_no_out_of_scope_exports():
    out = subprocess.run(["nm", "-D", "--defined-only", la.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert all(name.startswith(("libsais", "sagpu")) for name in exported)
    # nothing exported beyond what the headers declare
    assert exported <= header_functions()


def test_python_binding_table_matches_headers():
    bound = {name for name, _, _ in la.ABI}
    assert header_functions() == bound


T5 = np.frombuffer(b"abcab", np.uint8).copy()


def _calls(mod, T, SA, U, A, P, L, n, fs, freq):
    """Every entry point with the given arguments; returns the list of codes."""
    return [
        mod.libsais(T, SA, n, fs, freq),
        mod.libsais_gsa(T, SA, n, fs, freq),
        mod.libsais_bwt(T, U, A, n, fs, freq),
        mod.libsais_plcp(T, SA, P, n),
        mod.libsais_plcp_gsa(T, SA, P, n),
        mod.libsais_lcp(P, SA, L, n),
    ]


def test_invalid_arguments_return_minus_one():
    SA = np.zeros(8, np.int32)
    U = np.zeros(8, np.uint8)
    A = np.zeros(8, np.int32)
    P = np.zeros(8, np.int32)
    L = np.zeros(8, np.int32)
    assert la.libsais(None, SA, 5, 0, None) == -1
    assert la.libsais(T5, None, 5, 0, None) == -1
    assert la.libsais(T5, SA, -1, 0, None) == -1
    assert la.libsais(T5, SA, 5, -1, None) == -1
    assert la.libsais_gsa(T5, SA, 5, 0, None) == -1  # T[n-1] != 0
    assert la.libsais_bwt(T5, None, A, 5, 0, None) == -1
    assert la.libsais_bwt(T5, U, None, 5, 0, None) == -1
    assert la.libsais_bwt(T5, U, A, 5, -3, None) == -1
    assert la.libsais_plcp(T5, SA, None, 5) == -1
    assert la.libsais_plcp(T5, SA, P, -2) == -1
    assert la.libsais_plcp_gsa(T5, SA, P, 5) == -1  # T[n-1] != 0
    assert la.libsais_lcp(None, SA, L, 5) == -1
    assert la.libsais_lcp(P, SA, None, 5) == -1
    S64 = np.zeros(8, np.int64)
    assert la.libsais64(None, S64, 5, 0, None) == -1
    assert la.libsais64(T5, S64, -1, 0, None) == -1
    assert la.libsais64_gsa(T5, S64, 5, 0, None) == -1
    assert la.libsais64_bwt(T5, U, None, 5, 0, None) == -1
    assert la.libsais64_plcp(T5, S64, None, 5) == -1
    assert la.libsais64_plcp_gsa(T5, S64, S64, 5) == -1
    assert la.libsais64_lcp(S64, S64, None, 5) == -1


def test_small_n_cases():
    for n in (0, 1):
        T = np.frombuffer(b"\x07\x00", np.uint8).copy()
        SA = np.full(2, 99, np.int32)
        freq = np.full(256, 5, np.int32)
        assert la.libsais(T, SA, n, 0, freq) == 0
        expect = np.zeros(256, np.int32)
        expect[7] = n
        assert np.array_equal(freq, expect)
        if n == 1:
            assert SA[0] == 0
        U = np.full(2, 9, np.uint8)
        A = np.zeros(2, np.int32)
        freq[:] = 5
        assert la.libsais_bwt(T, U, A, n, 0, freq) == n
        if n == 1:
            assert U[0] == 7 and freq[7] == 1
        P = np.full(2, 9, np.int32)
        assert la.libsais_plcp(T, SA, P, n) == 0
        if n == 1:
            assert P[0] == 0
        L = np.full(2, 9, np.int32)
        Pk = np.array([42, 1], np.int32)
        S0 = np.array([0, 1], np.int32)
        assert la.libsais_lcp(Pk, S0, L, n) == 0
        if n == 1:
            assert L[0] == 42
    G = np.zeros(1, np.uint8)
    SA = np.zeros(1, np.int32)
    assert la.libsais_gsa(G, SA, 1, 0, None) == 0 and SA[0] == 0
    S64 = np.full(1, 9, np.int64)
    f64 = np.full(256, 3, np.int64)
    assert la.libsais64(G, S64, 1, 0, f64) == 0 and S64[0] == 0 and f64[0] == 1 and f64[1:].sum() == 0


@pytest.mark.skipif(not ol.have_ref(), reason="reference build oracle/_ref absent")
def test_edge_cases_match_reference():
    r = ol.ref()

    class Ref:
        pass

    ref = Ref()
    for name in ("libsais", "libsais_gsa", "libsais_bwt", "libsais_plcp", "libsais_plcp_gsa", "libsais_lcp"):
        f = getattr(r, name)

        def mk(f):
            return lambda *a: f(*[x.ctypes.data if isinstance(x, np.ndarray) else x for x in a])

        setattr(ref, name, mk(f))
    for n in (-1, 0, 1):
        for fs in (-1, 0):
            for tbytes in (b"\x05\x00", b"\x05\x06"):
                outs = []
                for mod in (la, ref):
                    T = np.frombuffer(tbytes, np.uint8).copy()
                    SA = np.array([1, 0], np.int32)  # stays a valid index for the n == 1 lcp call
                    U = np.full(2, 77, np.uint8)
                    A = np.zeros(2, np.int32)
                    P = np.array([3, 4], np.int32)
                    L = np.full(2, 77, np.int32)
                    freq = np.full(256, 9, np.int32)
                    codes = _calls(mod, T, SA, U, A, P, L, n, fs, freq)
                    outs.append((codes, SA.tolist(), U.tolist(), P.tolist(), L.tolist(), freq.tolist()))
                assert outs[0] == outs[1], (n, fs, tbytes)


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    monkeypatch.setattr(la, "_lib", None)
    monkeypatch.setattr(la, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(la.EngineNotBuilt):
        la.libsais(T5, np.zeros(5, np.int32), 5, 0, None)


def _aux_calls(mod, T, U, A, I, n, fs, r, freq):
    """bwt_aux / unbwt / unbwt_aux with the given arguments; the codes and outputs."""
    return [
        mod.libsais_bwt_aux(T, U, A, n, fs, freq, r, I),
        mod.libsais_unbwt(T, U, A, n, None, int(I[0])),
        mod.libsais_unbwt_aux(T, U, A, n, None, r, I),
    ]


def test_aux_and_unbwt_argument_checks():
    U = np.zeros(8, np.uint8)
    A = np.zeros(9, np.int32)
    I = np.array([1, 2, 3], np.int32)
    assert la.libsais_bwt_aux(T5, U, A, 5, 0, None, 3, I) == -1   # r not a power of two
    assert la.libsais_bwt_aux(T5, U, A, 5, 0, None, 1, I) == -1   # r < 2
    assert la.libsais_bwt_aux(T5, U, A, 5, 0, None, 2, None) == -1
    assert la.libsais_unbwt_aux(T5, U, A, 5, None, 3, I) == -1
    assert la.libsais_unbwt_aux(T5, U, None, 5, None, 2, I) == -1
    bad = np.array([0, 2, 3], np.int32)                          # I[t] must lie in [1, n]
    assert la.libsais_unbwt_aux(T5, U, A, 5, None, 2, bad) == -1
    assert la.libsais_unbwt(T5, U, A, 5, None, 6) == -1
    assert la.libsais_unbwt(T5, U, A, 5, None, 0) == -1
    I64 = np.array([0, 2, 3], np.int64)
    A64 = np.zeros(9, np.int64)
    assert la.libsais64_unbwt_aux(T5, U, A64, 5, None, 2, I64) == -1
    assert la.libsais64_bwt_aux(T5, U, A64, 5, 0, None, 6, I64) == -1
    assert la.libsais64_unbwt(T5, U, A64, 5, None, -1) == -1


@pytest.mark.skipif(not ol.have_ref(), reason="reference build oracle/_ref absent")
def test_aux_edge_cases_match_reference():
    r = ol.ref()

    class Ref:
        pass

    ref = Ref()
    for name in ("libsais_bwt_aux", "libsais_unbwt", "libsais_unbwt_aux"):
        f = getattr(r, name)

        def mk(f):
            return lambda *a: f(*[x.ctypes.data if isinstance(x, np.ndarray) else x for x in a])

        setattr(ref, name, mk(f))
    for n in (-1, 0, 1):
        for fs in (-1, 0):
            for rr in (1, 2, 3, 4):
                for i0 in (0, 1, 2):
                    outs = []
                    for mod in (la, ref):
                        T = np.frombuffer(b"\x05\x06", np.uint8).copy()
                        U = np.full(2, 77, np.uint8)
                        A = np.zeros(3, np.int32)
                        I = np.array([i0, 1], np.int32)
                        freq = np.full(256, 9, np.int32)
                        codes = _aux_calls(mod, T, U, A, I, n, fs, rr, freq)
                        outs.append((codes, U.tolist(), I.tolist(), freq.tolist()))
                    assert outs[0] == outs[1], (n, fs, rr, i0)


def test_context_and_omp_forms_check_like_the_reference():
    T = np.frombuffer(b"abcab", np.uint8).copy()
    SA = np.zeros(5, np.int32)
    ctx = la.libsais_create_ctx()
    assert ctx
    assert la.libsais_ctx(None, T, SA, 5, 0, None) == -1
    assert la.libsais_ctx(ctx, None, SA, 5, 0, None) == -1
    assert la.libsais_ctx(ctx, T, SA, 1, 0, None) == 0 and SA[0] == 0
    la.libsais_free_ctx(ctx)
    la.libsais_free_ctx(None)
    assert la.libsais_create_ctx_omp(-1) is None
    c2 = la.libsais_create_ctx_omp(0)
    assert c2
    la.libsais_free_ctx(c2)
    u = la.libsais_unbwt_create_ctx()
    assert u
    la.libsais_unbwt_free_ctx(u)
    assert la.libsais_omp(T, SA, 5, 0, None, -1) == -1
    assert la.libsais_int_omp(np.zeros(5, np.int32), SA, 5, 3, 0, -1) == -1
    assert la.libsais_int(np.zeros(5, np.int32), SA, -1, 3, 0) == -1
    I = np.array([1], np.int32)
    U = np.zeros(5, np.uint8)
    A = np.zeros(6, np.int32)
    assert la.libsais_unbwt_aux_omp(T, U, A, 5, None, 2, I, -2) == -1
    assert la.libsais_plcp_omp(T, SA, SA, 5, -1) == -1
    S64 = np.zeros(5, np.int64)
    assert la.libsais64_long(np.zeros(1, np.int64), S64, 1, 5, 0) == 0 and S64[0] == 0
    assert la.libsais64_omp(T, S64, 5, 0, None, -1) == -1


# ------------------------------------------------------------ 16-bit alphabet
def test_16bit_argument_checks():
    T = np.array([5, 9, 5, 9, 5], np.uint16)
    SA = np.zeros(5, np.int32)
    U = np.zeros(5, np.uint16)
    A = np.zeros(6, np.int32)
    I = np.array([1, 2, 3], np.int32)
    assert la.libsais16(None, SA, 5, 0, None) == -1
    assert la.libsais16(T, SA, -1, 0, None) == -1
    assert la.libsais16(T, SA, 5, -1, None) == -1
    assert la.libsais16_gsa(T, SA, 5, 0, None) == -1            # T[n-1] != 0
    assert la.libsais16_bwt(T, U, None, 5, 0, None) == -1
    assert la.libsais16_bwt_aux(T, U, A, 5, 0, None, 3, I) == -1  # r not a power of two
    assert la.libsais16_unbwt_aux(T, U, A, 5, None, 2, np.array([0, 2, 3], np.int32)) == -1
    assert la.libsais16_unbwt(T, U, A, 5, None, 6) == -1
    assert la.libsais16_plcp_gsa(T, SA, SA, 5) == -1
    assert la.libsais16_lcp(None, SA, SA, 5) == -1
    assert la.libsais16_ctx(None, T, SA, 5, 0, None) == -1
    assert la.libsais16_omp(T, SA, 5, 0, None, -1) == -1
    assert la.libsais16_create_ctx_omp(-1) is None
    c = la.libsais16_create_ctx()
    assert c
    la.libsais16_free_ctx(c)
    S64 = np.zeros(5, np.int64)
    assert la.libsais16x64(T, S64, 5, -1, None) == -1
    assert la.libsais16x64_gsa(T, S64, 5, 0, None) == -1
    assert la.libsais16x64_bwt_aux(T, U, S64, 5, 0, None, 6, S64) == -1
    assert la.libsais16x64_lcp_omp(S64, S64, S64, 5, -1) == -1


def test_16bit_small_n_cases():
    T = np.array([40000, 0], np.uint16)
    for n in (0, 1):
        SA = np.full(2, 99, np.int32)
        freq = np.full(65536, 5, np.int32)
        assert la.libsais16(T, SA, n, 0, freq) == 0
        expect = np.zeros(65536, np.int32)
        expect[40000] = n
        assert np.array_equal(freq, expect)
        U = np.full(2, 9, np.uint16)
        A = np.zeros(2, np.int32)
        assert la.libsais16_bwt(T, U, A, n, 0, None) == n
        if n == 1:
            assert SA[0] == 0 and U[0] == 40000
        I = np.array([7], np.int32)
        assert la.libsais16_bwt_aux(T, U, A, n, 0, None, 4, I) == 0 and I[0] == n
        f64 = np.full(65536, 3, np.int64)
        S64 = np.full(2, 9, np.int64)
        assert la.libsais16x64(T, S64, n, 0, f64) == 0 and f64.sum() == n


@pytest.mark.skipif(not ol.have_ref(), reason="reference build oracle/_ref absent")
def test_16bit_edge_cases_match_reference():
    r = ol.ref()
    table = {name: (res, args) for name, res, args in la.ABI_16 + la.ABI_16X64}

    def ref_fn(name):
        f = getattr(r, name)
        f.restype, f.argtypes = table[name]
        return lambda *a: f(*[x.ctypes.data if isinstance(x, np.ndarray) else x for x in a])

    def calls(get, it, T, SA, U, A, P, L, I, n, fs, rr, i0, freq):
        return [
            get(f"libsais16{it}")(T, SA, n, fs, freq),
            get(f"libsais16{it}_gsa")(T, SA, n, fs, freq),
            get(f"libsais16{it}_bwt")(T, U, A, n, fs, freq),
            get(f"libsais16{it}_bwt_aux")(T, U, A, n, fs, freq, rr, I),
            get(f"libsais16{it}_unbwt")(T, U, A, n, None, i0),
            get(f"libsais16{it}_unbwt_aux")(T, U, A, n, None, rr, I),
            get(f"libsais16{it}_plcp")(T, SA, P, n),
            get(f"libsais16{it}_plcp_gsa")(T, SA, P, n),
            get(f"libsais16{it}_lcp")(P, SA, L, n),
        ]

    for it, dt in (("", np.int32), ("x64", np.int64)):
        for n in (-1, 0, 1):
            for fs in (-1, 0):
                for tv in ((5, 0), (5, 6)):
                    for rr, i0 in ((2, 1), (3, 0), (4, 2)):
                        outs = []
                        for get in (lambda name: getattr(la, name), ref_fn):
                            T = np.array(tv, np.uint16)
                            SA = np.array([1, 0], dt)
                            U = np.full(2, 77, np.uint16)
                            A = np.zeros(3, dt)
                            P = np.array([3, 4], dt)
                            L = np.full(2, 77, dt)
                            I = np.array([i0, 1], dt)
                            freq = np.full(65536, 9, dt)
                            codes = calls(get, it, T, SA, U, A, P, L, I, n, fs, rr, i0, freq)
                            outs.append((codes, SA.tolist(), U.tolist(), P.tolist(), L.tolist(), I.tolist(),
                                         freq[:8].tolist(), int(freq.sum())))
                        assert outs[0] == outs[1], (it, n, fs, tv, rr, i0)
