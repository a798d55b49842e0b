"""The context and OpenMP forms of the API against the same forms of the reference.

Every `_ctx` and `_omp` entry point of libsais.h and libsais64.h (reference
src/libsais.c:7065-7095, 7148-7197, 7211-7322, 8035-8112, 8436-8517;
src/libsais64.c:7213-7379, 8067-8099, 8357-8417) and the 16-bit context forms
(src/libsais16.c) is called in this library and in the reference compiled from its
sources, with the same inputs: the return codes and every output array must be
identical.  The reference's own contexts are used on its side.
"""
import numpy as np
import pytest

import libsais_amd as la
from libsais_amd import workloads
from tests import oracle_lib as ol

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not ol.have_ref(), reason="compiled reference not built")]

_SIG = {name: (res, args) for name, res, args in la.ABI}


def _ref(name):
    f = getattr(ol.ref(), name)
    f.restype, f.argtypes = _SIG[name]
    return f


def _ours(name):
    return getattr(la.load(), name)


class Ctx:
    """A context of each library: ("ctx" = libsais_create_ctx, "uctx" = the unbwt one,
    "ctx16" = libsais16_create_ctx)."""

    MAKE = {"ctx": ("libsais_create_ctx", "libsais_free_ctx"),
            "uctx": ("libsais_unbwt_create_ctx", "libsais_unbwt_free_ctx"),
            "ctx16": ("libsais16_create_ctx", "libsais16_free_ctx"),
            "uctx16": ("libsais16_unbwt_create_ctx", "libsais16_unbwt_free_ctx")}

    def __init__(self, kind):
        self.kind = kind


def _arg(a, side, ctxs):
    if isinstance(a, Ctx):
        return ctxs[a.kind]
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a


def _both(name, args, outs):
    """Runs `name` in both libraries; args[i] for i in outs are output arrays (each side
    gets a zeroed copy).  Asserts equal return codes and outputs."""
    results = []
    for side, get in (("ours", _ours), ("ref", _ref)):
        ctxs = {}
        for a in args:
            if isinstance(a, Ctx) and a.kind not in ctxs:
                make = get(Ctx.MAKE[a.kind][0]) if side == "ours" else _ref(Ctx.MAKE[a.kind][0])
                ctxs[a.kind] = make()
                assert ctxs[a.kind]
        local = [np.zeros_like(a) if i in outs else a for i, a in enumerate(args)]
        try:
            rc = get(name)(*[_arg(a, side, ctxs) for a in local])
        finally:
            for kind, c in ctxs.items():
                (get(Ctx.MAKE[kind][1]) if side == "ours" else _ref(Ctx.MAKE[kind][1]))(c)
        results.append((rc, [local[i] for i in outs]))
    (rc_o, out_o), (rc_r, out_r) = results
    assert rc_o == rc_r, (name, rc_o, rc_r, la.last_error())
    for k, (x, y) in enumerate(zip(out_o, out_r)):
        assert np.array_equal(x, y), (name, k)
    return rc_o, out_o


@pytest.fixture(scope="module")
def data():
    T = workloads.dna_like(300000, seed=12)
    n = len(T)
    rc, SA, _ = ol.ref_sa(T)
    assert rc == 0
    P = ol.ref_plcp(T, SA)
    U, idx = ol.ref_bwt(T)
    rng = np.random.default_rng(77)
    G = rng.integers(1, 5, 200000).astype(np.uint8)
    G[rng.integers(1, 199999, 300)] = 0
    G[1::1000] = 7
    G = np.concatenate([G, np.zeros(1, np.uint8)])
    G[np.flatnonzero((G[:-1] == 0) & (G[1:] == 0))] = 3   # no empty strings
    G[0] = 2
    assert ol.gsa_valid(G)
    rc, GSA, _ = ol.ref_sa(G, gsa=True)
    assert rc == 0
    Ti = rng.integers(0, 70000, 150000).astype(np.int32)
    return dict(T=T, n=n, SA=SA, P=P, U=U, idx=idx, G=G, GSA=GSA, Ti=Ti)


def test_ctx_forms_32(data):
    T, n, SA, U, idx = data["T"], data["n"], data["SA"], data["U"], data["idx"]
    G = data["G"]
    i32 = np.int32
    _both("libsais_ctx", [Ctx("ctx"), T, np.zeros(n, i32), n, 0, np.zeros(256, i32)], {2, 5})
    _both("libsais_gsa_ctx", [Ctx("ctx"), G, np.zeros(len(G), i32), len(G), 0, np.zeros(256, i32)], {2, 5})
    _both("libsais_bwt_ctx", [Ctx("ctx"), T, np.zeros(n, np.uint8), np.zeros(n, i32), n, 0, np.zeros(256, i32)],
          {2, 6})
    _both("libsais_bwt_aux_ctx", [Ctx("ctx"), T, np.zeros(n, np.uint8), np.zeros(n, i32), n, 0, None, 64,
                                  np.zeros((n - 1) // 64 + 1, i32)], {2, 8})
    _both("libsais_unbwt_ctx", [Ctx("uctx"), U, np.zeros(n, np.uint8), np.zeros(n + 1, i32), n, None, idx], {2})
    rc, (Ua, Ia) = _both("libsais_bwt_aux", [T, np.zeros(n, np.uint8), np.zeros(n, i32), n, 0, None, 256,
                                             np.zeros((n - 1) // 256 + 1, i32)], {1, 7})
    _both("libsais_unbwt_aux_ctx", [Ctx("uctx"), Ua, np.zeros(n, np.uint8), np.zeros(n + 1, i32), n, None, 256, Ia],
          {2})
    # invalid arguments are rejected the same way on both sides
    _both("libsais_ctx", [Ctx("ctx"), T, np.zeros(n, i32), -1, 0, None], {2})
    _both("libsais_gsa_ctx", [Ctx("ctx"), T, np.zeros(n, i32), n, 0, None], {2})


@pytest.mark.parametrize("threads", [0, 1, 4])
def test_omp_forms_32(data, threads):
    T, n, SA, P, U, idx = data["T"], data["n"], data["SA"], data["P"], data["U"], data["idx"]
    G, GSA, Ti = data["G"], data["GSA"], data["Ti"]
    i32 = np.int32
    _both("libsais_omp", [T, np.zeros(n, i32), n, 0, np.zeros(256, i32), threads], {1, 4})
    _both("libsais_gsa_omp", [G, np.zeros(len(G), i32), len(G), 0, np.zeros(256, i32), threads], {1, 4})
    _both("libsais_int_omp", [Ti.copy(), np.zeros(len(Ti), i32), len(Ti), 70000, 0, threads], {1})
    _both("libsais_bwt_omp", [T, np.zeros(n, np.uint8), np.zeros(n, i32), n, 0, np.zeros(256, i32), threads], {1, 5})
    _both("libsais_bwt_aux_omp", [T, np.zeros(n, np.uint8), np.zeros(n, i32), n, 0, None, 128,
                                  np.zeros((n - 1) // 128 + 1, i32), threads], {1, 7})
    _both("libsais_unbwt_omp", [U, np.zeros(n, np.uint8), np.zeros(n + 1, i32), n, None, idx, threads], {1})
    _both("libsais_plcp_omp", [T, SA, np.zeros(n, i32), n, threads], {2})
    _both("libsais_plcp_gsa_omp", [G, GSA, np.zeros(len(G), i32), len(G), threads], {2})
    rc, SAi, _ = ol.ref_sa_int(Ti, 70000)
    assert rc == 0
    _both("libsais_plcp_int_omp", [Ti, SAi, np.zeros(len(Ti), i32), len(Ti), threads], {2})
    _both("libsais_lcp_omp", [P, SA, np.zeros(n, i32), n, threads], {2})
    # a negative thread count is rejected by both
    _both("libsais_omp", [T, np.zeros(n, i32), n, 0, None, -1], {1})
    _both("libsais_plcp_omp", [T, SA, np.zeros(n, i32), n, -1], {2})


@pytest.mark.parametrize("threads", [0, 2])
def test_omp_forms_64(data, threads):
    T, n, SA, P, U, idx = data["T"], data["n"], data["SA"], data["P"], data["U"], data["idx"]
    G = data["G"]
    i64 = np.int64
    SA64, P64 = SA.astype(i64), P.astype(i64)
    _both("libsais64_omp", [T, np.zeros(n, i64), n, 0, np.zeros(256, i64), threads], {1, 4})
    _both("libsais64_gsa_omp", [G, np.zeros(len(G), i64), len(G), 0, np.zeros(256, i64), threads], {1, 4})
    Tl = data["Ti"].astype(i64)
    _both("libsais64_long_omp", [Tl.copy(), np.zeros(len(Tl), i64), len(Tl), 70000, 0, threads], {1})
    _both("libsais64_bwt_omp", [T, np.zeros(n, np.uint8), np.zeros(n, i64), n, 0, np.zeros(256, i64), threads],
          {1, 5})
    _both("libsais64_bwt_aux_omp", [T, np.zeros(n, np.uint8), np.zeros(n, i64), n, 0, None, 32,
                                    np.zeros((n - 1) // 32 + 1, i64), threads], {1, 7})
    _both("libsais64_unbwt_omp", [U, np.zeros(n, np.uint8), np.zeros(n + 1, i64), n, None, idx, threads], {1})
    _both("libsais64_plcp_omp", [T, SA64, np.zeros(n, i64), n, threads], {2})
    _both("libsais64_plcp_gsa_omp", [G, data["GSA"].astype(i64), np.zeros(len(G), i64), len(G), threads], {2})
    _both("libsais64_lcp_omp", [P64, SA64, np.zeros(n, i64), n, threads], {2})


def test_ctx_forms_16(data):
    rng = np.random.default_rng(16)
    T = rng.integers(0, 3000, 120000).astype(np.uint16)
    n = len(T)
    i32 = np.int32
    _both("libsais16_ctx", [Ctx("ctx16"), T, np.zeros(n, i32), n, 0, np.zeros(65536, i32)], {2, 5})
    G = ol.gsa_set16(rng, 300, 200, 500)
    _both("libsais16_gsa_ctx", [Ctx("ctx16"), G, np.zeros(len(G), i32), len(G), 0, np.zeros(65536, i32)], {2, 5})
    rc, (U,) = _both("libsais16_bwt_ctx", [Ctx("ctx16"), T, np.zeros(n, np.uint16), np.zeros(n, i32), n, 0, None],
                     {2})
    _both("libsais16_unbwt_ctx", [Ctx("uctx16"), U, np.zeros(n, np.uint16), np.zeros(n + 1, i32), n, None, rc], {2})
    _both("libsais16_omp", [T, np.zeros(n, i32), n, 0, np.zeros(65536, i32), 3], {1, 4})
