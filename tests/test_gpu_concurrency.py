"""Concurrent callers of the drop-in API.

The reference is reentrant: every call works in its own buffers, and a context
serves one thread (include/libsais.h:54, src/libsais.c:7008-7016).  Here calls
without a context take one of the engine's slot contexts (a stream and a pool of
device buffers each), and every libsais context owns an engine context of its own,
so calls from several host threads run at the same time on the device.

Eight host threads call libsais + libsais_bwt at once, first without contexts, then
each through a context of its own (libsais_ctx / libsais_bwt_ctx).  Every SA, BWT
and primary index equals the reference libsais's on the same text, and the eight
concurrent calls take less wall time than eight calls one after another.
"""
import threading
import time

import numpy as np
import pytest

import libsais_amd as la
from libsais_amd import workloads
from tests import oracle_lib as ol

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not ol.have_ref(), reason="compiled reference not built")]

THREADS = 8
N = 4 << 20


@pytest.fixture(scope="module")
def texts():
    out = []
    for k in range(THREADS):
        T = workloads.dna_like(N + 977 * k, seed=100 + k)
        rc, SA, f = ol.ref_sa(T, threads=4)
        assert rc == 0
        U, idx = ol.ref_bwt(T, threads=4)
        out.append((T, SA, f, U, idx))
    return out


def _call(T, use_ctx):
    n = len(T)
    SA = np.empty(n, np.int32)
    f = np.zeros(256, np.int32)
    U = np.empty(n, np.uint8)
    A = np.empty(1, np.int32)
    if use_ctx:
        ctx = la.libsais_create_ctx()
        assert ctx
        try:
            rc = la.libsais_ctx(ctx, T, SA, n, 0, f)
            idx = la.libsais_bwt_ctx(ctx, T, U, A, n, 0, None)
        finally:
            la.libsais_free_ctx(ctx)
    else:
        rc = la.libsais(T, SA, n, 0, f)
        idx = la.libsais_bwt(T, U, A, n, 0, None)
    return rc, SA, f, U, idx


def _run_all(texts, use_ctx):
    results = [None] * len(texts)
    errors = []

    def work(k):
        try:
            results[k] = _call(texts[k][0], use_ctx)
        except Exception as e:  # reported below with the thread's index
            errors.append((k, repr(e)))

    ths = [threading.Thread(target=work, args=(k,)) for k in range(len(texts))]
    t0 = time.perf_counter()
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    wall = time.perf_counter() - t0
    assert not errors, errors
    return results, wall


def _check(texts, results):
    for k, ((T, SA_r, f_r, U_r, idx_r), (rc, SA, f, U, idx)) in enumerate(zip(texts, results)):
        assert rc == 0, (k, la.last_error())
        assert np.array_equal(SA, SA_r), k
        assert np.array_equal(f, f_r), k
        assert idx == idx_r, k
        assert np.array_equal(U, U_r), k


@pytest.mark.parametrize("use_ctx", [False, True], ids=["no_context", "context_per_thread"])
def test_concurrent_calls_match_reference(texts, use_ctx):
    print(f"[concurrency] {THREADS} threads, n={N}, contexts={use_ctx}", flush=True)
    _run_all(texts, use_ctx)                       # warm-up: contexts and pools created
    results, wall = _run_all(texts, use_ctx)
    _check(texts, results)
    t0 = time.perf_counter()
    for T, *_ in texts:
        rc, *_ = _call(T, use_ctx)
        assert rc == 0
    serial = time.perf_counter() - t0
    print(f"[concurrency] {THREADS} concurrent: {wall * 1e3:.1f} ms, one after another: {serial * 1e3:.1f} ms "
          f"({serial / wall:.2f}x)", flush=True)
    assert wall < serial, (wall, serial)
