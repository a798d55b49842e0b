"""The batched and served configurations at their full sizes.

Batched GSA: the 20,000-string set of about 512 MiB (log-normal lengths, median
4 KiB, clipped to 256 B - 4 MiB, 10% with a shared 1-20 KiB prefix) in one
sagpu_gsa_batch32 call.  Every string's SA and LCP must equal the reference's
libsais_gsa + libsais_plcp_gsa + libsais_lcp on that string alone (reference
src/libsais.c:7034-7048, 8381-8397, 8417-8432), run string by string on 16 worker
threads.

Serving: a trace of order-2 Markov ACGT jobs from 4 KiB to 64 MiB (SA, SA+BWT,
SA+LCP), submitted all at once to the server (slots, HBM admission, continuous
batching of the small jobs).  Every job's result must equal the same job run alone
on the plain path, and a sample that includes the 64 MiB job must equal the
reference.
"""
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import libsais_amd as la
from libsais_amd import batch, serving, workloads
from tests import oracle_lib as ol

pytestmark = [pytest.mark.gpu, pytest.mark.slow,
              pytest.mark.skipif(not ol.have_ref(), reason="compiled reference not built")]

WORKERS = 16


def _log(msg):
    print(f"[fullsize {time.strftime('%H:%M:%S')}] {msg}", flush=True)


@pytest.mark.timeout(900)
def test_batched_gsa_20000_strings_against_reference():
    T, offs = workloads.gsa_batch(count=20000, seed=5)
    count = len(offs) - 1
    n = int(offs[-1])
    _log(f"batched GSA: {count} strings, {n} bytes")
    t = time.perf_counter()
    SA, L = batch.batched(T, offs)
    _log(f"sagpu_gsa_batch32 + downloads {time.perf_counter() - t:.2f}s")
    SA_r = np.empty(n, np.int32)
    L_r = np.empty(n, np.int32)

    def one(k):
        o, e = int(offs[k]), int(offs[k + 1])
        S = T[o:e]
        rc, s, _ = ol.ref_sa(S, gsa=True)
        if rc != 0:
            return k
        SA_r[o:e] = s
        L_r[o:e] = ol.ref_lcp(ol.ref_plcp(S, s, gsa=True), s)
        return None

    t = time.perf_counter()
    with ThreadPoolExecutor(WORKERS) as ex:
        bad = [k for k in ex.map(one, range(count)) if k is not None]
    _log(f"reference per string on {WORKERS} threads {time.perf_counter() - t:.2f}s")
    assert not bad, bad[:10]
    if not np.array_equal(SA, SA_r):
        k = int(np.searchsorted(offs, int(np.flatnonzero(SA != SA_r)[0]), side="right")) - 1
        pytest.fail(f"SA of string {k} differs")
    if not np.array_equal(L, L_r):
        k = int(np.searchsorted(offs, int(np.flatnonzero(L != L_r)[0]), side="right")) - 1
        pytest.fail(f"LCP of string {k} differs")
    _log(f"all {count} strings identical (SA and LCP)")


def _trace():
    spec = serving.TraceSpec(jobs=150, rate=200.0, min_bytes=4 << 10, max_bytes=64 << 20, seed=6)
    reqs = serving.make_trace(spec)
    # the largest sizes of the range, one of each output kind
    reqs.append(serving.Request(workloads.markov(64 << 20, seed=606, order=2), serving.JOB_LCP))
    reqs.append(serving.Request(workloads.markov(48 << 20, seed=607, order=2), serving.JOB_BWT))
    reqs.append(serving.Request(workloads.markov(40 << 20, seed=608, order=2), serving.JOB_SA))
    return reqs


def _same(a, b):
    assert a.keys() == b.keys()
    for key in a:
        if isinstance(a[key], np.ndarray):
            assert np.array_equal(a[key], b[key]), key
        else:
            assert a[key] == b[key], key


@pytest.mark.timeout(900)
def test_served_jobs_up_to_64MiB_equal_alone_and_reference():
    reqs = _trace()
    total = sum(len(r.T) for r in reqs)
    _log(f"serving: {len(reqs)} jobs, {total} bytes, largest {max(len(r.T) for r in reqs)}")
    for r in reqs:
        r.arrival = 0.0
    with serving.Server(slots=4) as sv:
        t = time.perf_counter()
        serving.replay(reqs, sv)
        _log(f"served all at once {time.perf_counter() - t:.2f}s, "
             f"{sum(r.job.batched for r in reqs)} jobs batched")
    served = [r.outputs() for r in reqs]
    served = [{k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in o.items()} for o in served]
    t = time.perf_counter()
    for r, s in zip(reqs, served):
        assert serving.run_alone(r) == 0, la.last_error()
        _same(r.outputs(), s)
    _log(f"every job identical alone and in the mix (alone: {time.perf_counter() - t:.2f}s)")
    order = sorted(range(len(reqs)), key=lambda q: len(reqs[q].T))
    sample = sorted(set(order[:3] + order[len(order) // 2:len(order) // 2 + 2] + order[-4:]))
    for q in sample:
        r, s = reqs[q], served[q]
        n = len(r.T)
        rc, SA_r, _ = ol.ref_sa(r.T, threads=WORKERS)
        assert rc == 0 and np.array_equal(s["SA"], SA_r), q
        if r.kind == serving.JOB_BWT:
            U_r, idx_r = ol.ref_bwt(r.T, threads=WORKERS)
            assert s["primary"] == idx_r and np.array_equal(s["U"], U_r), q
        if r.kind == serving.JOB_LCP:
            P_r = ol.ref_plcp(r.T, SA_r, threads=WORKERS)
            assert np.array_equal(s["LCP"], ol.ref_lcp(P_r, SA_r, threads=WORKERS)), q
        _log(f"job {q} (n={n}, kind={r.kind}) identical to the reference")
