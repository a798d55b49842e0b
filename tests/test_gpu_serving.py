"""Serving: jobs admitted while others run, small ones batched, slots overlapping on
the device -- and every job's result identical to running it alone."""
import numpy as np
import pytest

from libsais_amd import serving, workloads
from tests import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _mix(seed=3, count=60):
    rng = np.random.default_rng(seed)
    reqs = []
    for q in range(count):
        size = int(np.exp(rng.uniform(np.log(2), np.log(2 << 20))))
        T = workloads.markov(size, seed=1000 + q, order=2)
        if q % 9 == 0:
            T = rng.integers(0, 3, size).astype(np.uint8)  # contains 0 bytes: cannot be batched
        reqs.append(serving.Request(T, q % 3))
    for n in (0, 1, 2):
        reqs.append(serving.Request(np.full(n, 65, np.uint8), serving.JOB_BWT))
    return reqs


def _same(a, b):
    oa, ob = a.outputs(), b.outputs()
    assert oa.keys() == ob.keys()
    for k in oa:
        if isinstance(oa[k], np.ndarray):
            assert np.array_equal(oa[k], ob[k]), k
        else:
            assert oa[k] == ob[k], k


def test_results_in_mix_equal_results_alone():
    mixed = _mix()
    alone = [serving.Request(r.T, r.kind) for r in mixed]
    with serving.Server(slots=4, batch_bytes=8 << 20) as sv:
        for r in mixed:
            sv.submit(r)
        sv.drain()
    assert all(r.job.status == 0 for r in mixed)
    assert sum(r.job.batched for r in mixed) > 10
    for r in alone:
        assert serving.run_alone(r) == 0
    for a, b in zip(mixed, alone):
        _same(a, b)
    # and against the oracle for a few
    for r in mixed[:12]:
        if len(r.T) < 2:
            continue
        rc, SA = ol.sa(r.T)
        assert np.array_equal(r.SA[:len(r.T)], SA)
        if r.kind == serving.JOB_BWT:
            U, idx = ol.bwt(r.T, SA)
            assert np.array_equal(r.U[:len(r.T)], U) and r.job.primary == idx
        if r.kind == serving.JOB_LCP:
            assert np.array_equal(r.LCP[:len(r.T)], ol.lcp(ol.plcp(r.T, SA), SA))


def test_admission_rejects_job_that_never_fits():
    small = [serving.Request(workloads.markov(5000, seed=q, order=2), serving.JOB_SA) for q in range(5)]
    big = serving.Request(workloads.markov(16 << 20, seed=77, order=2), serving.JOB_SA)
    with serving.Server(slots=2, hbm_budget=200 << 20, batch_bytes=1 << 20) as sv:
        sv.submit(big)
        for r in small:
            sv.submit(r)
        sv.drain()
    assert big.job.status == -2
    assert all(r.job.status == 0 for r in small)


def test_tight_budget_serialises_but_completes():
    reqs = [serving.Request(workloads.markov(1 << 20, seed=q, order=2), q % 3) for q in range(8)]
    alone = [serving.Request(r.T, r.kind) for r in reqs]
    budget = 80 * ((1 << 20) + 1) + (64 << 20) + 4096  # room for exactly one such job (80 n + 64 MiB each)
    with serving.Server(slots=4, hbm_budget=budget, batch_bytes=0) as sv:
        for r in reqs:
            sv.submit(r)
        for r in reqs:
            assert sv.wait(r) == 0
    for r in alone:
        assert serving.run_alone(r) == 0
    for a, b in zip(reqs, alone):
        _same(a, b)


def test_trace_replay_small():
    spec = serving.TraceSpec(jobs=80, rate=400.0, min_bytes=4 << 10, max_bytes=1 << 20, seed=6)
    reqs = serving.make_trace(spec)
    with serving.Server(slots=4) as sv:
        t0 = serving.replay(reqs, sv)
    s = serving.summarize(reqs, t0)
    assert s["jobs"] == 80 and s["p99_ms"] >= s["p50_ms"] > 0
    for r in reqs[:5]:
        rc, SA = ol.sa(r.T)
        assert np.array_equal(r.SA[:len(r.T)], SA)
