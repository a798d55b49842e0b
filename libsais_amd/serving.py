"""Serving loop: Poisson-arrival SA / SA+BWT / SA+LCP jobs through the native server.

The server (sagpu_server_*, engine.hip) owns the slots, the HBM admission and the
continuous batching; this module builds the job trace, replays it with the
arrival times, and summarises latency (arrival -> results in host memory) and
aggregate throughput.  The plain path replays the same trace one job at a time on
a single stream (sagpu_run_job), which is what a caller of the libsais API gets.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass

import numpy as np

from . import last_error, load
from . import workloads

JOB_SA, JOB_BWT, JOB_LCP = 0, 1, 2


class Job(ctypes.Structure):
    _fields_ = [
        ("T", ctypes.c_void_p),
        ("n", ctypes.c_int64),
        ("kind", ctypes.c_int32),
        ("SA", ctypes.c_void_p),
        ("U", ctypes.c_void_p),
        ("LCP", ctypes.c_void_p),
        ("primary", ctypes.c_int64),
        ("status", ctypes.c_int32),
        ("batched", ctypes.c_int32),
        ("t_submit", ctypes.c_double),
        ("t_start", ctypes.c_double),
        ("t_done", ctypes.c_double),
        ("batchable", ctypes.c_int32),
        ("reserved_", ctypes.c_int32),
    ]


def _declare():
    return load()


class Request:
    """One job with its host buffers (kept alive while the server uses them)."""

    def __init__(self, T: np.ndarray, kind: int, arrival: float = 0.0):
        self.T = np.ascontiguousarray(T, dtype=np.uint8)
        n = len(self.T)
        self.kind = kind
        self.arrival = arrival
        self.SA = np.empty(max(n, 1), np.int32)
        self.U = np.empty(max(n, 1), np.uint8) if kind == JOB_BWT else None
        self.LCP = np.empty(max(n, 1), np.int32) if kind == JOB_LCP else None
        self.job = Job()
        self.job.T = self.T.ctypes.data
        self.job.n = n
        self.job.kind = kind
        self.job.SA = self.SA.ctypes.data
        self.job.U = self.U.ctypes.data if self.U is not None else None
        self.job.LCP = self.LCP.ctypes.data if self.LCP is not None else None

    def outputs(self):
        n = len(self.T)
        out = {"SA": self.SA[:n]}
        if self.U is not None:
            out["U"] = self.U[:n]
            out["primary"] = int(self.job.primary)
        if self.LCP is not None:
            out["LCP"] = self.LCP[:n]
        return out


class Server:
    def __init__(self, slots: int = 4, hbm_budget: int = 0, batch_bytes: int = 32 << 20):
        self.lib = _declare()
        self.h = self.lib.sagpu_server_create(slots, hbm_budget, batch_bytes)
        if not self.h:
            raise RuntimeError(f"sagpu_server_create failed: {last_error()}")

    def submit(self, req: Request):
        rc = self.lib.sagpu_server_submit(self.h, ctypes.addressof(req.job))
        if rc != 0:
            raise RuntimeError(f"submit failed ({rc})")

    def wait(self, req: Request) -> int:
        return self.lib.sagpu_server_wait(self.h, ctypes.addressof(req.job))

    def drain(self):
        self.lib.sagpu_server_drain(self.h)

    def close(self):
        if self.h:
            self.lib.sagpu_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def run_alone(req: Request) -> int:
    """The plain path: one job, synchronously, on the default stream."""
    lib = _declare()
    return lib.sagpu_run_job(ctypes.addressof(req.job))


@dataclass
class TraceSpec:
    jobs: int = 2000
    rate: float = 200.0          # Poisson arrivals per second
    min_bytes: int = 4 << 10
    max_bytes: int = 64 << 20
    seed: int = 6
    mix: tuple = (0.4, 0.4, 0.2)  # SA, SA+BWT, SA+LCP


def make_trace(spec: TraceSpec):
    """Requests with arrival offsets (s), sizes log-uniform, order-2 Markov ACGT."""
    rng = np.random.default_rng(spec.seed)
    gaps = rng.exponential(1.0 / spec.rate, spec.jobs)
    arrivals = np.cumsum(gaps) - gaps[0]
    sizes = np.exp(rng.uniform(np.log(spec.min_bytes), np.log(spec.max_bytes), spec.jobs)).astype(np.int64)
    kinds = rng.choice(3, spec.jobs, p=list(spec.mix))
    reqs = []
    for q in range(spec.jobs):
        T = workloads.markov(int(sizes[q]), seed=spec.seed * 1_000_003 + q, order=2)
        reqs.append(Request(T, int(kinds[q]), float(arrivals[q])))
    return reqs


def replay(reqs, server: Server | None):
    """Submits every request at its arrival time (server) or runs them in arrival
    order one at a time (server=None).  Returns the start clock."""
    lib = _declare()
    t0 = lib.sagpu_clock()
    for r in reqs:
        target = t0 + r.arrival
        while True:
            d = target - lib.sagpu_clock()
            if d <= 0:
                break
            time.sleep(min(d, 0.002))
        if server is not None:
            server.submit(r)
            r.job.t_submit = t0 + r.arrival
        else:
            r.job.t_submit = t0 + r.arrival
            rc = run_alone(r)
            if rc != 0:
                raise RuntimeError(f"job failed ({rc}): {last_error()}")
    if server is not None:
        server.drain()
        for r in reqs:
            if r.job.status != 0:
                raise RuntimeError(f"job failed ({r.job.status})")
    return t0


def summarize(reqs, t0: float):
    lat = np.array([r.job.t_done - (t0 + r.arrival) for r in reqs])
    total = sum(len(r.T) for r in reqs)
    # bytes over the host link: the text in; SA (4 B per position) out, plus the
    # BWT (1 B) or the LCP (4 B) by kind
    link = sum(len(r.T) * (1 + 4 + (1 if r.kind == JOB_BWT else 4 if r.kind == JOB_LCP else 0)) for r in reqs)
    end = max(r.job.t_done for r in reqs)
    span = end - t0
    return {
        "jobs": len(reqs),
        "bytes": int(total),
        "makespan_s": round(span, 4),
        "aggregate_MBps": round(total / span / 1e6, 2),
        "link_bytes": int(link),
        "link_GBps": round(link / span / 1e9, 2),
        "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
        "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
        "max_ms": round(float(lat.max()) * 1e3, 3),
        "batched_jobs": int(sum(r.job.batched for r in reqs)),
    }
