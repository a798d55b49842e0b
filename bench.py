"""Benchmark: suffix array + BWT + PLCP + LCP of a 256 MiB DNA-like string on one MI355X.

One step = one fused sagpu_build32() call on a device-resident text (order-3 Markov
ACGT with 2% tandem repeats and 5% copied segments, seed 2): SA, BWT + primary
index, PLCP and LCP, all int32, outputs left in HBM.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--no-cpu]

roofline: the kernel with the largest device time per build, among the kernels
the engine times (HIP events recorded on the engine stream around every launch,
sagpu_stats.kernel_ms, in up to 5 extra steps after the timed ones: the timed
steps run without events, with the doubling rounds replaying captured graphs).  achieved = its algorithmic
bytes per launch (reads + writes at element size, counted by the engine per
launch: e.g. k_round_tile 28 B per active list entry, k_isa_scatter 12 B per
position) / its average launch duration.  roofline.kernels lists the five kernels
with the largest time per build the same way.  traffic is the HBM bytes per
launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE (scaled by a copy kernel of
known size), read from profiles/pmc_traffic.json.
pipeline: B_model = 269 n + 244 sum_r a_r bytes per step (prefix-doubling yardstick
with 8 initial symbols; a_r counts suffixes with max(LCP[j], LCP[j+1]) >= 8*2^(r-1),
computed from the LCP this run produced), divided by the step time.
extras: batched GSA (20,000 strings), the serving trace (600 jobs, served against
one job at a time), and the 64 MiB high-LCP inputs (all 'A', period 17, Fibonacci)
with the reference beside them.
cpu_baseline: the reference libsais 2.10.4 (built from its sources by oracle/Makefile)
with OpenMP on this host's cores, on the same input: libsais_omp + libsais_bwt_omp
+ libsais_plcp_omp + libsais_lcp_omp, each timed, summed; median and p99 of 15
runs on all cores, and one run each at 1 and 8 threads on the full input.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# sagpu_stats.kernel_ms index -> kernel (include/sagpu.h)
KERNELS = ["k_radix_pass<u32 key, u32 value>", "k_radix_pass (generating first pass)",
           "k_radix_pass<u64 key, u32 value>", "k_init_emit", "k_scatter_windows (initial ranks)", "k_round_gather",
           "k_round_tile",
           "k_plcp_pairs", "k_plcp_irr_pairs", "k_lcp_fix", "k_bwt_seeded"]
METRIC = "SA+BWT input MB/s @256 MiB DNA; batched vs plain aggregate MB/s; p99 latency"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def copy_peak(lib, nbytes: int = 1 << 30, reps: int = 5):
    """Device copy bandwidth (read + write bytes / time) of the engine's copy kernel."""
    src = lib.sagpu_malloc(nbytes)
    dst = lib.sagpu_malloc(nbytes)
    if not src or not dst:
        raise MemoryError("device allocation failed")
    try:
        lib.sagpu_memset(src, 1, nbytes)
        assert lib.sagpu_memcpy_d2d(dst, src, nbytes) == 0
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            assert lib.sagpu_memcpy_d2d(dst, src, nbytes) == 0
            best = min(best, time.perf_counter() - t0)
    finally:
        lib.sagpu_free(src)
        lib.sagpu_free(dst)
    return 2 * nbytes / best / 1e9


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of a kernel from the committed PMC summary (rocprofv3
    FETCH_SIZE and WRITE_SIZE runs, scaled by the copy-kernel calibration), or None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    if d.get("workload") != workload:
        return None
    return d.get("bytes_per_launch", {}).get(kernel)


def a_r_from_lcp(L: np.ndarray, h0: int = 8):
    """a_r = #{j : max(LCP[j], LCP[j+1]) >= h0 * 2^(r-1)}, LCP[n] := 0."""
    n = len(L)
    m = np.maximum(L, np.append(L[1:], 0).astype(L.dtype))
    out = []
    h = h0
    while True:
        a = int(np.count_nonzero(m >= h))
        if a == 0:
            break
        out.append(a)
        h *= 2
        if h > 4 * n:
            break
    return out


def cpu_baseline(T: np.ndarray, reps: int, degenerate: bool = False, batched: bool = False,
                 serving_jobs: list | None = None):
    """Reference libsais with OpenMP, or the oracle restatement if the reference
    build is absent.  With the reference, also the reference side of the extra
    sections that ran (degenerate inputs, batched GSA, serving trace)."""
    from tests import oracle_lib as ol

    n = len(T)
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = min(cores, 64)
    if ol.have_ref():
        def run(th):
            t0 = time.perf_counter()
            rc, SA, _ = ol.ref_sa(T, threads=th)
            t1 = time.perf_counter()
            ol.ref_bwt(T, threads=th)
            t2 = time.perf_counter()
            P = ol.ref_plcp(T, SA, threads=th)
            t3 = time.perf_counter()
            ol.ref_lcp(P, SA, threads=th)
            t4 = time.perf_counter()
            return t1 - t0, t2 - t1, t3 - t2, t4 - t3

        times, parts = [], []
        for r in range(reps):
            p = run(cores)
            parts.append(p)
            times.append(sum(p))
            log(f"cpu reference run {r + 1}/{reps}: sa {p[0]:.2f}s bwt {p[1]:.2f}s plcp {p[2]:.2f}s lcp {p[3]:.2f}s")
        t = float(np.median(times))
        t99 = float(np.percentile(times, 99))
        # thread sweep on the full input, one run per count (the reference saturates
        # near the memory-channel count), single-threaded included
        sweep = {str(cores): round(n / t / 1e6, 2)}
        for th in sorted({1, 8} - {cores}):
            p = run(th)
            sweep[str(th)] = round(n / sum(p) / 1e6, 2)
            log(f"cpu reference, {th} threads: {sum(p):.2f}s")
        med = np.median(np.array(parts), axis=0)
        out = {"value": round(n / t / 1e6, 2), "unit": "MB/s", "cores": cores, "kind": "reference",
               "sample": f"full {n}-byte input, SA+BWT+PLCP+LCP via libsais_omp/_bwt_omp/_plcp_omp/_lcp_omp "
                         f"(each timed, summed), median of {reps} runs on {cores} threads",
               "p99_MBps": round(n / t99 / 1e6, 2),
               "median_s": {"sa": round(float(med[0]), 3), "bwt": round(float(med[1]), 3),
                            "plcp": round(float(med[2]), 3), "lcp": round(float(med[3]), 3)},
               "cpu_model": _cpu_model(),
               "thread_sweep_MBps": {"sample_bytes": n, **{k: sweep[k] for k in sorted(sweep, key=int)}}}
        ol.ref()   # loaded once, before any worker thread calls it
        if degenerate:
            out["degenerate"] = cpu_degenerate(ol, cores)
        if batched:
            out["batched"] = cpu_batched(ol, cores)
        if serving_jobs:
            out["serving"] = cpu_serving(ol, cores, serving_jobs)
        return out
    # oracle port, single thread, on a bounded prefix
    m = min(n, 1 << 24)
    S = np.ascontiguousarray(T[:m])
    t0 = time.perf_counter()
    rc, SA = ol.sa(S)
    ol.bwt(S, SA)
    P = ol.plcp(S, SA)
    ol.lcp(P, SA)
    t = time.perf_counter() - t0
    return {"value": round(m / t / 1e6, 2), "unit": "MB/s", "cores": 1, "kind": "port",
            "sample": f"first {m} bytes, SA+BWT+PLCP+LCP with the oracle restatement"}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_degenerate(ol, cores: int):
    """Reference with OpenMP on each 64 MiB high-LCP input, once."""
    out = {}
    for name, T in degenerate_inputs(1 << 26):
        t0 = time.perf_counter()
        rc, SA, _ = ol.ref_sa(T, threads=cores)
        ol.ref_bwt(T, threads=cores)
        P = ol.ref_plcp(T, SA, threads=cores)
        ol.ref_lcp(P, SA, threads=cores)
        out[name] = round(len(T) / (time.perf_counter() - t0) / 1e6, 1)
    log(f"cpu reference, degenerate inputs: {out}")
    return {"MBps": out, "sample": "each 64 MiB input once: libsais_omp + _bwt_omp + _plcp_omp + _lcp_omp"}


def cpu_batched(ol, cores: int):
    """The batched GSA set through the reference: every string on its own
    (libsais_gsa + libsais_plcp_gsa + libsais_lcp, single-threaded, on `cores`
    worker threads), and a 64 MiB prefix of whole strings as one GSA with OpenMP."""
    from concurrent.futures import ThreadPoolExecutor

    from libsais_amd import workloads

    T, offs = workloads.gsa_batch(count=20000, seed=5)
    count = len(offs) - 1

    def one(k):
        S = T[int(offs[k]):int(offs[k + 1])]   # ends with its 0 separator
        rc, SA, _ = ol.ref_sa(S, gsa=True)
        P = ol.ref_plcp(S, SA, gsa=True)
        ol.ref_lcp(P, SA)

    t0 = time.perf_counter()
    with ThreadPoolExecutor(cores) as ex:
        list(ex.map(one, range(count)))
    t_per = time.perf_counter() - t0
    k = int(np.searchsorted(offs, 64 << 20, side="right")) - 1
    C = np.ascontiguousarray(T[:int(offs[k])])
    t0 = time.perf_counter()
    rc, SA, _ = ol.ref_sa(C, gsa=True, threads=cores)
    P = ol.ref_plcp(C, SA, gsa=True, threads=cores)
    ol.ref_lcp(P, SA, threads=cores)
    t_cat = time.perf_counter() - t0
    out = {"per_string_MBps": round(int(offs[-1]) / t_per / 1e6, 1),
           "per_string_sample": f"all {count} strings ({int(offs[-1])} bytes), one reference call chain per "
                                f"string on {cores} worker threads",
           "concatenation_MBps": round(len(C) / t_cat / 1e6, 1),
           "concatenation_sample": f"first {k} strings ({len(C)} bytes) as one GSA: libsais_gsa_omp + "
                                   f"_plcp_gsa_omp + _lcp_omp"}
    log(f"cpu reference, batched: {out}")
    return out


def cpu_serving(ol, cores: int, jobs: list):
    """The serving trace replayed on CPU workers: all jobs at once, in arrival
    order, one job per worker thread with the single-threaded reference (SA:
    libsais; SA+BWT: libsais + libsais_bwt; SA+LCP: libsais + libsais_plcp +
    libsais_lcp)."""
    from concurrent.futures import ThreadPoolExecutor

    def one(job):
        T, kind = job
        rc, SA, _ = ol.ref_sa(T)
        if kind == 1:
            ol.ref_bwt(T)
        elif kind == 2:
            P = ol.ref_plcp(T, SA)
            ol.ref_lcp(P, SA)

    t0 = time.perf_counter()
    with ThreadPoolExecutor(cores) as ex:
        list(ex.map(one, jobs))
    t = time.perf_counter() - t0
    total = sum(len(T) for T, _ in jobs)
    out = {"aggregate_MBps": round(total / t / 1e6, 2), "makespan_s": round(t, 3),
           "sample": f"the {len(jobs)}-job trace all at once on {cores} worker threads"}
    log(f"cpu reference, serving: {out}")
    return out


def bench_batched(reps: int = 2, plain_sample: int = 1000):
    """Batched GSA of 20,000 separator-terminated strings (log-normal lengths, median
    4 KiB, clipped to 256 B - 4 MiB, ~512 MiB, 10% with a shared 1-20 KiB prefix):
    one sagpu_gsa_batch32 call (SA + LCP per string) against the per-string path
    (one build per string) on a random sample of the same strings."""
    import libsais_amd as la
    from libsais_amd import batch, workloads

    from libsais_amd import autotune

    T, offs = workloads.gsa_batch(count=20000, seed=5)
    # autotune on the first 2000 strings of the same set (outputs checked identical
    # for every point), then run the whole set with the chosen knobs
    k = 2000
    tuned = autotune.tune_batched(T[:int(offs[k])].copy(), offs[:k + 1].copy(), reps=2)
    dev = batch.DeviceText(T, want_lcp=True)
    try:
        rc = batch.batched_on_device(dev, offs)  # warm-up
        if rc != 0:
            raise RuntimeError(f"batched build failed ({rc}): {la.last_error()}")
        t_default = batch.timed(batch.batched_on_device, dev, offs, reps=reps)
        autotune.apply(tuned["best"])
        batch.batched_on_device(dev, offs)
        t_b = batch.timed(batch.batched_on_device, dev, offs, reps=reps)
        autotune.apply(autotune.DEFAULT)
        rng = np.random.default_rng(0)
        sample = sorted(rng.choice(len(offs) - 1, min(plain_sample, len(offs) - 1), replace=False).tolist())
        sample_bytes = int(sum(offs[k + 1] - offs[k] for k in sample))
        batch.plain_on_device(dev, offs, sample[:20])  # warm-up
        t_p = batch.timed(batch.plain_on_device, dev, offs, sample)
    finally:
        dev.free()
    b_mbps = len(T) / t_b / 1e6
    p_mbps = sample_bytes / t_p / 1e6
    return {
        "workload": "gsa_batch_20000_strings_512MiB_sa_lcp",
        "strings": int(len(offs) - 1),
        "bytes": int(len(T)),
        "batched_ms": round(t_b * 1e3, 2),
        "default_knobs_ms": round(t_default * 1e3, 2),
        "autotune": {"sample": f"first {k} strings, {int(offs[k])} bytes", "best": tuned["best"],
                     "grid": autotune.GRID, "results": tuned["results"]},
        "batched_MBps": round(b_mbps, 1),
        "plain_MBps": round(p_mbps, 1),
        "plain_sample": f"{len(sample)} random strings, {sample_bytes} bytes, one build per string",
        "ratio": round(b_mbps / p_mbps, 2),
    }


def bench_degenerate(lib, reps: int = 3):
    """The 64 MiB high-LCP inputs (all 'A', period 17 with one mutated byte at n/2,
    Fibonacci word): fused SA+BWT+PLCP+LCP on a device-resident text, median of
    `reps` builds, with the doubling rounds taken."""
    import libsais_amd as la
    from libsais_amd import workloads

    n = 1 << 26
    out = {}
    dT = lib.sagpu_malloc(n)
    bufs = [lib.sagpu_malloc(4 * n), lib.sagpu_malloc(n), lib.sagpu_malloc(4 * n), lib.sagpu_malloc(4 * n)]
    try:
        for name, T in degenerate_inputs(n):
            assert lib.sagpu_memcpy_h2d(dT, T.ctypes.data, n) == 0
            f64 = np.zeros(256, np.int64)
            times = []
            for _ in range(reps + 1):
                lib.sagpu_synchronize()
                t0 = time.perf_counter()
                r = lib.sagpu_build32(dT, n, 0, bufs[0], bufs[1], bufs[2], bufs[3], f64.ctypes.data)
                lib.sagpu_synchronize()
                times.append(time.perf_counter() - t0)
                if r < 0:
                    raise RuntimeError(f"sagpu_build32 failed ({r}): {la.last_error()}")
            t = float(np.median(times[1:]))
            st = la.stats()
            out[name] = {"ms": round(t * 1000, 2), "MBps": round(n / t / 1e6, 1), "rounds": st.rounds,
                         "big_rounds": st.big_rounds, "big_members": st.big_elems,
                         "pivot_rounds": st.big_pivot_rounds, "pivot_members": st.big_pivot_elems}
            log(f"degenerate {name}: {out[name]}")
    finally:
        lib.sagpu_free(dT)
        for b in bufs:
            lib.sagpu_free(b)
    return {"workload": "64MiB_all_a_period17_fibonacci_sa_bwt_plcp_lcp", "n": n, **out}


def degenerate_inputs(n: int):
    from libsais_amd import workloads

    return [("all_a", workloads.all_a(n)), ("period17_mutated", workloads.periodic(n, 17, 3)),
            ("fibonacci", workloads.fibonacci(n))]


def bench_serving(jobs: int, max_bytes: int, slots: int = 4, keep: list | None = None):
    """Jobs with sizes log-uniform from 4 KiB, order-2 Markov ACGT, SA / SA+BWT /
    SA+LCP = 0.4/0.4/0.2, replayed twice: Poisson arrivals at 200 req/s (latency
    under the specified load) and all at once (throughput when the GPU is the
    bottleneck).  Served = slots + HBM admission + continuous batching; plain =
    one job at a time on one stream, same order."""
    from libsais_amd import serving

    spec = serving.TraceSpec(jobs=jobs, rate=200.0, min_bytes=4 << 10, max_bytes=max_bytes, seed=6)
    reqs = serving.make_trace(spec)
    out = {"workload": f"{jobs}_jobs_4KiB_{max_bytes >> 20}MiB_markov2_sa_bwt_lcp", "slots": slots}
    with serving.Server(slots=slots) as sv:
        warm = [serving.Request(r.T, r.kind) for r in reqs[:8]]
        for r in warm:
            sv.submit(r)
        sv.drain()
        t0 = serving.replay(reqs, sv)
        out["poisson200_served"] = serving.summarize(reqs, t0)
        arrivals = [r.arrival for r in reqs]
        for r in reqs:
            r.arrival = 0.0
        t0 = serving.replay(reqs, sv)
        out["saturated_served"] = serving.summarize(reqs, t0)
    t1 = serving.replay(reqs, None)
    out["saturated_plain"] = serving.summarize(reqs, t1)
    for r, a in zip(reqs, arrivals):
        r.arrival = a
    t1 = serving.replay(reqs, None)
    out["poisson200_plain"] = serving.summarize(reqs, t1)
    out["saturated_ratio_MBps"] = round(out["saturated_served"]["aggregate_MBps"] /
                                        out["saturated_plain"]["aggregate_MBps"], 2)
    if keep is not None:   # texts and kinds for the reference replay of the same trace
        keep.extend((r.T, r.kind) for r in reqs)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 28)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--cpu-reps", type=int, default=15)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the batched, serving and degenerate-input sections")
    ap.add_argument("--mode", choices=["main", "serving-full", "batched", "copy", "degenerate"], default="main")
    ap.add_argument("--set", action="append", default=[], metavar="NAME=VALUE",
                    help="engine option (sagpu_set_option), repeatable")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on exactly one GPU: sharding a text over several GPUs "
                                   "is designed but not built, so only --gpus 1 is measured",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import libsais_amd as la
    from libsais_amd import workloads

    lib = la.load()
    options = {}
    for kv in args.set:
        name, _, val = kv.partition("=")
        if lib.sagpu_set_option(name.encode(), int(val)) != 0:
            raise SystemExit(f"unknown option or value: {kv}")
        options[name] = int(val)
    if args.mode == "serving-full":
        print(json.dumps({"serving": bench_serving(2000, 64 << 20)}), flush=True)
        return
    if args.mode == "batched":
        print(json.dumps({"batched": bench_batched()}), flush=True)
        return
    if args.mode == "degenerate":
        print(json.dumps({"degenerate": bench_degenerate(lib)}), flush=True)
        return
    if args.mode == "copy":
        print(json.dumps({"copy_GBs": round(copy_peak(lib), 1), "bytes": 1 << 30}), flush=True)
        return
    n = args.n
    t0 = time.perf_counter()
    T = workloads.dna_like(n, seed=args.seed)
    log(f"generated {n} bytes in {time.perf_counter() - t0:.1f}s")

    if lib.sagpu_init() != 0:
        raise RuntimeError(la.last_error())
    dT = lib.sagpu_malloc(n)
    dSA = lib.sagpu_malloc(4 * n)
    dU = lib.sagpu_malloc(n)
    dP = lib.sagpu_malloc(4 * n)
    dL = lib.sagpu_malloc(4 * n)
    if not all((dT, dSA, dU, dP, dL)):
        raise MemoryError("device allocation failed")
    assert lib.sagpu_memcpy_h2d(dT, T.ctypes.data, n) == 0
    freq = np.zeros(256, np.int64)

    def step():
        r = lib.sagpu_build32(dT, n, 0, dSA, dU, dP, dL, freq.ctypes.data)
        if r < 0:
            raise RuntimeError(f"sagpu_build32 failed ({r}): {la.last_error()}")
        return r

    lib.sagpu_set_timing(0)
    for i in range(args.warmup):
        step()
        log(f"warmup {i} done")

    # the timed steps run as in production (no timing events; the rounds replay their
    # captured graphs) ...
    lib.sagpu_synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        primary = step()
    lib.sagpu_synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1000.0 / args.steps
    value = n / (ms_per_step / 1000.0) / 1e6
    log(f"timed {args.steps} steps: {ms_per_step:.2f} ms/step, {value:.1f} MB/s")
    graphs = la.stats().round_graph_launches

    # ... then the same steps again with HIP events around every kernel launch (kernels
    # one by one), for the per-kernel rooflines and the family times
    psteps = max(2, min(args.steps, 5))
    lib.sagpu_set_timing(1)
    sweep_ms, sweep_launch, sweep_elems, sweep_bytes = 0.0, 0, 0, 0
    k_ms = [0.0] * len(KERNELS)
    k_launch = [0] * len(KERNELS)
    k_bytes = [0] * len(KERNELS)
    dev_ms = 0.0
    fam = {"initial": 0.0, "rounds": 0.0, "bwt": 0.0, "plcp": 0.0, "lcp": 0.0}
    t0 = time.perf_counter()
    for i in range(psteps):
        step()
        st = la.stats()
        sweep_ms += st.ms_onesweep
        sweep_launch += st.onesweep_launches
        sweep_elems += st.sorted_elems
        sweep_bytes += st.sorted_bytes
        for q in range(len(KERNELS)):
            k_ms[q] += st.kernel_ms[q]
            k_launch[q] += st.kernel_launches[q]
            k_bytes[q] += st.kernel_bytes[q]
        dev_ms += st.ms_total
        fam["initial"] += st.ms_initial
        fam["rounds"] += st.ms_rounds
        fam["bwt"] += st.ms_bwt
        fam["plcp"] += st.ms_plcp
        fam["lcp"] += st.ms_lcp
    lib.sagpu_synchronize()
    timed_wall_ms = (time.perf_counter() - t0) * 1000.0 / psteps
    lib.sagpu_set_timing(0)

    st = la.stats()
    rounds = st.rounds
    active = [int(st.active[r]) for r in range(min(rounds, 64))]

    workload = "dna_like_256MiB_sa_bwt_plcp_lcp" if n == 1 << 28 else f"dna_like_{n}B_sa_bwt_plcp_lcp"
    cpeak = copy_peak(lib)

    # per-kernel rooflines: algorithmic bytes per launch over the average launch time,
    # kernels ordered by device time per build
    kernels = []
    for q, name in enumerate(KERNELS):
        if k_launch[q] == 0 or k_ms[q] <= 0:
            continue
        avg_ms = k_ms[q] / k_launch[q]
        bpl = k_bytes[q] / k_launch[q]
        ach = bpl / (avg_ms / 1000.0) / 1e9
        tr = measured_traffic(workload, name)
        kernels.append({"kernel": name, "ms_per_build": round(k_ms[q] / psteps, 4),
                        "launches_per_build": round(k_launch[q] / psteps, 2),
                        "avg_launch_ms": round(avg_ms, 4), "algorithmic_bytes_per_launch": int(bpl),
                        "achieved": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4),
                        "traffic": tr, "traffic_over_algorithmic": round(tr / bpl, 2) if tr else None})
    kernels.sort(key=lambda k: -k["ms_per_build"])
    top = kernels[0] if kernels else None
    timed_ms = sum(k["ms_per_build"] for k in kernels)

    # pipeline yardstick from this run's LCP
    L = np.empty(n, np.int32)
    lib.sagpu_memcpy_d2h(L.ctypes.data, dL, 4 * n)
    a_r = a_r_from_lcp(L)
    b_model = 269 * n + 244 * sum(a_r)
    del L

    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "MB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {
            "workload": workload,
            "n": n,
            "seed": args.seed,
            "outputs": ["SA", "BWT+primary", "PLCP", "LCP"],
            "index": "int32",
            "inputs_resident": "HBM",
            **({"options": options} if options else {}),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": top["kernel"] if top else None,
            "achieved": top["achieved"] if top else None,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": top["frac"] if top else None,
            "traffic": top["traffic"] if top else None,
            "copy_peak_GBs": round(cpeak, 1),
            "frac_of_copy": round(top["achieved"] / cpeak, 4) if top else None,
            "launches": int(top["launches_per_build"] * psteps) if top else 0,
            "avg_launch_ms": top["avg_launch_ms"] if top else None,
            "algorithmic_bytes_per_launch": top["algorithmic_bytes_per_launch"] if top else None,
            "share_of_build": round(top["ms_per_build"] / timed_wall_ms, 4) if top else None,
            "timing_steps": psteps,
            "ms_per_step_with_events": round(timed_wall_ms, 3),
            "kernels": kernels,
            "timed_kernels_ms_per_build": round(timed_ms, 3),
            "all_radix_passes": {"launches": sweep_launch, "ms": round(sweep_ms, 3), "bytes": sweep_bytes},
        },
        "pipeline": {
            "b_model_bytes": b_model,
            "a_r": a_r,
            "achieved_GBs": round(b_model / (ms_per_step / 1000.0) / 1e9, 1),
            "frac": round(b_model / (ms_per_step / 1000.0) / 1e9 / HBM_PEAK_GBS, 4),
            "device_ms_per_step": round(dev_ms / psteps, 3),
            "family_ms_per_step": {k: round(v / psteps, 3) for k, v in fam.items()},
            "round_graph_replays_per_build": graphs,
            "engine_rounds": rounds,
            "engine_active": active,
            "initial_symbols": st.initial_symbols,
            "primary": int(primary),
        },
    }
    for p in (dT, dSA, dU, dP, dL):
        lib.sagpu_free(p)

    # the libsais API on host buffers (PCIe transfers included): libsais + libsais_bwt
    # + libsais_plcp + libsais_lcp, as a caller of the reference would run them
    SAh = np.empty(n, np.int32)
    Uh = np.empty(n, np.uint8)
    Ph = np.empty(n, np.int32)
    Lh = np.empty(n, np.int32)
    Ah = np.empty(1, np.int32)
    la.libsais(T, SAh, n, 0, None)
    t0 = time.perf_counter()
    assert la.libsais(T, SAh, n, 0, None) == 0
    assert la.libsais_bwt(T, Uh, Ah, n, 0, None) > 0
    assert la.libsais_plcp(T, SAh, Ph, n) == 0
    assert la.libsais_lcp(Ph, SAh, Lh, n) == 0
    t_api = time.perf_counter() - t0
    result["host_api"] = {"calls": "libsais + libsais_bwt + libsais_plcp + libsais_lcp on host buffers",
                          "seconds": round(t_api, 3), "MBps": round(n / t_api / 1e6, 1)}
    del SAh, Uh, Ph, Lh

    jobs = None
    if not args.no_extras:
        log("batched GSA ...")
        result["batched"] = bench_batched()
        log(f"batched: {result['batched']}")
        log("serving trace ...")
        jobs = [] if not args.no_cpu else None
        result["serving"] = bench_serving(600, 64 << 20, keep=jobs)
        log(f"serving: {result['serving']}")
        log("degenerate inputs ...")
        result["degenerate"] = bench_degenerate(lib)
    if not args.no_cpu:
        log("cpu baseline ...")
        extras = not args.no_extras
        result["cpu_baseline"] = cpu_baseline(T, args.cpu_reps, degenerate=extras, batched=extras,
                                              serving_jobs=jobs if extras else None)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
