"""Autotuner for the engine's launch knobs.

The knobs are process-wide options of the engine (sagpu_set_option):

  radix_bits     digit width of the radix sort: 4, 8 or 11 bits (16, 256 or 2048
                 buckets per pass; 4 and 11 run the onesweep variant on 256-thread
                 tiles)
  radix_items    keys per thread of a radix tile: 8, 16, 32 (tiles of 2048, 4096,
                 8192 keys; 16384 with 512 threads and 8-bit digits)
  radix_variant  0 = onesweep (decoupled look-back), 1 = reduce-then-scan
  radix_threads  onesweep workgroup size with 32 keys per thread: 256 or 512
  tile_range     list entries per workgroup of the doubling-round tile kernel:
                 256, 512, 1024

The tuner runs a workload once per point of the grid on device-resident data,
checks that every point produces output identical to the default configuration's
(the knobs change speed only), and keeps the fastest.  The default grid sweeps
digit width x radix tile x tile range.  Active lists are compacted every round at
no extra pass (the tile kernel writes the survivors of its groups as it ranks them,
in one contiguous run per workgroup), so there is no compaction threshold to tune:
leaving retired entries in place would save no traffic.

  python -m libsais_amd.autotune --workload dna --n 67108864 --out tune.json
  python -m libsais_amd.autotune --workload batched --out tune.json
  autotune.apply(autotune.load("tune.json"))
"""
from __future__ import annotations

import argparse
import itertools
import json
import time

import numpy as np

from . import SAGPU_FLAG_GSA, last_error, load as _load_lib

GRID = {
    "radix_bits": (4, 8, 11),
    "radix_items": (8, 16, 32),
    "tile_range": (256, 512, 1024),
}
DEFAULT = {"radix_bits": 8, "radix_items": 32, "radix_variant": 0, "radix_threads": 512, "tile_range": 512}


def apply(cfg: dict) -> None:
    lib = _load_lib()
    for name, value in cfg.items():
        if lib.sagpu_set_option(name.encode(), int(value)) != 0:
            raise ValueError(f"engine rejected option {name}={value}")


def save(path: str, result: dict) -> None:
    with open(path, "w") as f:
        json.dump(result, f, indent=1)


def load(path: str) -> dict:
    with open(path) as f:
        return json.load(f)["best"]


def _points(grid: dict):
    names = list(grid)
    for values in itertools.product(*(grid[k] for k in names)):
        yield dict(zip(names, values))


def tune(run, fetch, grid: dict | None = None, reps: int = 2) -> dict:
    """run() executes the workload once on the device; fetch() returns its outputs
    (a tuple of arrays).  Returns {"best": cfg, "best_ms": t, "results": [...]}."""
    lib = _load_lib()
    grid = grid or GRID
    apply(DEFAULT)
    run()
    expect = fetch()
    results = []
    try:
        for cfg in _points(grid):
            apply({**DEFAULT, **cfg})
            run()  # warm-up (and pool growth) for this point
            outs = fetch()
            same = all(np.array_equal(a, b) for a, b in zip(outs, expect))
            lib.sagpu_synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                run()
            lib.sagpu_synchronize()
            ms = (time.perf_counter() - t0) * 1e3 / reps
            results.append({"config": cfg, "ms": round(ms, 3), "identical": bool(same)})
    finally:
        apply(DEFAULT)
    ok = [r for r in results if r["identical"]]
    if len(ok) != len(results):
        bad = [r["config"] for r in results if not r["identical"]]
        raise RuntimeError(f"configurations changed the output: {bad}")
    best = min(ok, key=lambda r: r["ms"])
    return {"best": {**DEFAULT, **best["config"]}, "best_ms": best["ms"], "results": results}


def tune_text(T: np.ndarray, gsa: bool = False, want_lcp: bool = True, grid=None, reps: int = 2) -> dict:
    """Tunes the fused build (SA, and LCP when want_lcp) of one device-resident text."""
    lib = _load_lib()
    n = len(T)
    dT = lib.sagpu_malloc(n)
    dSA = lib.sagpu_malloc(4 * n)
    dL = lib.sagpu_malloc(4 * n) if want_lcp else None
    if not dT or not dSA or (want_lcp and not dL):
        raise MemoryError("device allocation failed")
    try:
        assert lib.sagpu_memcpy_h2d(dT, np.ascontiguousarray(T).ctypes.data, n) == 0
        flags = SAGPU_FLAG_GSA if gsa else 0

        def run():
            r = lib.sagpu_build32(dT, n, flags, dSA, None, None, dL, None)
            if r < 0:
                raise RuntimeError(f"sagpu_build32 failed ({r}): {last_error()}")

        def fetch():
            SA = np.empty(n, np.int32)
            lib.sagpu_memcpy_d2h(SA.ctypes.data, dSA, 4 * n)
            if not dL:
                return (SA,)
            L = np.empty(n, np.int32)
            lib.sagpu_memcpy_d2h(L.ctypes.data, dL, 4 * n)
            return SA, L

        return tune(run, fetch, grid, reps)
    finally:
        for p in (dT, dSA, dL):
            if p:
                lib.sagpu_free(p)


def tune_batched(T: np.ndarray, offsets: np.ndarray, grid=None, reps: int = 2) -> dict:
    """Tunes sagpu_gsa_batch32 (SA + LCP) on a device-resident string set."""
    from . import batch

    dev = batch.DeviceText(T, want_lcp=True)
    try:
        def run():
            rc = batch.batched_on_device(dev, offsets)
            if rc != 0:
                raise RuntimeError(f"batched build failed ({rc}): {last_error()}")

        return tune(run, dev.result, grid, reps)
    finally:
        dev.free()


def main(argv=None):
    from . import workloads

    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--workload", choices=["dna", "batched"], default="dna")
    ap.add_argument("--n", type=int, default=1 << 26, help="text length for --workload dna")
    ap.add_argument("--strings", type=int, default=2000, help="strings for --workload batched")
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    if args.workload == "dna":
        res = tune_text(workloads.dna_like(args.n, seed=2), reps=args.reps)
    else:
        T, offs = workloads.gsa_batch(count=args.strings, seed=5, target_total=(512 << 20) * args.strings / 20000)
        res = tune_batched(T, offs, reps=args.reps)
    print(json.dumps(res))
    if args.out:
        save(args.out, res)


if __name__ == "__main__":
    main()
