"""The allocation-free build, sagpu_build32_noalloc, against sagpu_build32.

sagpu_build32_noalloc takes every working buffer from one static device array
per GPU instead of the engine's pool, for callers that cannot allocate during a
call.  It has no counterpart in the reference; its outputs must equal those of
sagpu_build32 (SA, BWT and primary index, PLCP, LCP, freq) bit for bit, whose
own parity with the reference and the oracle the other GPU tests pin.

Two calls in a row, the second on a smaller text (the array keeps the first
call's contents), then the largest accepted n on the all-'A' text (the largest
working set: every round sorts one big group), and n beyond the limit rejected.
No call may make a device or pinned allocation (sagpu_stats.device_allocs).
"""
import numpy as np
import pytest

import libsais_amd as la
from libsais_amd import workloads

pytestmark = pytest.mark.gpu

MAX_N = 4194304                   # SAGPU_NOALLOC_MAX_N (include/sagpu.h)
SCRATCH = 288 << 20               # SAGPU_NOALLOC_SCRATCH_BYTES


class DevBufs:
    """Device text and outputs of one build (allocated by the caller, as the
    allocation-free contract expects)."""

    def __init__(self, lib, T, gsa=False):
        self.lib, self.n, self.gsa = lib, len(T), gsa
        n = self.n
        self.dT = lib.sagpu_malloc(n)
        self.out = [lib.sagpu_malloc(4 * n), None if gsa else lib.sagpu_malloc(n), lib.sagpu_malloc(4 * n),
                    lib.sagpu_malloc(4 * n)]
        assert self.dT and all(b for b in self.out if b is not None)
        assert lib.sagpu_memcpy_h2d(self.dT, T.ctypes.data, n) == 0

    def run(self, fn):
        """fn = sagpu_build32 or sagpu_build32_noalloc; returns (code, SA, U, PLCP, LCP, freq)."""
        lib, n = self.lib, self.n
        for b in self.out:
            if b is not None:
                assert lib.sagpu_memset(b, 0xA5, n) == 0      # stale bytes must not leak through
        f = np.zeros(256, np.int64)
        r = fn(self.dT, n, la.SAGPU_FLAG_GSA if self.gsa else 0, *self.out, f.ctypes.data)
        res = [r]
        for b, dt in zip(self.out, (np.int32, np.uint8, np.int32, np.int32)):
            if b is None:
                res.append(None)
                continue
            o = np.empty(n, dt)
            assert lib.sagpu_memcpy_d2h(o.ctypes.data, b, o.nbytes) == 0
            res.append(o)
        res.append(f)
        return res

    def free(self):
        for p in [self.dT] + self.out:
            if p is not None:
                self.lib.sagpu_free(p)


def _same(a, b, what):
    assert a[0] == b[0], (what, "return code / primary index", a[0], b[0])
    for k, name in enumerate(("SA", "U", "PLCP", "LCP", "freq"), start=1):
        if a[k] is None:
            assert b[k] is None
            continue
        assert np.array_equal(a[k], b[k]), (what, name)


def test_noalloc_matches_pooled_build_two_calls_then_max_n():
    lib = la.load()
    assert lib.sagpu_init() == 0, la.last_error()
    texts = [("dna_like_3MiB", workloads.dna_like(3 << 20, seed=41), False),
             ("acgt_smaller", workloads.iid(700001, seed=42, alphabet=b"ACGT"), False),
             ("gsa_set", _gsa_text(), True),
             ("all_a_max_n", workloads.all_a(MAX_N), False)]
    bufs = [DevBufs(lib, T, gsa) for _, T, gsa in texts]
    try:
        print("[no-alloc] expected outputs from sagpu_build32", flush=True)
        expect = [b.run(lib.sagpu_build32) for b in bufs]
        for (name, _, _), e in zip(texts, expect):
            assert e[0] >= 0, (name, la.last_error())
        allocs0 = la.stats().device_allocs
        for (name, T, _), b, e in zip(texts, bufs, expect):
            print(f"[no-alloc] {name} n={len(T)}", flush=True)
            got = b.run(lib.sagpu_build32_noalloc)
            assert got[0] >= 0, (name, la.last_error())
            st = la.stats()
            assert 0 < st.peak_bytes <= SCRATCH, (name, st.peak_bytes)
            print(f"[no-alloc] {name}: identical check, scratch peak {st.peak_bytes} B "
                  f"({st.peak_bytes / len(T):.1f} B per position)", flush=True)
            _same(got, e, name)
        assert la.stats().device_allocs == allocs0, "the allocation-free build allocated device or pinned memory"
    finally:
        for b in bufs:
            b.free()


def test_noalloc_rejects_n_beyond_limit():
    lib = la.load()
    # rejected from the arguments alone: no text is read, so a one-byte buffer serves
    d = lib.sagpu_malloc(1)
    try:
        r = lib.sagpu_build32_noalloc(d, MAX_N + 1, 0, None, None, None, None, None)
        assert r == -1
        assert "SAGPU_NOALLOC_MAX_N" in la.last_error()
    finally:
        lib.sagpu_free(d)


def _gsa_text():
    rng = np.random.default_rng(43)
    parts = []
    for _ in range(300):
        parts.append(rng.integers(1, 5, int(rng.integers(1, 4000))).astype(np.uint8))
        parts.append(np.zeros(1, np.uint8))
    return np.concatenate(parts)
