"""Batched generalized suffix arrays (many separator-terminated strings).

batched(): one sagpu_gsa_batch32 call for the whole set (device resident).
plain():   the per-string path, one sagpu_build32(GSA) call per string on the same
           device buffers -- what a caller without batching would do.
Both return per-string SA (local indices) and LCP concatenated in text order:
string k's arrays occupy [offsets[k], offsets[k+1]).  They equal, string by string,
libsais_gsa + libsais_plcp_gsa + libsais_lcp on that string alone.
"""
from __future__ import annotations

import time

import numpy as np

from . import SAGPU_FLAG_GSA, last_error, load


class DeviceText:
    """A text and output buffers resident in device memory."""

    def __init__(self, T: np.ndarray, want_lcp: bool = True):
        self.lib = load()
        self.n = len(T)
        self.dT = self.lib.sagpu_malloc(self.n)
        self.dSA = self.lib.sagpu_malloc(4 * self.n)
        self.dLCP = self.lib.sagpu_malloc(4 * self.n) if want_lcp else None
        if not self.dT or not self.dSA or (want_lcp and not self.dLCP):
            self.free()
            raise MemoryError("device allocation failed")
        assert self.lib.sagpu_memcpy_h2d(self.dT, np.ascontiguousarray(T).ctypes.data, self.n) == 0

    def result(self):
        SA = np.empty(self.n, np.int32)
        self.lib.sagpu_memcpy_d2h(SA.ctypes.data, self.dSA, 4 * self.n)
        L = None
        if self.dLCP:
            L = np.empty(self.n, np.int32)
            self.lib.sagpu_memcpy_d2h(L.ctypes.data, self.dLCP, 4 * self.n)
        return SA, L

    def free(self):
        for p in (self.dT, self.dSA, self.dLCP):
            if p:
                self.lib.sagpu_free(p)
        self.dT = self.dSA = self.dLCP = None


def batched_on_device(dev: DeviceText, offsets: np.ndarray) -> int:
    offs = np.ascontiguousarray(offsets, dtype=np.int64)
    return dev.lib.sagpu_gsa_batch32(dev.dT, dev.n, offs.ctypes.data, len(offs) - 1, dev.dSA, dev.dLCP)


def plain_on_device(dev: DeviceText, offsets: np.ndarray, strings=None) -> int:
    """One GSA build per string (the unbatched path); strings = iterable of indices."""
    lib = dev.lib
    offs = np.asarray(offsets, dtype=np.int64)
    idx = range(len(offs) - 1) if strings is None else strings
    for k in idx:
        o, m = int(offs[k]), int(offs[k + 1] - offs[k])
        r = lib.sagpu_build32(dev.dT + o, m, SAGPU_FLAG_GSA, dev.dSA + 4 * o, None, None,
                              (dev.dLCP + 4 * o) if dev.dLCP else None, None)
        if r < 0:
            return int(r)
    return 0


def batched(T: np.ndarray, offsets: np.ndarray, want_lcp: bool = True):
    dev = DeviceText(T, want_lcp)
    try:
        rc = batched_on_device(dev, offsets)
        if rc != 0:
            raise RuntimeError(f"sagpu_gsa_batch32 returned {rc}: {last_error()}")
        return dev.result()
    finally:
        dev.free()


def plain(T: np.ndarray, offsets: np.ndarray, want_lcp: bool = True):
    dev = DeviceText(T, want_lcp)
    try:
        rc = plain_on_device(dev, offsets)
        if rc != 0:
            raise RuntimeError(f"per-string build returned {rc}: {last_error()}")
        return dev.result()
    finally:
        dev.free()


def timed(fn, *args, reps=1):
    lib = load()
    lib.sagpu_synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        rc = fn(*args)
        if rc != 0:
            raise RuntimeError(f"{fn.__name__} returned {rc}: {last_error()}")
    lib.sagpu_synchronize()
    return (time.perf_counter() - t0) / reps
