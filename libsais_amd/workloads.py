"""Seeded synthetic inputs for the suffix-array workloads (C generators in csrc/gen.c).

Names follow what the inputs are:
  iid_acgt(n, seed)             i.i.d. uniform ACGT (1 MiB: seed 1; 4 GiB: seed 4)
  dna_like(n, seed)             order-3 Markov ACGT, GC 0.35-0.60, 2% tandem repeats,
                                5% copied segments with 1% substitutions (256 MiB: seed 2)
  all_a(n)                      n x 'A'
  periodic(n, period, seed)     period-17 string, one mutated byte at n/2 (seed 3)
  fibonacci(n)                  Fibonacci word over {a,b}
  gsa_batch(count, ...)         separator-terminated string set, log-normal lengths
  markov(n, seed, order)        order-k Markov ACGT (serving jobs)
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import GEN_PATH

_gen = None


def _lib():
    global _gen
    if _gen is None:
        if not os.path.exists(GEN_PATH):
            raise RuntimeError(f"{GEN_PATH} is missing; build with make -C libsais_amd/csrc")
        g = ctypes.CDLL(GEN_PATH)
        vp, u64, i32, d = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_double
        g.sagen_iid.argtypes = [vp, u64, u64, vp, i32]
        g.sagen_iid.restype = None
        g.sagen_dna_like.argtypes = [vp, u64, u64]
        g.sagen_dna_like.restype = i32
        g.sagen_markov.argtypes = [vp, u64, u64, i32]
        g.sagen_markov.restype = None
        g.sagen_periodic.argtypes = [vp, u64, u64, u64]
        g.sagen_periodic.restype = None
        g.sagen_fibonacci.argtypes = [vp, u64]
        g.sagen_fibonacci.restype = None
        g.sagen_gsa_batch.argtypes = [vp, u64, vp, i32, u64, d, d, d, d, d]
        g.sagen_gsa_batch.restype = ctypes.c_int64
        _gen = g
    return _gen


def iid(n: int, seed: int, alphabet: bytes = b"ACGT") -> np.ndarray:
    out = np.empty(n, dtype=np.uint8)
    al = np.frombuffer(alphabet, dtype=np.uint8).copy()
    _lib().sagen_iid(out.ctypes.data, n, seed, al.ctypes.data, len(al))
    return out


def iid_acgt(n: int = 1 << 20, seed: int = 1) -> np.ndarray:
    return iid(n, seed, b"ACGT")


def dna_like(n: int = 1 << 28, seed: int = 2) -> np.ndarray:
    out = np.empty(n, dtype=np.uint8)
    rc = _lib().sagen_dna_like(out.ctypes.data, n, seed)
    if rc != 0:
        raise MemoryError("sagen_dna_like failed")
    return out


def markov(n: int, seed: int, order: int = 2) -> np.ndarray:
    out = np.empty(n, dtype=np.uint8)
    _lib().sagen_markov(out.ctypes.data, n, seed, order)
    return out


def all_a(n: int = 1 << 26) -> np.ndarray:
    return np.full(n, ord("A"), dtype=np.uint8)


def periodic(n: int = 1 << 26, period: int = 17, seed: int = 3) -> np.ndarray:
    out = np.empty(n, dtype=np.uint8)
    _lib().sagen_periodic(out.ctypes.data, n, period, seed)
    return out


def fibonacci(n: int = 1 << 26) -> np.ndarray:
    out = np.empty(n, dtype=np.uint8)
    _lib().sagen_fibonacci(out.ctypes.data, n)
    return out


def gsa_batch(count: int = 20000, seed: int = 5, median: float = 4096.0, min_len: float = 256.0,
              max_len: float = float(4 << 20), target_total: float = float(512 << 20),
              shared_frac: float = 0.10):
    """Concatenated separator-terminated strings and their start offsets (count+1)."""
    cap = int(target_total * 1.5) + count * 2 + (8 << 20)
    out = np.empty(cap, dtype=np.uint8)
    offs = np.empty(count + 1, dtype=np.int64)
    total = _lib().sagen_gsa_batch(out.ctypes.data, cap, offs.ctypes.data, count, seed, median, min_len,
                                   max_len, target_total, shared_frac)
    if total < 0:
        raise MemoryError("sagen_gsa_batch: capacity exceeded")
    return out[:total].copy(), offs
