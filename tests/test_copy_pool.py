"""The host copy pool behind the staged transfers of the libsais calls, exercised on the
CPU through sagpu_host_memcpy: every byte arrives for ragged sizes and any thread
count, several callers can share the pool at once, and a forked child (which has
none of the parent's worker threads) gets a working pool of its own."""
import os
import threading

import numpy as np
import pytest

import libsais_amd as la


def _copy(dst, src, threads):
    assert la.load().sagpu_host_memcpy(dst.ctypes.data, src.ctypes.data, src.nbytes, threads) == 0


@pytest.mark.parametrize("size", [0, 1, 4095, 1 << 20, (1 << 20) + 1, (37 << 20) + 3])
@pytest.mark.parametrize("threads", [0, 1, 2, 3, 16])
def test_copy_sizes(size, threads):
    rng = np.random.default_rng(size + threads)
    src = rng.integers(0, 256, size, dtype=np.uint8)
    dst = np.zeros(size, np.uint8)
    _copy(dst, src, threads)
    assert np.array_equal(dst, src)


def test_negative_threads():
    a = np.zeros(8, np.uint8)
    assert la.load().sagpu_host_memcpy(a.ctypes.data, a.ctypes.data, 8, -1) == -1


def test_concurrent_callers():
    """Four Python threads copy through the pool at once (ctypes releases the GIL)."""
    rng = np.random.default_rng(5)
    srcs = [rng.integers(0, 256, (9 << 20) + k, dtype=np.uint8) for k in range(4)]
    dsts = [np.zeros_like(s) for s in srcs]
    errors = []

    def work(k):
        try:
            for _ in range(5):
                dsts[k][:] = 0
                _copy(dsts[k], srcs[k], 6)
                if not np.array_equal(dsts[k], srcs[k]):
                    errors.append(k)
        except Exception as e:  # pragma: no cover - reported below
            errors.append(repr(e))

    ts = [threading.Thread(target=work, args=(k,)) for k in range(4)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(60)
    assert not any(t.is_alive() for t in ts)
    assert errors == []


def test_fork_child_gets_a_pool():
    src = np.arange(3 << 20, dtype=np.uint8)
    dst = np.zeros_like(src)
    _copy(dst, src, 4)          # the parent's workers exist now
    pid = os.fork()
    if pid == 0:                # child: copy again, exit 0 on success
        code = 1
        try:
            d2 = np.zeros_like(src)
            _copy(d2, src, 4)
            code = 0 if np.array_equal(d2, src) else 2
        finally:
            os._exit(code)
    _, status = os.waitpid(pid, 0)
    assert os.WIFEXITED(status) and os.WEXITSTATUS(status) == 0
