"""The radix sort primitive on device buffers, against numpy's stable sort."""
import ctypes

import numpy as np
import pytest

import libsais_amd as la

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[(0, 256, 8), (0, 512, 8), (1, 256, 8), (0, 256, 4), (0, 256, 11)],
                ids=["onesweep", "onesweep_512", "reduce_then_scan", "digits4", "digits11"])
def variant(request):
    lib = la.load()
    v, threads, bits = request.param
    assert lib.sagpu_set_option(b"radix_variant", v) == 0
    assert lib.sagpu_set_option(b"radix_threads", threads) == 0
    assert lib.sagpu_set_option(b"radix_bits", bits) == 0
    yield request.param
    lib.sagpu_set_option(b"radix_variant", 0)
    lib.sagpu_set_option(b"radix_threads", 512)
    lib.sagpu_set_option(b"radix_bits", 8)


def _dev(arr):
    lib = la.load()
    p = lib.sagpu_malloc(arr.nbytes)
    assert p
    assert lib.sagpu_memcpy_h2d(p, arr.ctypes.data, arr.nbytes) == 0
    return p


def _host(p, like):
    out = np.empty_like(like)
    assert la.load().sagpu_memcpy_d2h(out.ctypes.data, p, out.nbytes) == 0
    return out


@pytest.mark.parametrize("n", [1, 2, 1000, 4096, 4097, 100_000, 3_000_001])
@pytest.mark.parametrize("dtype,bits", [(np.uint32, (0, 32)), (np.uint32, (4, 20)), (np.uint64, (0, 64)),
                                        (np.uint64, (13, 47))])
def test_sort_pairs_matches_stable_sort(variant, n, dtype, bits):
    rng = np.random.default_rng(n + bits[0])
    if dtype == np.uint32:
        keys = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    else:
        keys = rng.integers(0, 2**63, n, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, n, dtype=np.uint64)
    keys[: n // 3] = keys[0]  # many equal keys: stability matters
    vals = np.arange(n, dtype=np.uint32)
    lo, hi = bits
    mask = (1 << (hi - lo)) - 1
    digit = (keys.astype(np.uint64) >> np.uint64(lo)) & np.uint64(mask)
    order = np.argsort(digit, kind="stable")
    lib = la.load()
    dk, dv = _dev(keys), _dev(vals)
    f = lib.sagpu_sort_pairs_u32 if dtype == np.uint32 else lib.sagpu_sort_pairs_u64
    assert f(dk, dv, n, lo, hi) == 0, la.last_error()
    got_k, got_v = _host(dk, keys), _host(dv, vals)
    lib.sagpu_free(dk)
    lib.sagpu_free(dv)
    assert np.array_equal(got_v, vals[order])
    assert np.array_equal(got_k, keys[order])


def test_sort_bench_runs(variant):
    lib = la.load()
    ms, mp = ctypes.c_double(), ctypes.c_double()
    assert lib.sagpu_sort_bench(1 << 22, 4, 32, 2, ctypes.byref(ms), ctypes.byref(mp)) == 0
    assert ms.value > 0 and mp.value > 0
