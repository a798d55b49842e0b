"""Batched generalized suffix arrays: every string's SA and LCP equal libsais_gsa +
libsais_plcp_gsa + libsais_lcp on that string alone (oracle), and the batched call
equals the per-string path bit for bit."""
import numpy as np
import pytest

import libsais_amd as la
from libsais_amd import batch, workloads
from tests import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _check_against_oracle(T, offs, SA, L, sample=None):
    ks = range(len(offs) - 1) if sample is None else sample
    for k in ks:
        o, e = int(offs[k]), int(offs[k + 1])
        Tk = np.ascontiguousarray(T[o:e])
        rc, SAk = ol.sa(Tk, gsa=True)
        assert rc == 0
        assert np.array_equal(SA[o:e], SAk), k
        Lk = ol.lcp(ol.plcp(Tk, SAk, gsa=True), SAk)
        assert np.array_equal(L[o:e], Lk), k


def test_small_batch_matches_oracle():
    T, offs = workloads.gsa_batch(count=300, seed=5, median=200, min_len=1, max_len=5000, target_total=80_000,
                                  shared_frac=0.2)
    SA, L = batch.batched(T, offs)
    _check_against_oracle(T, offs, SA, L)


def test_batched_equals_plain_path():
    T, offs = workloads.gsa_batch(count=150, seed=8, median=1000, min_len=16, max_len=20000, target_total=200_000,
                                  shared_frac=0.3)
    a = batch.batched(T, offs)
    b = batch.plain(T, offs)
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1], b[1])


def test_degenerate_strings_in_batch():
    parts = [b"a", b"aaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaa" * 20, b"ab" * 700, b"b", b"abracadabra" * 50]
    T = np.frombuffer(b"".join(p + b"\x00" for p in parts), np.uint8).copy()
    offs = np.zeros(len(parts) + 1, np.int64)
    offs[1:] = np.cumsum([len(p) + 1 for p in parts])
    SA, L = batch.batched(T, offs)
    _check_against_oracle(T, offs, SA, L)


def test_invalid_offsets_rejected():
    T = np.frombuffer(b"ab\x00cd\x00", np.uint8).copy()
    dev = batch.DeviceText(T)
    try:
        for offs in ([0, 2, 6], [0, 3, 5], [0, 1, 6], [1, 3, 6]):
            assert batch.batched_on_device(dev, np.array(offs, np.int64)) == -1
        assert batch.batched_on_device(dev, np.array([0, 3, 6], np.int64)) == 0
    finally:
        dev.free()


def test_medium_batch_sampled_against_reference():
    T, offs = workloads.gsa_batch(count=2000, seed=5, median=4096, min_len=256, max_len=1 << 20,
                                  target_total=16 << 20, shared_frac=0.1)
    SA, L = batch.batched(T, offs)
    rng = np.random.default_rng(0)
    sample = sorted(rng.choice(len(offs) - 1, 60, replace=False))
    if ol.have_ref():
        for k in sample:
            o, e = int(offs[k]), int(offs[k + 1])
            Tk = np.ascontiguousarray(T[o:e])
            rc, SAk, _ = ol.ref_sa(Tk, gsa=True)
            assert rc == 0 and np.array_equal(SA[o:e], SAk), k
            Pk = ol.ref_plcp(Tk, SAk, gsa=True)
            assert np.array_equal(L[o:e], ol.ref_lcp(Pk, SAk)), k
    else:
        _check_against_oracle(T, offs, SA, L, sample)
