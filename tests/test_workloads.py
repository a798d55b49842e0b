"""Workload generators: determinism and the structural properties each promises."""
import hashlib

import numpy as np

from libsais_amd import workloads
from tests import oracle_lib as ol


def _sha(a):
    return hashlib.sha256(a.tobytes()).hexdigest()


def test_deterministic():
    for f in (lambda: workloads.iid_acgt(100000, 1), lambda: workloads.dna_like(1 << 18, 2),
              lambda: workloads.markov(50000, 6, 2), lambda: workloads.periodic(10000, 17, 3)):
        assert _sha(f()) == _sha(f())
    assert _sha(workloads.iid_acgt(1000, 1)) != _sha(workloads.iid_acgt(1000, 2))


def test_iid_acgt_is_uniform():
    T = workloads.iid_acgt(1 << 20, 1)
    counts = np.bincount(T, minlength=256)
    assert set(np.nonzero(counts)[0]) == set(b"ACGT")
    for c in b"ACGT":
        assert abs(counts[c] / len(T) - 0.25) < 0.005


def test_dna_like_shape():
    n = 1 << 22
    T = workloads.dna_like(n, 2)
    counts = np.bincount(T, minlength=256)
    assert set(np.nonzero(counts)[0]) <= set(b"ACGT")
    gc = (counts[ord("C")] + counts[ord("G")]) / n
    assert 0.30 < gc < 0.65
    # copied segments and tandem repeats leave long repeated substrings
    rc, SA = ol.sa(T)
    P = ol.plcp(T, SA)
    assert (P >= 1000).sum() > n // 100


def test_periodic_and_fibonacci():
    n = 10000
    T = workloads.periodic(n, 17, 3)
    mism = [i for i in range(17, n) if T[i] != T[i - 17]]
    assert mism == [n // 2, n // 2 + 17]
    F = workloads.fibonacci(n)
    # reference construction: s_k = s_{k-1} s_{k-2}
    a, b = "a", "ab"
    while len(b) < n:
        a, b = b, b + a
    assert bytes(F) == b[:n].encode()
    assert (workloads.all_a(50) == 65).all()


def test_gsa_batch_structure():
    T, offs = workloads.gsa_batch(count=300, seed=5, median=512, min_len=16, max_len=4096,
                                  target_total=200_000, shared_frac=0.1)
    assert offs[-1] == len(T)
    assert T[-1] == 0
    assert ol.gsa_valid(T)
    lens = np.diff(offs) - 1
    assert lens.min() >= 16 and lens.max() <= 4096
    for s in range(300):
        assert T[offs[s + 1] - 1] == 0
        assert (T[offs[s]:offs[s + 1] - 1] != 0).all()
    assert abs(lens.sum() - 200_000) / 200_000 < 0.2
