"""The rule by which the initial grouping marks slot boundaries (k_init_emit's
lcp_mark, rounds.hip), checked on the CPU against the oracle's LCP.

After the initial sort by the packed first ksym symbols, slot e and slot e-1 hold
different keys or the same one.  A known mark must equal the final LCP[e]; an
unknown mark's bound, capped as the PLCP pass caps it (by the shorter of the two
final suffixes), must not exceed LCP[e].  The rule: the lcp of the keys, reduced
to a suffix's length where that suffix is alone in its group (its key is padded
past its end); a slot of a larger group whose key has trailing code-0 symbols
reaching back to the keys' lcp leaves the pair unknown.  The slot before the first
slot of each span of the grouping pass counts as grouped.  This restates the
rule, not the kernel: the GPU parity tests check the kernel's output.
"""
import numpy as np
import pytest

from tests import oracle_lib as ol


def _marks_ok(T, span):
    n = len(T)
    alpha = sorted(set(T.tolist()))
    code = np.zeros(256, np.int64)
    for i, c in enumerate(alpha):
        code[c] = i
    bits = max(1, (len(alpha) - 1).bit_length())
    ksym = 32 // bits
    pad = np.concatenate([code[T], np.zeros(ksym, np.int64)])
    keys = np.zeros(n, np.int64)
    for t in range(ksym):
        keys = (keys << bits) | pad[t:t + n]
    order = np.lexsort((np.arange(n), keys))   # the stable initial sort
    ks = keys[order]
    rc, SA = ol.sa(T)
    L = ol.lcp(ol.plcp(T, SA), SA)
    alone = np.ones(n, bool)
    alone[1:] &= ks[1:] != ks[:-1]
    alone[:-1] &= ks[1:] != ks[:-1]
    mask = (1 << bits) - 1

    def pad_tail(k):
        c = 0
        while c < ksym and ((int(k) >> (bits * c)) & mask) == 0:
            c += 1
        return c

    known = 0
    for e in range(1, n):
        x = int(ks[e]) ^ int(ks[e - 1])
        if x == 0:
            continue                                  # one group: LCP_UNKNOWN | ksym
        l = (32 - x.bit_length() - (32 - ksym * bits)) // bits
        onep = alone[e - 1] and e % span != 0
        m, unknown = l, False
        if alone[e]:
            m = min(m, n - order[e])
        elif pad_tail(ks[e]) + l >= ksym:
            unknown = True
        if onep:
            m = min(m, n - order[e - 1])
        elif pad_tail(ks[e - 1]) + l >= ksym:
            unknown = True
        if unknown:
            assert min(l, n - max(SA[e], SA[e - 1])) <= L[e], (e, l, L[e])
        else:
            assert m == L[e], (e, m, L[e])
            known += 1
    return known


def _cases():
    rng = np.random.default_rng(11)
    out = []
    for a in (b"AC", b"ACGT", b"ACGTN", bytes(range(40)), bytes(range(256))):
        sym = np.frombuffer(a, np.uint8)
        for n in (2, 3, 5, 17, 33, 100, 1000, 5000):
            out.append((f"iid{len(a)}_{n}", sym[rng.integers(0, len(a), n)].copy()))
            tail = np.full(int(rng.integers(1, 40)), a[0], np.uint8)   # runs of code 0 at the end
            out.append((f"tailrun{len(a)}_{n}", np.concatenate([sym[rng.integers(0, len(a), n)], tail])))
    unit = np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, 48)]
    out.append(("tail_repeat", np.concatenate([np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, 3000)], unit,
                                               np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, 2000)], unit[:20]])))
    out.append(("all_a", np.full(3000, 65, np.uint8)))
    out.append(("acgt17", np.frombuffer(b"GCCTAAGACATATAAAA", np.uint8).copy()))
    return out


@pytest.mark.parametrize("name,T", _cases(), ids=[c[0] for c in _cases()])
def test_initial_marks_against_oracle(name, T):
    for span in (64, 1024):
        _marks_ok(T, span)
