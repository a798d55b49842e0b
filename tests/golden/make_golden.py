"""Writes the golden vectors in this directory from the compiled reference libsais.

Run from the repository root after `make -C oracle ref` (needs the reference
sources, so it runs in the development container, not on the GPU host):

    python tests/golden/make_golden.py

Outputs
  small.npz     inputs and reference outputs (SA, BWT, primary, PLCP, LCP, freq) of
                small named strings, plus GSA cases (SA, PLCP)
  golden.json   sha256 digests of the reference outputs for the 1 MiB i.i.d. ACGT
                string (seed 1) and the 64-bit API on the same input
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from libsais_amd import workloads  # noqa: E402
from tests import oracle_lib as ol  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def small_cases():
    rng = np.random.default_rng(2024)
    cases = {
        "banana": np.frombuffer(b"banana", np.uint8),
        "mississippi": np.frombuffer(b"mississippi", np.uint8),
        "abracadabra": np.frombuffer(b"abracadabra", np.uint8),
        "all_a_37": np.full(37, 65, np.uint8),
        "period17_500": workloads.periodic(500, 17, 3),
        "fibonacci_610": workloads.fibonacci(610),
        "acgt_1000": workloads.iid(1000, 7, b"ACGT"),
        "bytes256_2000": rng.integers(0, 256, 2000).astype(np.uint8),
        "with_zeros_700": rng.integers(0, 3, 700).astype(np.uint8),
        "binary_300": rng.integers(0, 2, 300).astype(np.uint8),
    }
    gsa = {
        "gsa_three": np.frombuffer(b"banana\x00ananas\x00nab\x00", np.uint8),
        "gsa_acgt": np.concatenate([np.append(workloads.iid(int(rng.integers(1, 30)), s, b"ACGT"), 0)
                                    for s in range(40)]).astype(np.uint8),
        "gsa_repeat": np.tile(np.frombuffer(b"ab\x00", np.uint8), 50),
    }
    return cases, gsa


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    if not ol.have_ref():
        raise SystemExit("oracle/_ref/libsais_ref.so missing: run `make -C oracle ref` first")
    cases, gsa = small_cases()
    out = {}
    for name, T in cases.items():
        T = np.ascontiguousarray(T, dtype=np.uint8)
        rc, SA, freq = ol.ref_sa(T)
        assert rc == 0
        U, idx = ol.ref_bwt(T)
        P = ol.ref_plcp(T, SA)
        L = ol.ref_lcp(P, SA)
        out[f"{name}__T"] = T
        out[f"{name}__SA"] = SA
        out[f"{name}__U"] = U
        out[f"{name}__primary"] = np.array([idx], np.int64)
        out[f"{name}__PLCP"] = P
        out[f"{name}__LCP"] = L
        out[f"{name}__freq"] = freq
    for name, T in gsa.items():
        T = np.ascontiguousarray(T, dtype=np.uint8)
        rc, SA, freq = ol.ref_sa(T, gsa=True)
        assert rc == 0, name
        P = ol.ref_plcp(T, SA, gsa=True)
        out[f"{name}__T"] = T
        out[f"{name}__SA"] = SA
        out[f"{name}__PLCP"] = P
        out[f"{name}__freq"] = freq
    np.savez_compressed(os.path.join(HERE, "small.npz"), **out)

    T = workloads.iid_acgt(1 << 20, seed=1)
    rc, SA, freq = ol.ref_sa(T)
    U, idx = ol.ref_bwt(T)
    P = ol.ref_plcp(T, SA)
    L = ol.ref_lcp(P, SA)
    golden = {
        "acgt_1MiB_seed1": {
            "n": len(T),
            "T_sha256": sha(T),
            "SA_int32_sha256": sha(SA),
            "U_sha256": sha(U),
            "primary": int(idx),
            "PLCP_int32_sha256": sha(P),
            "LCP_int32_sha256": sha(L),
            "SA_int64_sha256": sha(SA.astype(np.int64)),
            "freq": [int(x) for x in freq],
            "source": "reference libsais 2.10.4 (oracle/_ref) libsais, libsais_bwt, libsais_plcp, libsais_lcp",
        }
    }
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(golden, f, indent=1)
    print("wrote", len(out), "arrays and golden.json")


if __name__ == "__main__":
    main()
