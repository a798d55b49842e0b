"""The autotuner: every grid point gives output identical to the default knobs
(and to the oracle), and the chosen configuration is one of the measured ones."""
import numpy as np
import pytest

import libsais_amd as la
from libsais_amd import autotune, workloads
from tests import oracle_lib as ol

pytestmark = pytest.mark.gpu


def test_tune_text_grid_identical_outputs():
    T = workloads.dna_like(1 << 20, seed=21)
    res = autotune.tune_text(T, reps=1)
    assert len(res["results"]) == 27 and all(r["identical"] for r in res["results"])
    assert {r["config"]["radix_bits"] for r in res["results"]} == {4, 8, 11}
    assert res["best"] in [{**autotune.DEFAULT, **r["config"]} for r in res["results"]]
    # non-default points against the oracle
    autotune.apply({"radix_items": 8, "radix_variant": 1, "tile_range": 256})
    try:
        rc, SA_ref = ol.sa(T)
        assert np.array_equal(la.suffix_array(T), SA_ref)
    finally:
        autotune.apply(autotune.DEFAULT)
    autotune.apply({"radix_bits": 11, "radix_items": 16, "tile_range": 1024})
    try:
        rc, SA_ref = ol.sa(T)
        assert np.array_equal(la.suffix_array(T), SA_ref)
    finally:
        autotune.apply(autotune.DEFAULT)


def test_tune_batched_small():
    T, offs = workloads.gsa_batch(count=300, seed=7, median=1000, min_len=16, max_len=20000,
                                  target_total=300_000, shared_frac=0.2)
    grid = {"radix_items": (16, 32), "tile_range": (256, 1024)}
    res = autotune.tune_batched(T, offs, grid=grid, reps=1)
    assert len(res["results"]) == 4 and res["best"]["tile_range"] in (256, 1024)
    autotune.apply(res["best"])
    try:
        from libsais_amd import batch
        SA, L = batch.batched(T, offs)
        for k in range(0, len(offs) - 1, 29):
            o, e = int(offs[k]), int(offs[k + 1])
            rc, SAk = ol.sa(np.ascontiguousarray(T[o:e]), gsa=True)
            assert np.array_equal(SA[o:e], SAk)
    finally:
        autotune.apply(autotune.DEFAULT)


def test_knob_validation():
    lib = la.load()
    for name, bad in ((b"radix_items", 12), (b"tile_range", 300), (b"radix_variant", 2), (b"radix_threads", 384),
                      (b"radix_bits", 6), (b"round_graphs", 2), (b"windowed_scatter", -1),
                      (b"big_pivot_min", -1)):
        assert lib.sagpu_set_option(name, bad) == -1
