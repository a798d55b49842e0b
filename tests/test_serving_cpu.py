"""Serving trace construction (CPU only): deterministic and shaped as specified."""
import numpy as np

from libsais_amd import serving


def test_trace_is_deterministic_and_shaped():
    spec = serving.TraceSpec(jobs=200, rate=200.0, min_bytes=4 << 10, max_bytes=1 << 20, seed=6)
    a = serving.make_trace(spec)
    b = serving.make_trace(spec)
    assert [len(r.T) for r in a] == [len(r.T) for r in b]
    assert [r.kind for r in a] == [r.kind for r in b]
    assert all(np.array_equal(x.T, y.T) for x, y in zip(a[:20], b[:20]))
    sizes = np.array([len(r.T) for r in a])
    assert sizes.min() >= 4 << 10 and sizes.max() <= 1 << 20
    arr = np.array([r.arrival for r in a])
    assert arr[0] == 0 and (np.diff(arr) >= 0).all()
    assert 0.5 < arr[-1] / (len(a) / spec.rate) < 2.0
    kinds = np.bincount([r.kind for r in a], minlength=3) / len(a)
    assert abs(kinds[0] - 0.4) < 0.12 and abs(kinds[2] - 0.2) < 0.12
    for r in a:
        assert set(np.unique(r.T)) <= set(b"ACGT")


def test_job_struct_layout():
    import ctypes

    assert ctypes.sizeof(serving.Job) == 96
    assert serving.Job.status.offset == 56 and serving.Job.t_submit.offset == 64
