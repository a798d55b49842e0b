"""A C program built against the public headers and linked with -lsais.

tests/c_caller/caller.c includes include/libsais.h and include/libsais64.h as an
existing libsais user does and calls the library through the C ABI (no ctypes).
Its CPU half checks the argument checks and n <= 1 cases; its GPU half compares
SA, BWT, PLCP and LCP with the oracle on seeded texts, including U aliasing T.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CALLER = os.path.join(HERE, "c_caller", "caller")


def _caller():
    if not os.path.exists(CALLER):
        subprocess.run(["make", "-C", os.path.join(HERE, "c_caller")], check=True, capture_output=True)
    return CALLER


def _run(mode):
    r = subprocess.run([_caller(), mode], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().splitlines()[-1].startswith("ok ")


def test_c_caller_cpu_cases():
    _run("cpu")


@pytest.mark.gpu
def test_c_caller_gpu_cases():
    _run("gpu")
