"""Host code under sanitizers (CPU only).

The CPU tests of the C wrappers (test_abi.py: argument checks and n <= 1 cases,
compared with the reference), of the oracle (test_oracle.py) and of the host copy
pool (test_copy_pool.py) run again in a child process against the sanitizer
builds of the same sources:
  - AddressSanitizer + UndefinedBehaviorSanitizer: libsais_amd/lib/san/libsais.so
    (api.c, api16.c and engine.hip's host side instrumented) and
    oracle/lib/san/liboracle.so;
  - ThreadSanitizer: libsais_amd/lib/tsan/libsais.so, on the copy pool (its worker
    threads, and several callers sharing it).  The fork case is left out: the
    ThreadSanitizer runtime does not support starting threads in a child forked from
    a multi-threaded process.
Any sanitizer report ends the child with a failure.  The builds come from
`make -C libsais_amd/csrc sanitize` and `make -C oracle sanitize` (done by
__graft_entry__.build()); the runtimes are clang's, from the ROCm toolchain.
"""
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN_LIB = os.path.join(ROOT, "libsais_amd", "lib", "san", "libsais.so")
TSAN_LIB = os.path.join(ROOT, "libsais_amd", "lib", "tsan", "libsais.so")
SAN_ORACLE = os.path.join(ROOT, "oracle", "lib", "san", "liboracle.so")


def _runtime(name):
    hits = sorted(glob.glob(f"/opt/rocm/lib/llvm/lib/clang/*/lib/linux/libclang_rt.{name}-x86_64.so"))
    return hits[-1] if hits else None


def _run(env_extra, args):
    env = dict(os.environ)
    env.update(env_extra)
    cmd = [sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu"] + args
    return subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)


def _check_loaded(env_extra, lib):
    """The child really maps the instrumented library (not the plain one)."""
    code = ("import libsais_amd as la; la.load(); "
            "print(any(l.rstrip().endswith(%r) for l in open('/proc/self/maps')))" % lib)
    env = dict(os.environ)
    env.update(env_extra)
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, capture_output=True, text=True,
                         timeout=120)
    assert out.stdout.strip().endswith("True"), out.stdout + out.stderr


@pytest.mark.skipif(not (os.path.exists(SAN_LIB) and os.path.exists(SAN_ORACLE) and _runtime("asan")),
                    reason="sanitizer builds or the clang ASan runtime missing")
def test_address_and_undefined_sanitizers():
    env = {"LD_PRELOAD": _runtime("asan"),
           "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1",
           "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1",
           "LIBSAIS_AMD_LIB": SAN_LIB, "LIBSAIS_ORACLE_LIB": SAN_ORACLE}
    _check_loaded(env, SAN_LIB)
    r = _run(env, ["tests/test_abi.py", "tests/test_oracle.py", "tests/test_copy_pool.py"])
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr


@pytest.mark.skipif(not (os.path.exists(TSAN_LIB) and _runtime("tsan")),
                    reason="ThreadSanitizer build or the clang TSan runtime missing")
def test_thread_sanitizer_copy_pool():
    env = {"LD_PRELOAD": _runtime("tsan"), "TSAN_OPTIONS": "halt_on_error=1:report_signal_unsafe=0",
           "LIBSAIS_AMD_LIB": TSAN_LIB}
    _check_loaded(env, TSAN_LIB)
    r = _run(env, ["tests/test_copy_pool.py", "-k", "not fork"])
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "WARNING: ThreadSanitizer" not in r.stderr
