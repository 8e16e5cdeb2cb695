"""The host code of the C-ABI library and the oracle under AddressSanitizer
and UndefinedBehaviorSanitizer.

Both are rebuilt with ROCm's clang and host-only sanitizer flags
(`make -C mojoblas_amd/csrc sanitize`, `make -C oracle sanitize`, into
_build/san), and the CPU tests of the boundary and the oracle run again in a
child process with the sanitizer runtime preloaded and the sanitized builds
loaded in place of the shipped ones (tests/conftest.py): every validation
path, info code, quick return, buffer entry, thread-precision and
no-allocation check of the C-ABI, and every oracle and CPU-baseline routine.
Any sanitizer report fails the run.  A planted out-of-bounds read in a
child of its own shows that the runtime is live.
"""
import glob
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
RT = sorted(glob.glob("/opt/rocm/lib/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))


def _env():
    env = dict(os.environ)
    env.update({"MOJOBLAS_SANITIZED_BUILD": "1", "LD_PRELOAD": RT[-1],
                "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1",
                "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1"})
    return env


@pytest.fixture(scope="module")
def sanitized_builds():
    if not RT or not os.path.exists("/opt/rocm/bin/hipcc"):
        pytest.skip("ROCm clang sanitizer runtime or hipcc not available")
    subprocess.check_call(["make", "-s", "-j8", "-C", os.path.join(ROOT, "mojoblas_amd", "csrc"), "sanitize"],
                          stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"])


def test_boundary_and_oracle_cpu_tests_under_asan_ubsan(sanitized_builds):
    files = ["test_abi_cpu.py", "test_oracle.py", "test_oracle_level2.py", "test_oracle_pins.py",
             "test_mojo_shims.py"]
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider", "-s"]
                       + [os.path.join(HERE, f) for f in files], cwd=ROOT, env=_env(), capture_output=True,
                       text=True, timeout=900)
    tail = (r.stdout[-3000:] + r.stderr[-3000:])
    assert r.returncode == 0, tail
    assert "AddressSanitizer" not in r.stderr and "runtime error:" not in r.stderr, tail
    assert "SANITIZED_LIBS_LOADED" in r.stdout, tail  # the instrumented builds were the ones loaded


def test_sanitizer_catches_a_planted_overflow(sanitized_builds):
    code = ("import numpy as np, oracle, os\n"
            "oracle._LIB_PATH = os.path.join(%r, '_build', 'san', 'liboracle.so')\n"
            "a = np.zeros(4096, np.float64)\n"
            "c = np.zeros(64 * 64, np.float64)\n"
            "oracle.dgemm('N', 'N', 64, 64, 128, 1.0, a, 64, a, 128, 0.0, c, 64)\n"  # A needs 8192 elements
            "print('NOT CAUGHT')\n" % ROOT)
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=_env(), capture_output=True, text=True, timeout=300)
    assert r.returncode != 0 and "heap-buffer-overflow" in r.stderr, r.stderr[-2000:]
    assert "NOT CAUGHT" not in r.stdout
