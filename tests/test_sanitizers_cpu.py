"""The host C++ (host.cpp, io.cpp) and the CPU oracle under AddressSanitizer
and UndefinedBehaviorSanitizer: builds them with -fsanitize=address,undefined
(tfhe_amd.build.build_sanitized; the GPU code is the normal build's) and runs
the host-side test modules against that build in a child process, gcc's
sanitizer runtimes preloaded, so that an invalid access or undefined behaviour
in key generation, the key-file readers (malformed files included), the
levelisation or the argument checks fails the run. The slowest case (the
N = 2048 key round trip, 30 s instrumented) is left to the normal build."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _runtime(name):
    path = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.mark.skipif(not (_runtime("libasan.so") and _runtime("libubsan.so")), reason="gcc sanitizer runtimes absent")
def test_host_code_under_asan_and_ubsan():
    sys.path.insert(0, ROOT)
    from tfhe_amd import build as B
    san = B.build_sanitized()
    env = dict(os.environ)
    env.update(TFHE_AMD_LIB=os.path.join(san, "libtfhe_amd.so"), TFHE_AMD_ORACLE_LIB=os.path.join(san, "liboracle.so"),
               LD_PRELOAD=f"{_runtime('libasan.so')} {_runtime('libubsan.so')}",
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_io_cpu.py"), os.path.join(ROOT, "tests", "test_scheduler_cpu.py"),
                        os.path.join(ROOT, "tests", "test_library_cpu.py"), "-k", "not pbig"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert " passed" in r.stdout and "error" not in r.stdout.lower().split("passed")[-1]
