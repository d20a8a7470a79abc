"""The library's host code (checkpoint reader and writer, SA-AMG setup, RCM)
built with AddressSanitizer and UBSan runs clean on valid input, corrupted
checkpoint files and malformed CSR input (tests/sanitize_driver.py, in a
child process with the sanitizer runtime preloaded).  CPU only."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "simall_amd", "_build", "asan", "libsimall_host_asan.so")


def _runtime(name):
    out = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True)
    path = out.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.mark.skipif(not os.path.exists(LIB), reason="sanitized host library not built")
def test_host_code_under_asan_and_ubsan(tmp_path):
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    if not asan or not ubsan:
        pytest.skip("g++ sanitizer runtimes not found")
    env = dict(os.environ)
    env["LD_PRELOAD"] = f"{asan}:{ubsan}"
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "sanitize_driver.py"), LIB,
                        str(tmp_path)], env=env, capture_output=True, text=True, timeout=600)
    log = r.stdout + r.stderr
    assert r.returncode == 0, log[-4000:]
    assert "AddressSanitizer" not in log and "runtime error" not in log, log[-4000:]
    assert "corrupted files" in r.stdout
