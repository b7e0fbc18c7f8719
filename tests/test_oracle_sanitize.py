"""The FP64 oracle under AddressSanitizer and UndefinedBehaviorSanitizer (CPU): every
entry point the tests use, driven once on small domains by oracle/sanitize_main.cpp, must
run without a sanitizer report (out-of-bounds access, leak, undefined arithmetic)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_is_clean_under_asan_and_ubsan():
    if not shutil.which(os.environ.get("CXX", "g++")):
        pytest.skip("no host C++ compiler")
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"])
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(ROOT, "oracle", "_build", "oracle_sanitized")], capture_output=True, text=True,
                       env=env, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr
    assert "sanitized oracle run ok" in r.stdout
