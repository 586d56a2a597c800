"""The oracle's own tests again, against an AddressSanitizer +
UndefinedBehaviorSanitizer build of the oracle (host code only): every
out-of-bounds access, use after free, signed overflow or misaligned access in
the C restatement aborts the run.  The sanitized library replaces the plain
one through OQ_ORACLE_LIB in a child process that preloads the ASan runtime
(a library that dlopen loads cannot bring it in itself)."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ORACLE_TESTS = ["test_oracle_okada.py", "test_oracle_hex8.py", "test_oracle_rhs.py",
                "test_oracle_tsit5.py", "test_oracle_vcabm.py"]


def _runtime():
    p = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True)
    path = p.stdout.strip()
    return path if p.returncode == 0 and os.path.isabs(path) and os.path.exists(path) else None


def test_oracle_under_address_and_ub_sanitizers():
    asan = _runtime()
    if asan is None:
        pytest.skip("gcc's AddressSanitizer runtime is not installed")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    env = dict(os.environ)
    env["OQ_ORACLE_LIB"] = os.path.join(ROOT, "oracle", "lib", "liboetqf_oracle_san.so")
    env["LD_PRELOAD"] = asan + (" " + env["LD_PRELOAD"] if env.get("LD_PRELOAD") else "")
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"  # the interpreter's own allocations are not ours
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu"]
                       + [os.path.join(HERE, t) for t in ORACLE_TESTS],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    tail = (r.stdout + r.stderr)[-3000:]
    assert r.returncode == 0, tail
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, tail
    assert " passed" in r.stdout, tail
