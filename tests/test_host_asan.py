"""The host C layer (sba_host.c, sba_div.c) built with AddressSanitizer
(gcc, host code only: simd_bigint_amd/csrc/Makefile target `asan`) and
run on the CPU: the CPU multiply tiers against a schoolbook and the error
contract's fallback paths with injected GPU failures (tests/c/host_asan.c).
ASan aborts on any invalid access or leak."""
import os
import subprocess

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
CSRC = os.path.join(ROOT, "simd_bigint_amd", "csrc")
BIN = os.path.join(ROOT, "simd_bigint_amd", "_build", "host_asan_test")


def test_host_layer_under_asan():
    b = subprocess.run(["make", "-C", CSRC, "asan"], capture_output=True, text=True, timeout=900)
    if b.returncode != 0 and "fsanitize" in b.stderr and "not supported" in b.stderr:
        pytest.skip("no ASan runtime for the host compiler")
    assert b.returncode == 0, b.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0")
    p = subprocess.run([BIN], capture_output=True, text=True, timeout=600, env=env)
    assert p.returncode == 0, (p.stdout + p.stderr)[-4000:]
    assert "all checks passed" in p.stdout
    assert "ERROR: AddressSanitizer" not in p.stderr and "LeakSanitizer" not in p.stderr
