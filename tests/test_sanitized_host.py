"""Runs tests/sanitize_host.sh: the host-only library code (sweep schedules,
their CPU emulation, grid inference, generators) and the CPU oracle built with
AddressSanitizer and UndefinedBehaviorSanitizer, under the CPU tests that
exercise them.  A sanitizer report fails the run."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_host_code_under_asan_ubsan():
    cxx = os.environ.get("CXX", "g++")
    if shutil.which(cxx) is None:
        pytest.skip(f"{cxx} not available")
    asan = subprocess.run([cxx, "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    if not os.path.isabs(asan) or not os.path.exists(asan):
        pytest.skip("libasan not available")
    env = {k: v for k, v in os.environ.items() if k not in ("M4G_LIBRARY", "M4G_ORACLE_LIBRARY")}
    p = subprocess.run([os.path.join(HERE, "sanitize_host.sh"), "-x"], capture_output=True, text=True,
                       timeout=900, env=env)
    tail = (p.stdout + p.stderr)[-4000:]
    assert p.returncode == 0, tail
    assert "AddressSanitizer" not in tail and "runtime error" not in tail, tail
    assert " passed" in tail
