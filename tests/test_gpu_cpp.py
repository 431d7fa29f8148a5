"""Runs the C++ programs of tests/cpp (written against the reference's call
surface, built with the drop-in headers) on the GPU; each must exit 0."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "bin")
PROGRAMS = sorted(os.listdir(BIN)) if os.path.isdir(BIN) else []


def test_programs_built():
    assert len(PROGRAMS) >= 9, "tests/cpp not built (run __graft_entry__.build())"


@pytest.mark.parametrize("prog", PROGRAMS)
def test_cpp_program(prog):
    print(f"running {prog}", flush=True)
    r = subprocess.run([os.path.join(BIN, prog)], capture_output=True, text=True, timeout=100)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, (prog, r.returncode, r.stderr[-500:])
