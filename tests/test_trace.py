"""Trace ranges (the reference's Vampir trace events, vpt.hpp:29-83) through
roctx: M4G_TRACE_LEVEL selects the events as MTL_VPT_LEVEL does, unset leaves
the tracing library unloaded, and a bad level is reported and ignored.  The
calls here fail their argument checks, so no GPU is touched; the ranges open
and close around them all the same."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "mtl4-mirror_amd", "lib", "libmtl4gpu.so")

PROBE = r"""
import ctypes, sys
lib = ctypes.CDLL(sys.argv[1])
r = ctypes.c_double()
# two_norm (2039) wraps unary_dot; cg_without_pc (7001) with null arguments
st1 = lib.m4g_two_norm(None, ctypes.byref(r))
st2 = lib.m4g_cg(None, None, None, None, None)
maps = open("/proc/self/maps").read()
print(st1, st2, int("roctx" in maps))
"""


def run(level):
    env = dict(os.environ)
    env.pop("M4G_TRACE_LEVEL", None)
    if level is not None:
        env["M4G_TRACE_LEVEL"] = level
    p = subprocess.run([sys.executable, "-c", PROBE, LIB], env=env, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    st1, st2, loaded = (int(v) for v in p.stdout.split())
    return st1, st2, loaded, p.stderr


@pytest.fixture(scope="module", autouse=True)
def built():
    if not os.path.exists(LIB):
        pytest.skip("library not built")


def test_trace_off_by_default_loads_nothing():
    st1, st2, loaded, err = run(None)
    assert st1 != 0 and st2 != 0
    assert loaded == 0
    assert err == ""


def test_trace_level_loads_roctx_and_keeps_status():
    ref = run(None)
    st1, st2, loaded, err = run("2")
    assert (st1, st2) == ref[:2]
    assert loaded == 1
    assert err == ""


def test_trace_level_above_every_event_loads_nothing():
    # level 9: only events above 9000 (none of the library's) are recorded
    _, _, loaded, _ = run("9")
    assert loaded == 0


@pytest.mark.parametrize("bad", ["x", "-1", "10", "2.5"])
def test_trace_bad_level_is_reported_and_ignored(bad):
    st1, st2, loaded, err = run(bad)
    assert "M4G_TRACE_LEVEL" in err
    assert loaded == 0
