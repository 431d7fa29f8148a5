#!/bin/bash
# Builds the host-only parts of the library (sweep schedules, analysis and
# emulation, generators) and the CPU oracle with AddressSanitizer and
# UndefinedBehaviorSanitizer, then runs the CPU tests that exercise them
# (schedule emulation, grid inference, generators, the oracle's own tests)
# against those builds.  Any sanitizer report ends the run with an error.
#
#   tests/sanitize_host.sh [extra pytest arguments]
set -euo pipefail
ROOT="$(cd "$(dirname "$0")/.." && pwd)"
OUT="$ROOT/tests/_san"
CSRC="$ROOT/mtl4-mirror_amd/csrc"
CXX="${CXX:-g++}"
SAN="-fsanitize=address,undefined -fno-sanitize-recover=undefined -fno-omit-frame-pointer"
mkdir -p "$OUT"

$CXX -O1 -g -std=c++17 -fPIC -Wall $SAN -shared -o "$OUT/libmtl4gpu_host_san.so" \
    "$CSRC/analysis.cpp" "$CSRC/schedule.cpp" "$CSRC/generators.cpp" "$CSRC/errors.cpp" "$CSRC/trace.cpp" -lpthread -ldl
$CXX -O1 -g -std=c++17 -fPIC -ffp-contract=off -Wall $SAN -shared -o "$OUT/liboracle_san.so" \
    "$ROOT/oracle/mtl4_oracle.cpp"

ASAN_LIB="$($CXX -print-file-name=libasan.so)"
export LD_PRELOAD="$ASAN_LIB"
# Python's own allocations are not instrumented and never freed at exit
export ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:halt_on_error=1"
export UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1"
export M4G_LIBRARY="$OUT/libmtl4gpu_host_san.so"
export M4G_ORACLE_LIBRARY="$OUT/liboracle_san.so"
cd "$ROOT"
exec python3 -m pytest -q -p no:cacheprovider tests/test_host.py tests/test_oracle.py \
    --deselect tests/test_host.py::test_library_exports_every_declared_symbol "$@"
