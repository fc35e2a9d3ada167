#!/usr/bin/env bash
# Runs the CPU test suite (pytest -m "not gpu") on sanitizer builds of the
# host code: the CPU oracle (oracle/lib_san) and the C++ facade with its
# suites (latticeboltzmanncuda_amd/lib_san, bin_san), built with g++/gcc
# -fsanitize=address,undefined and every report fatal.  The HIP library is
# the ordinary build: no GPU code is instrumented and no GPU call is made.
#
#   tests/run_sanitized.sh [extra pytest arguments]
set -euo pipefail
ROOT="$(cd "$(dirname "${BASH_SOURCE[0]}")/.." && pwd)"
cd "$ROOT"
make -s -C oracle sanitize
python3 -m latticeboltzmanncuda_amd.build --sanitize > /dev/null
# Python itself is not instrumented: the sanitizer runtimes are preloaded so
# that the instrumented oracle can be loaded into it; leak checking is off
# (the interpreter keeps its allocations until exit)
ASAN_RT="$(gcc -print-file-name=libasan.so)"
UBSAN_RT="$(gcc -print-file-name=libubsan.so)"
export LD_PRELOAD="$ASAN_RT:$UBSAN_RT"
export ASAN_OPTIONS="detect_leaks=0:abort_on_error=1"
export UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1"
export LBM_ORACLE_LIB="$ROOT/oracle/lib_san/liblbm_oracle.so"
export LBM_FACADE_BIN="$ROOT/latticeboltzmanncuda_amd/bin_san"
exec python3 -m pytest tests -q -m "not gpu" -p no:cacheprovider "$@"
