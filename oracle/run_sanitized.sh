#!/bin/bash
# Runs the CPU test suite against the AddressSanitizer + UBSan build of the
# oracle (liboracle_asan.so).  CPU only: the GPU library is not involved.
#   bash oracle/run_sanitized.sh [extra pytest args]
set -e
HERE="$(cd "$(dirname "$0")" && pwd)"
REPO="$(dirname "$HERE")"
make -C "$HERE" liboracle_asan.so
ASAN_LIB="$(g++ -print-file-name=libasan.so)"
UBSAN_LIB="$(g++ -print-file-name=libubsan.so)"
cd "$REPO"
ORACLE_LIB="$HERE/liboracle_asan.so" LD_PRELOAD="$ASAN_LIB:$UBSAN_LIB" \
  ASAN_OPTIONS=detect_leaks=0:abort_on_error=1 UBSAN_OPTIONS=print_stacktrace=1:halt_on_error=1 \
  python -m pytest tests -m "not gpu" -q -p no:cacheprovider "$@"
