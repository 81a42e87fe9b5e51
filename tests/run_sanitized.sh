#!/bin/bash
# The CPU test suite under AddressSanitizer + UndefinedBehaviorSanitizer:
# builds the instrumented product library, driver and oracle
# (build_sanitized.py), preloads the shared sanitizer runtime into Python and
# points the loaders at the instrumented files.  CPU only: no GPU test runs.
# Usage: tests/run_sanitized.sh [extra pytest arguments]
set -euo pipefail
ROOT="$(cd "$(dirname "${BASH_SOURCE[0]}")/.." && pwd)"
PKG="$ROOT/amrex-euler-allmach-explicit-imex_amd"
eval "$(cd "$PKG" && python3 build_sanitized.py | sed 's/^/SAN_/')"
export EC_AMD_LIB="$SAN_lib" EC_ORACLE_LIB="$SAN_oracle" EC_AMD_DRIVER="$SAN_driver"
# Python and torch are not instrumented: leak reports would be theirs
export ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:halt_on_error=1"
export UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1"
cd "$ROOT"
LD_PRELOAD="$SAN_runtime" python3 -m pytest tests -q -m "not gpu" -p no:cacheprovider "$@"
