#!/bin/bash
# The CPU test suite (pytest -m "not gpu") against AddressSanitizer +
# UndefinedBehaviorSanitizer builds of the oracle and of the HIP library's
# host code (label parsing and validation, plan layout, bundles, view
# checks, the C ABI), as the reference runs its tests under
# -fsanitize=address,undefined.  Runs on a machine without a GPU.
#   bash tests/run_sanitized.sh [extra pytest args]
set -euo pipefail
ROOT="$(cd "$(dirname "$0")/.." && pwd)"
make -s -C "$ROOT/oracle" sanitize
make -s -j8 -C "$ROOT/t1m_amd" sanitize
RT="$(/opt/rocm/llvm/bin/clang -print-file-name=libclang_rt.asan-x86_64.so)"
export T1M_LIB="$ROOT/t1m_amd/_build_san/libt1m_hip.so"
export T1M_ORACLE_LIB="$ROOT/oracle/_build_san/libt1m_oracle.so"
# Python itself is not instrumented: leaks at exit are its own.
export ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:halt_on_error=1"
export UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1"
cd "$ROOT"
LD_PRELOAD="$RT" python -m pytest tests -x -q -m "not gpu" -p no:cacheprovider "$@"
