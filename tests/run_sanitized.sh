#!/bin/bash
# The CPU test suite (-m "not gpu") against AddressSanitizer + UndefinedBehaviorSanitizer
# builds of the oracle, the host code of libwavetool.so (C ABI, argument checks, host
# projector) and the host water stage (PNG loader, uniforms). No GPU is used.
# Run from anywhere:  bash tests/run_sanitized.sh [extra pytest args]
set -euo pipefail
ROOT="$(cd "$(dirname "$0")/.." && pwd)"
make -s -C "$ROOT/wave-tool_amd/csrc" asan
make -s -C "$ROOT/oracle" asan
RT="$(/opt/rocm/llvm/bin/clang++ -print-file-name=libclang_rt.asan-x86_64.so)"
# Python and torch are not instrumented: leak reports and the link-order check would be
# about them, not about the code under test
export ASAN_OPTIONS="detect_leaks=0:verify_asan_link_order=0:halt_on_error=1:abort_on_error=1"
export UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1"
export WT_LIBWAVETOOL="$ROOT/wave-tool_amd/csrc/_build_asan/libwavetool.so"
export WT_ORACLE_LIB="$ROOT/oracle/_asan/liboracle.so"
export WT_STAGE_CHECK="$ROOT/wave-tool_amd/csrc/_build_asan/water_stage_check"
cd "$ROOT"
LD_PRELOAD="$RT" python -m pytest tests/ -q -m "not gpu" -p no:cacheprovider "$@"
