#!/bin/bash
# The CPU test suite under AddressSanitizer + UndefinedBehaviorSanitizer, in
# two runs (each process can hold one sanitizer runtime):
#   1. the oracle's sanitized build (gcc's runtime), loaded via VJ_ORACLE_LIB;
#   2. the library's host code (clang's runtime), loaded via VJ_LIB.
# Usage: bash tests/sanitize.sh [extra pytest arguments]
set -euo pipefail
cd "$(dirname "$0")/.."
make -s -C oracle all asan
make -s -C cs677s26_amd/csrc sanitize
export ASAN_OPTIONS=detect_leaks=0:halt_on_error=1:abort_on_error=1
export UBSAN_OPTIONS=halt_on_error=1:print_stacktrace=1
gcc_asan=$(gcc -print-file-name=libasan.so)
clang_asan=$(/opt/rocm/llvm/bin/clang -print-file-name=libclang_rt.asan-x86_64.so)
if [ ! -f "$clang_asan" ]; then
    clang_asan=$(ls /opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so | head -1)
fi
echo "== oracle under ASan + UBSan (gcc runtime)"
LD_PRELOAD="$gcc_asan" VJ_ORACLE_LIB="$PWD/oracle/build/libvjoracle_asan.so" \
    python -m pytest tests -m "not gpu" -q -p no:cacheprovider "$@"
echo "== library host code under ASan + UBSan (clang runtime)"
LD_PRELOAD="$clang_asan" VJ_LIB="$PWD/cs677s26_amd/_build_san/libvj_hip.so" \
    python -m pytest tests -m "not gpu" -q -p no:cacheprovider "$@"
