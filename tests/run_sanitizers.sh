#!/bin/bash
# CPU only.  Runs the oracle's tests, the host greedy merge of libdtc
# (dtc_tca_greedy) and the C ABI's argument checks with the oracle and the
# host code of libdtc built under AddressSanitizer and
# UndefinedBehaviorSanitizer.  Both builds use ROCm's clang, whose ASan runtime
# (preloaded here) also carries the UBSan handlers.  The sanitized libdtc is
# preloaded as well, so the DTCSpMM extension, which links libdtc.so by
# soname, binds to it; ctypes callers find it through DTC_LIBRARY.
set -euo pipefail
cd "$(dirname "$0")/.."
make -s -C oracle sanitize
python dtc-spmm_asplos24_amd/build.py --sanitize
ASAN_RT="$(/opt/rocm/llvm/bin/clang -print-resource-dir)/lib/linux/libclang_rt.asan-x86_64.so"
SAN_LIB="$PWD/dtc-spmm_asplos24_amd/_build_san/libdtc.so"
export DTC_ORACLE_LIBRARY="$PWD/oracle/_build/san/libmetcf_oracle.so"
export DTC_LIBRARY="$SAN_LIB"
# leaks: the interpreter and torch keep allocations until exit by design
export ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:verify_asan_link_order=0"
export UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1"
LD_PRELOAD="$ASAN_RT:$SAN_LIB" python -m pytest -q -p no:cacheprovider -m "not gpu" \
  tests/test_oracle.py tests/test_reorder.py tests/test_capi.py "$@"
