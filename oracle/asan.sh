#!/bin/bash
# Builds the host library (compression_amd/host/*.c) and the CPU oracle with
# AddressSanitizer and UndefinedBehaviorSanitizer into oracle/_asan/ (gcc,
# CPU code only), then runs the CPU test suite against those builds.  Extra
# arguments go to pytest, e.g.  oracle/asan.sh -k lzss
set -e
cd "$(dirname "$0")/.."
OUT=oracle/_asan
mkdir -p "$OUT"
FL="-O1 -g -fno-omit-frame-pointer -fsanitize=address,undefined -fno-sanitize-recover=undefined \
    -fPIC -std=gnu11 -pthread -Iinclude"
gcc $FL -shared -o "$OUT/libmdz_host.so" compression_amd/host/*.c -lz -lm
gcc $FL -shared -o "$OUT/libmdz_oracle.so" oracle/rlev1.c oracle/brle.c oracle/rlev2.c \
    oracle/lzss.c oracle/inflate.c -L"$OUT" -lmdz_host -Wl,-rpath,'$ORIGIN' -lz -lm
PRE="$(gcc -print-file-name=libasan.so) $(gcc -print-file-name=libubsan.so)"
MDZ_HOST_LIB="$PWD/$OUT/libmdz_host.so" MDZ_ORACLE_LIB="$PWD/$OUT/libmdz_oracle.so" \
    LD_PRELOAD="$PRE" ASAN_OPTIONS=detect_leaks=0:abort_on_error=1 \
    UBSAN_OPTIONS=print_stacktrace=1 \
    python -m pytest tests -q -m "not gpu" -p no:cacheprovider "$@"
