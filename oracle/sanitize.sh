#!/bin/bash
# Host-side sanitizer run (CPU only): the oracle and the C++ host classes
# (libdfsph_host.so: scene files, fluid blocks, OBJ sampling, text particle
# files) built with AddressSanitizer + UndefinedBehaviorSanitizer by gcc/g++,
# then the CPU tests that drive them.  Outputs go to oracle/_asan/.
set -euo pipefail
ROOT=$(cd "$(dirname "$0")/.." && pwd)
OUT="$ROOT/oracle/_asan"
PKG="$ROOT/splishsplash_for_cuda_amd"
SAN="-fsanitize=address,undefined -fno-omit-frame-pointer -fno-sanitize-recover=undefined"
mkdir -p "$OUT"
gcc -O1 -g $SAN -fopenmp -fPIC -ffp-contract=off -std=c11 -Wall -Wno-unknown-pragmas -shared \
    -o "$OUT/liboracle.so" "$ROOT/oracle/dfsph_oracle.c" -lm
g++ -O1 -g $SAN -std=c++17 -fPIC -shared -Wall -I"$ROOT/include" -I"$PKG/csrc" \
    "$PKG/csrc/DFSPH_CUDA.cpp" "$PKG/csrc/SceneLoader.cpp" "$PKG/csrc/DynamicWindow.cpp" \
    -L"$PKG" -ldfsph_hip -Wl,-rpath,"$PKG" -o "$OUT/libdfsph_host.so"
export LD_PRELOAD="$(gcc -print-file-name=libasan.so) $(gcc -print-file-name=libubsan.so)"
export ASAN_OPTIONS=detect_leaks=0:halt_on_error=1
export UBSAN_OPTIONS=halt_on_error=1:print_stacktrace=1
export ORACLE_LIB="$OUT/liboracle.so" DFSPH_HOST_LIB="$OUT/libdfsph_host.so" OMP_NUM_THREADS=${OMP_NUM_THREADS:-4}
cd "$ROOT"
python -m pytest tests/test_oracle.py tests/test_oracle_bodies.py tests/test_scene.py tests/test_abi.py -q \
    -m "not gpu" -p no:cacheprovider "$@"
