#!/bin/bash
# Builds the C++ drop-in caller against include/hph_engine.hpp and hawkes_amd/libhph.so with the
# host compiler alone: tests/cpp/dropin by default (run path relative to the binary, so the tree
# can move), or the path given as the first argument (run path to this tree's library).
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
if [ -n "$1" ]; then
    OUT=$1
    RPATH="$ROOT/hawkes_amd"
else
    OUT=$HERE/dropin
    RPATH='$ORIGIN/../../hawkes_amd'
fi
g++ -O2 -std=c++17 -I"$ROOT/include" "$HERE/dropin_main.cpp" -L"$ROOT/hawkes_amd" -lhph \
    -Wl,-rpath,"$RPATH" -o "$OUT"
