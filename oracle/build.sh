#!/usr/bin/env bash
# Builds the CPU oracle (test infrastructure only) into oracle/_build/liboracle.so.
# No FMA contraction and no fast-math: the restatement keeps the reference's IEEE order.
# OpenMP splits independent columns and rows over threads; it does not change any result.
set -euo pipefail
here="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
mkdir -p "$here/_build"
g++ -std=c++17 -O2 -ffp-contract=off -fno-fast-math -fopenmp -fPIC -shared -Wall -Wno-sign-compare \
    "$here/oracle.cpp" -o "$here/_build/liboracle.so.tmp"
mv "$here/_build/liboracle.so.tmp" "$here/_build/liboracle.so"
