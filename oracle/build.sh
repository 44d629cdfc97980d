#!/usr/bin/env bash
# Builds the CPU oracle (test infrastructure) into oracle/liboracle.so.
# -ffp-contract=off keeps the C operation order (no fused multiply-adds), so
# the float oracle rounds where the reference source says it rounds.
set -euo pipefail
here="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
g++ -O2 -std=c++17 -fopenmp -ffp-contract=off -fPIC -shared \
    "$here/oracle.cpp" -o "$here/liboracle.so.tmp"
mv "$here/liboracle.so.tmp" "$here/liboracle.so"
