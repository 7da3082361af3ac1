#!/bin/bash
# Builds the oracle (CPU restatement; test infrastructure only) into
# oracle/liboracle_umappp.so.  No -march and no -ffast-math: the restatement
# must round exactly as the reference does on a baseline x86-64 build.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
g++ -O2 -std=c++17 -fPIC -shared -pthread -ffp-contract=off -Wall -Wextra \
    -o "$here/liboracle_umappp.so.tmp" "$here/umappp_oracle.cpp"
mv "$here/liboracle_umappp.so.tmp" "$here/liboracle_umappp.so"
