#!/bin/sh
# Builds the CPU oracle (test infrastructure) into oracle/_build/liboracle.so.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
mkdir -p "$HERE/_build"
gcc -std=gnu11 -O3 -ffp-contract=off -fPIC -shared -Wall -Wextra -Wno-unused-parameter \
    "$HERE/oracle.c" -o "$HERE/_build/liboracle.so.tmp" -lm
mv "$HERE/_build/liboracle.so.tmp" "$HERE/_build/liboracle.so"
