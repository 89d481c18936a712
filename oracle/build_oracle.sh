#!/bin/sh
# Builds the CPU oracle (test infrastructure) into oracle/_build/liboracle.so.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
mkdir -p "$HERE/_build"
gcc -std=c11 -O2 -fPIC -shared -Wall -Wextra -Wno-unused-parameter \
    "$HERE/oracle.c" -o "$HERE/_build/liboracle.so.tmp" -lm
mv "$HERE/_build/liboracle.so.tmp" "$HERE/_build/liboracle.so"
