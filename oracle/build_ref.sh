#!/bin/sh
# Builds the reference harness (oracle/ref_harness.cpp) against the reference's
# own header-only sources where they lie. Output goes to oracle/_ref/ only.
# Requires the reference tree (default /root/reference); exits 0 without
# building when it is absent (e.g. on a machine that only has this repository).
set -e
REF=${REFERENCE_ROOT:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
if [ ! -d "$REF/src" ]; then
  echo "build_ref.sh: reference sources not found at $REF; skipping" >&2
  exit 0
fi
mkdir -p "$HERE/_ref"
g++ -std=c++14 -O2 -w -I "$REF/src" -isystem "$REF/src/External" \
    "$HERE/ref_harness.cpp" -o "$HERE/_ref/ref_harness" -lpthread
echo "built $HERE/_ref/ref_harness"
# the reference's command-line parser (bunnykiller.cpp, with its vendored TinyXML for -parse-xml-file)
g++ -std=c++14 -O1 -w -I "$REF/src" -isystem "$REF/src/External" \
    "$HERE/ref_cli_probe.cpp" "$REF/src/External/TinyXML/tinyxml2.cpp" -o "$HERE/_ref/ref_cli_probe" -lpthread
echo "built $HERE/_ref/ref_cli_probe"
