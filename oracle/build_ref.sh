#!/bin/sh
# Builds oracle/_ref/tc_ref from the reference's own host sources where they lie
# in the reference checkout (read-only; nothing is copied) plus ref_harness.cpp.
# Test infrastructure only: the binary writes parity fixtures and is never
# part of the product.  Exits 0 without building when the checkout is absent.
set -e
REF=${BCX_REFERENCE_DIR:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
if [ ! -f "$REF/src/TreeCount/TreeDecomp.cpp" ]; then
  echo "build_ref: reference checkout not found at $REF; skipping" >&2
  exit 0
fi
mkdir -p "$HERE/_ref"
# The harness includes the reference's include/Graph.h by a path relative to
# the reference tree, so it is compiled with -iquote pointing there.
g++ -std=c++17 -O2 -w -iquote "$REF/src" \
  -include "$REF/include/Graph.h" \
  -o "$HERE/_ref/tc_ref" \
  -x c++ "$HERE/ref_harness.cpp" \
  "$REF/src/TreeCount/TreeDecomp.cpp" "$REF/src/TreeCount/Graph.cpp" \
  "$REF/src/Graph.cpp" "$REF/src/buildIndex.cpp"
echo "build_ref: built $HERE/_ref/tc_ref"
