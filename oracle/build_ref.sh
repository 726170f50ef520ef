#!/bin/bash
# Builds oracle/_ref/libref.so from the reference's own Karger-Stein sources
# (no Boost, no MPI needed) plus ref_driver.cpp.  Outputs only into oracle/_ref/.
# The rest of the reference (DisjointSets, AdjacencyListGraph, the MPI drivers)
# needs Boost and MPI, which this image lacks, and is not built.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
REF="${CAC_REFERENCE_DIR:-/root/reference}"
KS="$REF/src/karger-stein"
if [ ! -f "$KS/co_mincut.cpp" ]; then
  echo "build_ref.sh: reference tree not found at $REF; skipping" >&2
  exit 0
fi
mkdir -p "$HERE/_ref"
g++ -O2 -std=c++11 -DNDEBUG -mavx2 -fPIC -shared -w \
  -I"$REF/src" -I"$KS" -I"$REF/lib" \
  -o "$HERE/_ref/libref.so.tmp" \
  "$HERE/ref_driver.cpp" "$KS/co_mincut.cpp" "$KS/stack_allocator.cpp" "$KS/bulk_union_find.cpp"
mv "$HERE/_ref/libref.so.tmp" "$HERE/_ref/libref.so"
echo "built $HERE/_ref/libref.so"
