#!/usr/bin/env bash
# Builds oracle/_ref/pado_ref from the reference's own headers, in place under
# /root/reference/includes, plus oracle/ref_harness.cpp.
#
# TEST INFRASTRUCTURE ONLY: the binary is the CPU baseline and a golden-vector
# generator; nothing in the product links or runs it.
#
# The reference was written for icpc.  g++ and clang reject it at six
# `IndexType::Batch(` / `IndexType::DistanceIndexType(` sites in the unweighted
# header (a dependent type name used without `typename`).  The only change made
# is to insert that keyword, on a temporary copy outside the repository that is
# deleted after compiling; no reference source is kept, and nothing else in the
# headers is touched.
set -euo pipefail
REF=${PADO_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -d "$REF/includes" ]; then
    echo "build_ref.sh: $REF/includes not present; skipping the reference build" >&2
    exit 0
fi
mkdir -p "$OUT"
TMP=$(mktemp -d /tmp/pado_ref_build.XXXXXX)
trap 'rm -rf "$TMP"' EXIT
for h in globals.h graph.h pado_unw_unv_para.201912311607.parallel_bp_labeling.h pado_weighted_unv_para.h \
         pado_unw_unv_unp.h; do
    sed -e 's/push_back(IndexType::/push_back(typename IndexType::/g' "$REF/includes/$h" > "$TMP/$h"
done
CXX=${CXX:-g++}
$CXX -std=c++14 -O3 -march=x86-64-v3 -fopenmp -fpermissive -w -DPADO_REF_WEIGHTED \
    -I"$TMP" "$HERE/ref_harness.cpp" -o "$OUT/pado_ref.tmp"
mv "$OUT/pado_ref.tmp" "$OUT/pado_ref"
echo "built $OUT/pado_ref"
# The sequential class, whose index-file writer works (UNP:1120-1180).
$CXX -std=c++14 -O2 -march=x86-64-v3 -fopenmp -fpermissive -w -I"$TMP" "$HERE/ref_unp.cpp" -o "$OUT/pado_ref_unp.tmp"
mv "$OUT/pado_ref_unp.tmp" "$OUT/pado_ref_unp"
echo "built $OUT/pado_ref_unp"
