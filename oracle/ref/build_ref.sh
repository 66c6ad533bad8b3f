#!/bin/bash
# Compiles the reference's standalone sources in place from /root/reference
# (never copied) plus ref_driver.cc into oracle/_ref/libbwref.so.
# Test infrastructure; the output is git-ignored.
set -euo pipefail
R=/root/reference
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/../_ref"
mkdir -p "$OUT"
INC="-I$R/il -I$R/src -I$R/external/nlohmann -I$R/external"
g++ -std=c++17 -O2 -fPIC -shared -fopenmp $INC \
  "$HERE/ref_driver.cc" \
  "$R/src/hmat/cluster/cluster.cc" \
  "$R/src/hmat/hmatrix/h_pattern.cc" \
  "$R/src/elasticity/fullspace_iso_2d_segment/elastic_2dP0_segment.cc" \
  -o "$OUT/libbwref.so.tmp"
mv "$OUT/libbwref.so.tmp" "$OUT/libbwref.so"
