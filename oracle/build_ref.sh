#!/bin/bash
# Compiles oracle/ref_driver.cpp against the reference's own host headers, in
# place under the reference tree (utils/matrix.h, utils/cpu_spmm_fuse.h), into
# oracle/_ref/ref_driver.  Nothing from the reference is copied.  Skips quietly
# when the reference tree is absent (the GPU box has only the prebuilt binary).
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
REF="${COSPDNN_REFERENCE:-/root/reference}"
if [ ! -f "$REF/src/utils/matrix.h" ]; then
  echo "build_ref.sh: reference sources not found at $REF; oracle/_ref not rebuilt"
  exit 0
fi
mkdir -p "$HERE/_ref"
g++ -O2 -std=c++11 -w -I"$REF/src" -o "$HERE/_ref/ref_driver" "$HERE/ref_driver.cpp"
