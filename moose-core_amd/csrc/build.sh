#!/bin/bash
# Builds libhsolve_amd.so for gfx950 in-tree (next to this script's package).
#   -ffp-contract=off : the FP64 form reproduces the reference bit for bit
#   -packed-fp32-ops off : no v_pk_* FP32 instructions
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
out="$here/../libhsolve_amd.so"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
ARCH="${HSG_ARCH:-gfx950}"
"$HIPCC" --offload-arch="$ARCH" -O3 -std=c++17 -fPIC -shared \
    -ffp-contract=off -Xclang -target-feature -Xclang -packed-fp32-ops \
    -Wall -Wno-unused-function -Wno-unused-value -Wno-unused-result \
    -o "$out.tmp" \
    "$here/hsolve_kernels.hip" "$here/hsolve_setup.cpp" "$here/hsolve_engine.cpp"
mv "$out.tmp" "$out"
echo "built $out"
