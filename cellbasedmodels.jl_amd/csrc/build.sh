#!/bin/sh
# Builds libcbm_hip.so for gfx950 in-tree (next to the Python package).
set -e
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/../cbm/lib"
mkdir -p "$OUT"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
# the stage- and rule-kernel headers, embedded for hipRTC compilation of user bodies
INC="$OUT/cbm_rtc_src.inc"
{
  for h in cbm_device.h cbm_structs.h cbm_stage.h cbm_rule.h cbm_medium.h cbm_model.h; do
    v=$(echo "$h" | tr '.' '_')
    printf 'static const char* cbm_rtc_%s = R"CBMRTC(' "$v"
    cat "$HERE/$h"
    printf ')CBMRTC";\n'
  done
} > "$INC.tmp"
mv "$INC.tmp" "$INC"
FLAGS="--offload-arch=gfx950 -O3 -std=c++17 -fPIC -ffp-contract=off -fno-fast-math -Xclang -target-feature -Xclang -packed-fp32-ops -Wall -Wno-unused-function -Wno-unused-variable -I$OUT"
$HIPCC $FLAGS -c "$HERE/cbm_kernels.hip" -o "$OUT/cbm_kernels.o"
$HIPCC $FLAGS -c "$HERE/cbm_api.hip" -o "$OUT/cbm_api.o"
$HIPCC --offload-arch=gfx950 -shared -fPIC -o "$OUT/libcbm_hip.so.tmp" "$OUT/cbm_kernels.o" "$OUT/cbm_api.o" -lhiprtc
mv "$OUT/libcbm_hip.so.tmp" "$OUT/libcbm_hip.so"
rm -f "$OUT/cbm_kernels.o" "$OUT/cbm_api.o"
echo "built $OUT/libcbm_hip.so"
