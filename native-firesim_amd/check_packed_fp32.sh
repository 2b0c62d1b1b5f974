#!/bin/bash
# Fails if a gfx950 code object built here holds packed-FP32 VALU instructions
# (v_pk_add/mul/fma_f32). The HIP objects are compiled with the packed-fp32-ops
# target feature switched off; this confirms it on the disassembly of each
# object's own device code (hipcc's host pass warns that it does not know the
# feature, which is harmless: the device pass honours it).
set -euo pipefail
LLVM=${LLVM:-/opt/rocm/llvm/bin}
tmp=$(mktemp -d)
trap 'rm -rf "$tmp"' EXIT
status=0
for obj in "$@"; do
    base=$(basename "$obj")
    "$LLVM/llvm-objcopy" -O binary --only-section=.hip_fatbin "$obj" "$tmp/$base.fatbin"
    "$LLVM/clang-offload-bundler" --type=o --targets=hipv4-amdgcn-amd-amdhsa--gfx950 --input="$tmp/$base.fatbin" \
        --output="$tmp/$base.co" --unbundle
    n=$("$LLVM/llvm-objdump" -d --mcpu=gfx950 "$tmp/$base.co" | grep -c -E 'v_pk_(add|mul|fma)_f32' || true)
    if [ "$n" != "0" ]; then
        echo "check_packed_fp32: $obj holds $n packed-FP32 instructions" >&2
        status=1
    fi
done
exit $status
