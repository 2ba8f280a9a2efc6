#!/bin/bash
# Builds the reference's own HIP Day-Phase kernels from their sources where
# they lie (genesis-compute/src/amd/{physics,bindings}.hip of the reference
# checkout at /root/reference) into oracle/_ref/libgenesis_amd_ref.so, for
# gfx950. Test infrastructure only: the GPU cross-check in
# tests/test_gpu_reference.py runs it next to the CPU oracle. Nothing is
# copied into this repository; without /root/reference this does nothing.
set -euo pipefail
SRC=/root/reference/genesis-compute/src/amd
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/_ref"
if [ ! -f "$SRC/physics.hip" ] || [ ! -f "$SRC/bindings.hip" ]; then
  echo "reference sources not present; skipping oracle/_ref" >&2
  exit 0
fi
mkdir -p "$OUT"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
# -fgpu-rdc: bindings.hip refers to physics.hip's __constant__ table.
"$HIPCC" --offload-arch=gfx950 -O2 -fgpu-rdc -fPIC -shared -w \
  -Xclang -target-feature -Xclang -packed-fp32-ops \
  -o "$OUT/libgenesis_amd_ref.so.tmp" "$SRC/physics.hip" "$SRC/bindings.hip"
mv "$OUT/libgenesis_amd_ref.so.tmp" "$OUT/libgenesis_amd_ref.so"
echo "built $OUT/libgenesis_amd_ref.so"
