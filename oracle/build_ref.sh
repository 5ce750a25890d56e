#!/usr/bin/env bash
# Compiles the reference's own FP64 physics functions, from their sources in
# /root/reference/src/physics (kerr.cpp + header-only formulas), together with
# oracle/ref/ref_wrap.cpp and oracle/ref/ref_wrap_disk.cpp into
# oracle/_ref/libref.so, and the reference's
# render-ABI header into oracle/_ref/abi_layout (prints BH_LaunchParams
# offsets).  Test infrastructure only; outputs stay in oracle/_ref/.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
REF="${BLACKHOLE_REFERENCE:-/root/reference}"
OUT="$HERE/_ref"
if [ ! -d "$REF/src/physics" ]; then
  echo "build_ref.sh: $REF not present; skipping" >&2
  exit 0
fi
mkdir -p "$OUT"
CXX_REF="${CXX_REF:-/opt/rocm/llvm/bin/clang++}"
"$CXX_REF" -std=c++23 -O2 -ffp-contract=off -fPIC -shared -w \
  -I "$REF/src" "$HERE/ref/ref_wrap.cpp" "$HERE/ref/ref_wrap_disk.cpp" "$REF/src/physics/kerr.cpp" \
  -o "$OUT/libref.so"
cat > "$OUT/abi_layout.c" <<'CEOF'
#include <stdio.h>
#include <stddef.h>
#include "cuda/kernel_launch.h"
#define F(f) printf("\"%s\": %zu,\n", #f, offsetof(struct BH_LaunchParams, f));
int main(void) {
  printf("{\n");
#include "abi_fields.inc"
  printf("\"__sizeof__\": %zu,\n", sizeof(struct BH_LaunchParams));
  printf("\"BH_KERNEL_FP32_BASELINE\": %d, \"BH_KERNEL_FP32_COARSENED\": %d, "
         "\"BH_KERNEL_FP16_STORAGE\": %d, \"BH_KERNEL_FP16_H2_ILP\": %d, \"BH_KERNEL_COUNT\": %d\n}\n",
         BH_KERNEL_FP32_BASELINE, BH_KERNEL_FP32_COARSENED, BH_KERNEL_FP16_STORAGE,
         BH_KERNEL_FP16_H2_ILP, BH_KERNEL_COUNT);
  return 0;
}
CEOF
python3 - "$OUT/abi_fields.inc" <<'PYEOF'
import sys, re
sys.path.insert(0, __import__("os").path.dirname(__import__("os").path.dirname(sys.argv[1])) + "/..")
from blackhole_amd.abi import BH_LaunchParams
with open(sys.argv[1], "w") as f:
    for name, *_ in BH_LaunchParams._fields_:
        f.write(f"F({name})\n")
PYEOF
gcc -std=c99 -I "$REF/src" -I "$OUT" "$OUT/abi_layout.c" -o "$OUT/abi_layout"
echo "build_ref.sh: built $OUT/libref.so and $OUT/abi_layout"
