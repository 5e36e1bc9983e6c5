#!/usr/bin/env bash
# Builds kubomsd_amd/lib/libkmsd.so for gfx950 (hipcc cross-compiles without a GPU).
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
ROOT="$(cd "$HERE/../.." && pwd)"
OUT="$ROOT/kubomsd_amd/lib"
mkdir -p "$OUT" "$HERE/_build"
FLAGS=(--offload-arch=gfx950 -O3 -std=c++17 -fPIC -I"$ROOT/include"
       -Xclang -target-feature -Xclang -packed-fp32-ops -Wno-unused-command-line-argument)
SRCS=(kmsd_launch kmsd_api kmsd_metric kmsd_precond kmsd_pcg kmsd_engine kmsd_mu kmsd_mu_mfma kmsd_stats kmsd_noalloc)
pids=()
for s in "${SRCS[@]}"; do
  if [ -f "$HERE/$s.hip" ]; then
    hipcc "${FLAGS[@]}" -c "$HERE/$s.hip" -o "$HERE/_build/$s.o" & pids+=($!)
  fi
done
for p in "${pids[@]}"; do wait "$p"; done
objs=()
for s in "${SRCS[@]}"; do [ -f "$HERE/$s.hip" ] && objs+=("$HERE/_build/$s.o"); done
hipcc --offload-arch=gfx950 -shared -fPIC -o "$OUT/libkmsd.so.tmp" "${objs[@]}"
mv "$OUT/libkmsd.so.tmp" "$OUT/libkmsd.so"
echo "built $OUT/libkmsd.so"
