#!/bin/bash
# Builds tests/native/_build_asan/batch_asan_driver: every library source with
# its host code under AddressSanitizer (-Xarch_host: the GPU code is built as
# in the product, uninstrumented) plus batch_asan_driver.cpp.  Run from the
# repository root, then on a GPU:
#   ASAN_OPTIONS=detect_leaks=0:verify_asan_link_order=0 \
#     tests/native/_build_asan/batch_asan_driver 16
set -e
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/_build_asan"
SRC="$HERE/../../lambent_amd/csrc"
HIPCC=/opt/rocm/bin/hipcc
COMMON="-O1 -g -std=c++17 -fPIC -ffp-contract=off -fno-omit-frame-pointer -Wno-unused-command-line-argument"
HIPF="-x hip --offload-arch=gfx950 -Xclang -target-feature -Xclang -packed-fp32-ops -munsafe-fp-atomics"
mkdir -p "$OUT"
pids=()
for f in "$SRC"/*.cpp; do
  $HIPCC $COMMON -Xarch_host -fsanitize=address -c "$f" -o "$OUT/$(basename "$f").o" &
  pids+=($!)
done
for f in "$SRC"/*.hip; do
  $HIPCC $COMMON $HIPF -Xarch_host -fsanitize=address -c "$f" -o "$OUT/$(basename "$f").o" &
  pids+=($!)
done
for p in "${pids[@]}"; do wait "$p"; done
$HIPCC $COMMON -Xarch_host -fsanitize=address -c "$HERE/batch_asan_driver.cpp" -o "$OUT/driver.o"
$HIPCC --offload-arch=gfx950 -Xarch_host -fsanitize=address -o "$OUT/batch_asan_driver" "$OUT"/*.o -lpthread
rm -f "$OUT"/*.o
echo "built $OUT/batch_asan_driver"
