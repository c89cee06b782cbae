#!/bin/sh
# Regenerates tests/golden/philox_rocrand.txt from rocRAND's host engine.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$(mktemp -d)
/opt/rocm/bin/hipcc -O1 -x hip --offload-arch=gfx950 -o "$OUT/gen" "$HERE/gen_philox_rocrand.cpp"
"$OUT/gen" "$HERE/philox_rocrand.txt"
rm -rf "$OUT"
