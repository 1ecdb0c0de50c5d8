#!/bin/sh
# Regenerates rng_golden.json from the pcg-cpp header vendored with pyarrow and
# the system libstdc++ (see make_rng_golden.cpp).
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
PCG=/usr/local/lib/python3.10/dist-packages/pyarrow/include/arrow/vendored/pcg
OUT=$(mktemp -d)
g++ -std=c++17 -O1 -ffp-contract=off -I"$PCG" "$HERE/make_rng_golden.cpp" -o "$OUT/make_rng_golden"
"$OUT/make_rng_golden" > "$HERE/rng_golden.json"
rm -rf "$OUT"
