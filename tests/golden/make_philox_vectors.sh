#!/bin/sh
# Regenerates tests/golden/philox_rocrand.jsonl from rocRAND's Philox4x32-10 engine.
set -e
here=$(cd "$(dirname "$0")" && pwd)
out=$(mktemp -d)
hipcc -O1 -o "$out/philox_vec" "$here/philox_rocrand_vectors.cpp"
"$out/philox_vec" > "$here/philox_rocrand.jsonl"
rm -rf "$out"
