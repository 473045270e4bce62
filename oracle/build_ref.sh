#!/usr/bin/env bash
# Compiles the reference's hot-path headers (FEM/adaptive_fem_engine.hpp,
# FEM/constitutive_models.hpp) and CORE/lie_operator.cpp, from where they lie
# under /root/reference, together with oracle/ref_driver.cpp into
# oracle/_ref/libref.so.  Used only to write golden vectors in the container
# that holds the reference; nothing here is committed or shipped.
#
# The reference includes its headers as "fem/..." and "core/..." while the
# directories are FEM/ and CORE/; the include directory below holds two
# symlinks to those real directories (no header is written or copied).
set -euo pipefail
REF=${REF:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -d "$REF/FEM" ]; then
  echo "build_ref: $REF not present; skipping" >&2
  exit 0
fi
mkdir -p "$OUT/inc"
ln -sfn "$REF/FEM" "$OUT/inc/fem"
ln -sfn "$REF/CORE" "$OUT/inc/core"
g++ -std=c++20 -O3 -march=x86-64-v3 -fopenmp -fPIC -shared -w \
    -I"$OUT/inc" "$HERE/ref_driver.cpp" "$REF/CORE/lie_operator.cpp" -o "$OUT/libref.so"
echo "build_ref: $OUT/libref.so"
