#!/bin/bash
# Builds oracle/_ref/ref_driver: the harness oracle/ref_driver.cpp linked with
# the six reference source files it needs, compiled directly from
# /root/reference/src (never copied).  Output only under oracle/_ref/.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
REF=/root/reference/src
OUT="$HERE/_ref"
mkdir -p "$OUT"
FLAGS="-O2 -std=c++14 -D_ALIGNBYTES_=16 -D_USE_BACKEND_LOCAL_=1 -D_USE_BACKEND_NATIVE_=1 \
 -D_USE_DIM1_=0 -D_USE_DIM2_=0 -D_USE_DIM3_=1 -D_USE_DIM4_=0 -D_USE_FPZIP_=0 -D_USE_AMGX_=0 \
 -D_USE_AVX_=0 -D_USE_BACKEND_CUBISM_=0 -D_USE_HDF_=0 -D_USE_HYPRE_=0 -D_USE_MPI_=0 \
 -D_USE_OPENCL_=0 -D_USE_OPENMP_=0 -I$REF"
g++ $FLAGS -o "$OUT/ref_driver.tmp" "$HERE/ref_driver.cpp" \
  "$REF/geom/mesh.cpp" "$REF/util/suspender.cpp" "$REF/parse/vars.cpp" \
  "$REF/util/logger.cpp" "$REF/util/format.cpp"
mv "$OUT/ref_driver.tmp" "$OUT/ref_driver"

# oracle/_ref/ref_hip_driver: the same harness with the solver created through
# the reference's module registry, plus the "hip" module of
# aphros_amd/adapter linked with the product library (a drop-in check: the
# reference's own stage loop and plugin interface calling libaphpois).
REPO="$(cd "$HERE/.." && pwd)"
if [ -f "$REPO/aphros_amd/libaphpois.so" ]; then
  g++ $FLAGS -DAPH_HIP_MODULE -I"$REPO/include" -o "$OUT/ref_hip_driver.tmp" \
    "$HERE/ref_driver.cpp" "$REPO/aphros_amd/adapter/linear_hip.cpp" \
    "$REF/linear/linear.cpp" "$REF/geom/mesh.cpp" "$REF/util/suspender.cpp" \
    "$REF/parse/vars.cpp" "$REF/util/logger.cpp" "$REF/util/format.cpp" \
    -L"$REPO/aphros_amd" -laphpois -L/opt/rocm/lib -Wl,-rpath-link,/opt/rocm/lib \
    -Wl,-rpath,'$ORIGIN/../../aphros_amd'
  mv "$OUT/ref_hip_driver.tmp" "$OUT/ref_hip_driver"
fi

# oracle/_ref/ref_blocks_driver: the reference's modules and the "hip" module
# on a rank split into several blocks that share one lead (the local
# backend's exchanges between stages, served by the harness).
if [ -f "$REPO/aphros_amd/libaphpois.so" ]; then
  g++ $FLAGS -I"$REPO/include" -I"$REPO/aphros_amd/adapter" -o "$OUT/ref_blocks_driver.tmp" \
    "$HERE/ref_blocks_driver.cpp" "$REPO/aphros_amd/adapter/linear_hip.cpp" \
    "$REF/linear/linear.cpp" "$REF/geom/mesh.cpp" "$REF/util/suspender.cpp" \
    "$REF/parse/vars.cpp" "$REF/util/logger.cpp" "$REF/util/format.cpp" \
    -L"$REPO/aphros_amd" -laphpois -L/opt/rocm/lib -Wl,-rpath-link,/opt/rocm/lib \
    -Wl,-rpath,'$ORIGIN/../../aphros_amd'
  mv "$OUT/ref_blocks_driver.tmp" "$OUT/ref_blocks_driver"
fi
