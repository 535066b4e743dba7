#!/bin/sh
# TEST INFRASTRUCTURE: compiles the reference's Geometry translation units that
# build with this image's own headers, from where they lie under
# /root/reference, plus oracle/ref_harness.cpp, into oracle/_ref/. Output only
# goes to oracle/_ref/ (git-ignored). Exits 0 without building when the
# reference is absent (e.g. on a GPU host).
#
# OpenMP is on, as in the reference's own build: its pragmas are live (call
# ref_set_threads(1) for its serial list order).
# The reference relies on MSVC's transitive includes; the flags force-include
# the standard <cfloat>, <cmath> and <cstring> headers that it uses without
# including. No stand-in headers or libraries are supplied.
set -e
REF=${SOLIDCLOTH_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -f "$REF/Geometry/Collision3D.cpp" ]; then
    echo "build_ref: reference not found at $REF; skipping"
    exit 0
fi
mkdir -p "$OUT"
FLAGS="-std=c++17 -O2 -fPIC -ffp-contract=off -fopenmp -DLINUX -DDISABLE_PROFILER -include cfloat -include cmath -include cstring -I$REF"
g++ $FLAGS -shared -o "$OUT/libref_geometry.so" \
    "$REF/Geometry/Collision3D.cpp" "$REF/Geometry/Mesh.cpp" "$REF/Geometry/AabbTree.cpp" \
    "$REF/Geometry/ClosestPointOnMesh.cpp" "$REF/Geometry/VoronoiRegions.cpp" \
    "$REF/Geometry/ClosestPointsBetweenMeshes.cpp" "$HERE/ref_harness.cpp"
echo "build_ref: built $OUT/libref_geometry.so"
