#!/bin/sh
# Compiles the reference's leaf loader chain from its own sources where they
# lie under the (read-only) reference tree, plus ref_harness.cpp, into
# oracle/_ref/libg4ref.so.  Only these TUs build without external libraries;
# the hot-path TUs (BeamSolver, TrackBeam, FieldSolverADI, ...) include
# hdf5.h / mpi.h and are not built.  Usage: oracle/build_ref.sh [REFROOT]
set -e
REF=${1:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
[ -d "$REF/src" ] || { echo "reference tree not found at $REF; skipping"; exit 0; }
mkdir -p "$HERE/_ref"
g++ -O2 -std=c++17 -fPIC -shared -ffp-contract=off -w -I"$REF/include" \
    "$REF/src/Util/RandomU.cpp" "$REF/src/Util/Hammerslay.cpp" "$REF/src/Util/Inverfc.cpp" \
    "$REF/src/Util/BesselJ.cpp" "$REF/src/Loading/QuietLoading.cpp" \
    "$REF/src/Loading/ShotNoise.cpp" "$REF/src/Util/GaussHermite.cpp" \
    "$HERE/ref_harness.cpp" -o "$HERE/_ref/libg4ref.so"
echo "built $HERE/_ref/libg4ref.so"
