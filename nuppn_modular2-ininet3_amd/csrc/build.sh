#!/bin/bash
# Builds libnuppn_hip.so (gfx950) in-tree: nuppn_modular2-ininet3_amd/lib/libnuppn_hip.so
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
OUT="$HERE/../lib"
mkdir -p "$OUT" "$HERE/_build"
HIPCC=${HIPCC:-/opt/rocm/bin/hipcc}
FLAGS="-O3 -std=c++17 -fPIC --offload-arch=gfx950 -ffp-contract=off -Wall -Wno-unused-function \
 -Xclang -target-feature -Xclang -packed-fp32-ops -munsafe-fp-atomics"
"$HIPCC" $FLAGS -c "$HERE/symbolic.cpp" -o "$HERE/_build/symbolic.o"
"$HIPCC" $FLAGS -c "$HERE/colamd.cpp" -o "$HERE/_build/colamd.o"
"$HIPCC" $FLAGS -c "$HERE/nuppn_hip.hip" -o "$HERE/_build/nuppn_hip.o"
"$HIPCC" -shared -fPIC --offload-arch=gfx950 -o "$OUT/libnuppn_hip.so.tmp" \
  "$HERE/_build/nuppn_hip.o" "$HERE/_build/symbolic.o" "$HERE/_build/colamd.o"
mv -f "$OUT/libnuppn_hip.so.tmp" "$OUT/libnuppn_hip.so"
echo "built $OUT/libnuppn_hip.so"
# The Fortran side of the boundary: the nuppn_hip bind(c) module and the ppn-style driver that
# calls the library through it (flang; the module file goes to _build)
FC=${FC:-/opt/rocm/llvm/bin/flang}
if [ -x "$FC" ]; then
  FDIR="$HERE/../fortran"
  "$FC" -O2 -J "$HERE/_build" -c "$FDIR/nuppn_hip.F90" -o "$HERE/_build/nuppn_hip_f.o"
  "$FC" -O2 -J "$HERE/_build" "$FDIR/ppn_driver.F90" "$HERE/_build/nuppn_hip_f.o" \
    -o "$OUT/nuppn_ppn_driver" -L"$OUT" -lnuppn_hip -Wl,-rpath,'$ORIGIN'
  echo "built $OUT/nuppn_ppn_driver"
else
  echo "no Fortran compiler at $FC: the nuppn_hip module and ppn driver are not built" >&2
fi
