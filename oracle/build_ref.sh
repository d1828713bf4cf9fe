#!/bin/bash
# Builds oracle/_ref/ref_golden from the reference sources where they lie
# (/root/reference, read only) and this directory's ref_golden.cpp.
#
# The reference's path code is compiled from its own translation units with
# g++ directly (no CMake): every .cpp under yuzen/src except
# numerics/LinearSolverBackend.cpp (the only unit that needs Eigen 3.4.0 /
# AMGCL 1.4.4, which are not in this image) and mainProg.cpp. Flags follow
# the reference's build: C++17, -ffp-contract=off, _DEF_YUZEN_ (CMakeLists.txt
# :33-45, 179-185). Outputs go to oracle/_ref only.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
REF="${YUZEN_REFERENCE:-/root/reference}"
OUT="$HERE/_ref"
S="$REF/yuzen/src"
if [ ! -d "$S" ]; then
    echo "build_ref.sh: no reference tree at $REF; nothing built" >&2
    exit 0
fi
mkdir -p "$OUT/obj"
INC="-I$REF/third_party -I$S -I$S/IO -I$S/util -I$S/grid -I$S/model -I$S/commonFunctions \
 -I$S/model/well_mixed -I$S/model/well_mixed/WM1D -I$S/model/well_mixed/WM2D -I$S/model/ConcZ/CZ2D \
 -I$S/numerics -I$S/MyMath"
FLAGS="-std=c++17 -O2 -ffp-contract=off -D_DEF_YUZEN_ -fopenmp -w -ffunction-sections -fdata-sections"
SRCS=$(find "$S" -name '*.cpp' ! -name 'LinearSolverBackend.cpp' ! -name 'mainProg.cpp' | sort)
objs=()
for src in $SRCS; do
    obj="$OUT/obj/$(echo "${src#$S/}" | tr '/' '_' | sed 's/\.cpp$/.o/')"
    objs+=("$obj")
    if [ ! -f "$obj" ] || [ "$src" -nt "$obj" ]; then
        echo "$src -> $obj"
    fi
done | awk '{print $1" "$3}' | xargs -r -P 8 -n 2 sh -c 'g++ '"$FLAGS $INC"' -c "$0" -o "$1"'
rm -f "$OUT/libyuzencore_noeigen.a"
ar rcs "$OUT/libyuzencore_noeigen.a" "$OUT"/obj/*.o
g++ $FLAGS $INC "$HERE/ref_golden.cpp" "$OUT/libyuzencore_noeigen.a" -Wl,--gc-sections -o "$OUT/ref_golden"
echo "built $OUT/ref_golden"
