#!/bin/sh
# Build check of the CSimBox drop-in: the reference's src/SimBox.cpp with the three lines of
# CSimBox::Run that simbox_engine.h names replaced, compiled against the reference's headers
# together with simbox_engine.cpp, and linked with the reference's other objects (built by
# oracle/ref.mk) and libospreydpd_amd.so into a `dpd` executable.  Everything is written under
# $OUT (default /tmp/dpd_dropin); nothing of the reference enters the repository.
#
#   make -f oracle/ref.mk -j8 && make -C osprey-dpd_amd/csrc -j8
#   sh osprey-dpd_amd/integration/check_dropin.sh
set -e
REF=${REF:-/root/reference/src}
HERE=$(cd "$(dirname "$0")" && pwd)
REPO=$(cd "$HERE/../.." && pwd)
OUT=${OUT:-/tmp/dpd_dropin}
OBJ=$REPO/oracle/_ref/obj
LIB=$REPO/osprey-dpd_amd/lib
CXX=${CXX:-g++}
mkdir -p "$OUT"

# the three call sites of CSimBox::Run, each matched as the whole line it is
sed -e 's|^#include "SimBox.h"$|&\n#include "simbox_engine.h"|' \
    -e 's|^\([[:space:]]*\)UpdateRenormalisedMom();$|\1DPD_SIMBOX_ENGINE.RenormaliseMomenta();|' \
    -e 's|^\([[:space:]]*\)Evolve();$|\1DPD_SIMBOX_ENGINE.Step(TimeToSample() \|\| TimeToDisplay() \|\| TimeToRestart());|' \
    -e 's|^\([[:space:]]*\)CNTCellCheck();\(.*\)$|\1if (CSimBoxEngine::Synced()) CNTCellCheck();\2|' \
    "$REF/SimBox.cpp" > "$OUT/SimBox.cpp"
test "$(grep -c "DPD_SIMBOX_ENGINE\|CSimBoxEngine::Synced" "$OUT/SimBox.cpp")" -eq 3 || { echo "SimBox.cpp: call sites not found" >&2; exit 1; }

FLAGS="-O3 -std=c++11 -w -I$REF -I$REPO/include -I$HERE"
$CXX $FLAGS -c "$OUT/SimBox.cpp" -o "$OUT/SimBox.o"
$CXX $FLAGS -c "$HERE/simbox_engine.cpp" -o "$OUT/simbox_engine.o"
ls "$OBJ"/*.o | grep -v '/SimBox\.o$' > "$OUT/objs.txt"
$CXX -o "$OUT/dpd" "$OUT/SimBox.o" "$OUT/simbox_engine.o" $(cat "$OUT/objs.txt") \
    -L"$LIB" -lospreydpd_amd -Wl,-rpath,"$LIB" -pthread
# every engine entry point the drop-in calls resolves in the library
nm -D --undefined-only "$OUT/dpd" | grep ' dpd_' || true
echo "built $OUT/dpd"
