#!/bin/bash
# Builds the reference checker into oracle/_ref/ from the qTorch sources where
# they lie (header-only C++11; only g++ is needed).  Nothing from the reference
# is copied: ref_driver.cpp includes the headers from REFERENCE_SRC.
#   ref_driver     : the reference's own flags (Makefile:17-19 -> -g -std=c++11 -pthread, no -O)
#   ref_driver_O2  : same sources at -O2, a second, faster CPU data point
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
REFERENCE_SRC="${REFERENCE_SRC:-/root/reference/src}"
if [ ! -f "$REFERENCE_SRC/ContractionTools.h" ]; then
    echo "reference sources not found at $REFERENCE_SRC; oracle/_ref not rebuilt" >&2
    exit 0
fi
mkdir -p "$HERE/_ref"
g++ -g -std=c++11 -pthread -I"$REFERENCE_SRC" "$HERE/ref_driver.cpp" -o "$HERE/_ref/ref_driver"
g++ -O2 -g -std=c++11 -pthread -I"$REFERENCE_SRC" "$HERE/ref_driver.cpp" -o "$HERE/_ref/ref_driver_O2"
echo "built $HERE/_ref/ref_driver and ref_driver_O2"
