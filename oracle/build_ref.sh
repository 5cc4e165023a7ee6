#!/bin/bash
# Builds the reference checker into oracle/_ref/ from the qTorch sources where
# they lie (header-only C++11; only g++ is needed).  Nothing from the reference
# is copied: ref_driver.cpp includes the headers from REFERENCE_SRC.
#   ref_driver     : the reference's own flags (Makefile:17-19 -> -g -std=c++11 -pthread, no -O)
#   ref_driver_O2  : same sources at -O2, a second, faster CPU data point (the one
#                    the benchmark's CPU baseline runs; ref_driver stays here, it
#                    only checks at fixture time that -O0 and -O2 print the same bits)
# maxcut.h (the QAOA light cones) includes nlopt.hpp: the vendored nlopt API headers
# go on the include path; nothing of nlopt is compiled or linked.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
REFERENCE_SRC="${REFERENCE_SRC:-/root/reference/src}"
if [ ! -f "$REFERENCE_SRC/ContractionTools.h" ]; then
    echo "reference sources not found at $REFERENCE_SRC; oracle/_ref not rebuilt" >&2
    exit 0
fi
NLOPT_API="${NLOPT_API:-$REFERENCE_SRC/../nlopt-2.4.2/api}"
mkdir -p "$HERE/_ref"
g++ -g -std=c++11 -pthread -I"$REFERENCE_SRC" -I"$NLOPT_API" "$HERE/ref_driver.cpp" -o "$HERE/_ref/ref_driver"
g++ -O2 -g -std=c++11 -pthread -I"$REFERENCE_SRC" -I"$NLOPT_API" "$HERE/ref_driver.cpp" -o "$HERE/_ref/ref_driver_O2"
echo "built $HERE/_ref/ref_driver and ref_driver_O2"
