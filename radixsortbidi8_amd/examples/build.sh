#!/bin/bash
# Builds the C++ drop-in example (known-answer tests through include/rsbd8_hip.hpp) against the
# in-tree HIP library. Host-only C++: plain g++ with the HIP runtime headers.
set -e
HERE="$(cd "$(dirname "$0")" && pwd)"
LIBDIR="$HERE/../lib"
mkdir -p "$LIBDIR"
g++ -std=c++17 -O2 -Wall -D__HIP_PLATFORM_AMD__ -I/opt/rocm/include -I"$HERE/../../include" \
	-o "$LIBDIR/kat_main" "$HERE/kat_main.cpp" \
	-L"$LIBDIR" -lrsbd8_hip -L/opt/rocm/lib -lamdhip64 -Wl,-rpath,'$ORIGIN' -Wl,-rpath,/opt/rocm/lib
