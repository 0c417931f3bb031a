#!/bin/bash
# Builds the Fortran binding modules and the replay harness against the in-tree library.
set -e
HERE="$(cd "$(dirname "$0")" && pwd)"
FC="${FC:-/opt/rocm/bin/amdflang}"
mkdir -p "$HERE/_build"
cd "$HERE/_build"
"$FC" -O2 -c "$HERE/dynemol_hip.f90" -o dynemol_hip.o
"$FC" -O2 -c "$HERE/gpu_elhl_chebyshev.f90" -o gpu_elhl_chebyshev.o
"$FC" -O2 -c "$HERE/harness.f90" -o harness.o
"$FC" -o "$HERE/dh_harness" harness.o gpu_elhl_chebyshev.o dynemol_hip.o -L"$HERE/.." -ldynemol_hip \
    -Wl,-rpath,'$ORIGIN/..'
