#!/bin/bash
# Compile the Fortran ISO_C_BINDING module (self-contained) with amdflang;
# the object and .mod land next to the source.  The drop-in wrappers need the
# host program's own modules and are compiled with it (see oracle/Makefile,
# target `dropin`, for the test build).
set -e
HERE="$(cd "$(dirname "$0")" && pwd)"
FC="${FC:-/opt/rocm/bin/amdflang}"
cd "$HERE"
"$FC" -O2 -c cap_hip.f90 -o cap_hip.o
