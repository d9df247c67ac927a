#!/bin/sh
# Builds the FP64 CPU oracle (test infrastructure) into oracle/_build/.
# The reference itself (CUDA C++, needs nvcc and an NVIDIA GPU) is not
# buildable in this image, so there is no oracle/_ref build.
set -e
here=$(cd "$(dirname "$0")" && pwd)
mkdir -p "$here/_build"
g++ -O2 -std=c++17 -ffp-contract=off -fopenmp -fPIC -shared -Wall -Wextra -Wno-unused-parameter \
    -I"$here/../include" "$here/clip_oracle.cpp" -o "$here/_build/libclip_oracle.so"
