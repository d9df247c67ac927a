#!/bin/sh
# Host-only build of tests/node_check.hip (the per-node arithmetic of
# clip_amd/csrc/node.hpp) for the CPU tests of the adjoint algebra.
set -e
here=$(cd "$(dirname "$0")" && pwd)
mkdir -p "$here/_build"
hipcc=${HIPCC:-/opt/rocm/bin/hipcc}
"$hipcc" -x hip --offload-host-only -O2 -std=c++17 -fPIC -shared \
    -I"$here/../include" -I"$here/../clip_amd/csrc" "$here/node_check.hip" -o "$here/_build/libnode_check.so"
