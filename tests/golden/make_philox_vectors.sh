#!/bin/sh
# Rebuilds tests/golden/philox_vectors.bin from rocRAND's Philox (host code).
set -e
cd "$(dirname "$0")"
/opt/rocm/bin/hipcc -O1 -std=c++17 philox_vectors.cpp -o /tmp/philox_vectors
/tmp/philox_vectors philox_vectors.bin
