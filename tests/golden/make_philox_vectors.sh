#!/bin/sh
# Regenerates philox_rocrand.txt: the first 8 words of rocRAND's host-compiled
# philox4x32_10_engine for several (seed, subsequence) pairs.  Needs g++ and
# the ROCm rocRAND headers; writes only into this directory.
set -e
here=$(cd "$(dirname "$0")" && pwd)
tmp=$(mktemp -d)
g++ -O1 -std=c++17 -D__HIP_PLATFORM_AMD__ -I/opt/rocm/include "$here/philox_rocrand_kat.cpp" -o "$tmp/kat"
"$tmp/kat" > "$here/philox_rocrand.txt"
rm -rf "$tmp"
