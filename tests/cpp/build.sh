#!/bin/bash
# Builds the C++ test programs into tests/cpp/_bin (run by __graft_entry__.build(); the binaries
# travel with the tree, are git-ignored, and are run by the tests):
#   bitstream_trace  oracle/bitstream_trace.cpp against this repository's BitStream.h (g++, host only)
#   dropin_volume    tests/cpp/dropin_volume.cpp against include/cudaCompress/*.h, linked with
#                    cudacompress_amd/lib/libcudacompress_amd.so (hipcc; host code only)
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
ROOT="$(cd "$HERE/../.." && pwd)"
BIN="$HERE/_bin"
mkdir -p "$BIN"
g++ -std=c++17 -O2 -Wall -I"$ROOT/include" "$ROOT/oracle/bitstream_trace.cpp" -o "$BIN/bitstream_trace"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
"$HIPCC" -std=c++17 -O2 -Wall -I"$ROOT/include" "$HERE/dropin_volume.cpp" \
    -L"$ROOT/cudacompress_amd/lib" -lcudacompress_amd -Wl,-rpath,'$ORIGIN/../../../cudacompress_amd/lib' \
    -o "$BIN/dropin_volume"
