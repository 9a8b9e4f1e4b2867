#!/bin/bash
# Builds libshear_rk3.so for gfx950 (MI355X) next to the package sources.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
out="${1:-$here/../libshear_rk3.so}"
hipcc --offload-arch=gfx950 -O3 -std=c++17 -fPIC -shared \
  -Xclang -target-feature -Xclang -packed-fp32-ops \
  -Wno-unused-result \
  -o "$out.tmp" "$here/snf_api.hip"
mv "$out.tmp" "$out"
