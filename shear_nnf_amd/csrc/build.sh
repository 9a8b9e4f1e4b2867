#!/bin/bash
# Builds libshear_rk3.so for gfx950 (MI355X) next to the package sources.
# The host API and x-z/banded kernels (snf_api.hip) and the spectral y plans
# (snf_yplans_*.hip, a few line plans each) compile in parallel, then link.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
out="${1:-$here/../libshear_rk3.so}"
jobs="${SNF_BUILD_JOBS:-$(nproc)}"
tmp="$(mktemp -d)"
trap 'rm -rf "$tmp"' EXIT
flags=(--offload-arch=gfx950 -O3 -std=c++17 -fPIC -fvisibility=hidden
       -Xclang -target-feature -Xclang -packed-fp32-ops -Wno-unused-result)
srcs=("$here/snf_api.hip" "$here"/snf_yplans_*.hip)
printf '%s\n' "${srcs[@]}" | xargs -P "$jobs" -I{} sh -c \
  'hipcc "$@" -c "{}" -o "'"$tmp"'/$(basename "{}" .hip).o"' _ "${flags[@]}"
hipcc --offload-arch=gfx950 -shared -fPIC -o "$out.tmp" "$tmp"/*.o
mv "$out.tmp" "$out"
