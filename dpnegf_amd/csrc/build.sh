#!/bin/bash
# Builds libdpnegf_amd.so for gfx950 in-tree (dpnegf_amd/_lib/).
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
out="$here/../_lib"
bld="$here/../_build"
mkdir -p "$out" "$bld"
FLAGS="--offload-arch=gfx950 -O3 -std=c++17 -fPIC -Xclang -target-feature -Xclang -packed-fp32-ops -I$here -I$here/../../include"
objs=()
pids=()
for f in rgf_transmission.hip lead_self_energy.hip rgf_full.hip capi.cpp; do
  o="$bld/${f%.*}.o"
  rm -f "$o"
  extra=()
  # The batched product's accumulators are loop-carried across a runtime K
  # loop; in the default accumulator-register form the compiler copies all of
  # them between register files every iteration.  No other kernel in this file
  # uses the matrix cores.
  if [[ "$f" == rgf_full.hip ]]; then
    extra=(-mllvm -amdgpu-mfma-vgpr-form)
  fi
  if [[ "$f" == *.cpp ]]; then
    hipcc $FLAGS -x hip -c "$here/$f" -o "$o" &
  else
    hipcc $FLAGS "${extra[@]}" -c "$here/$f" -o "$o" &
  fi
  pids+=($!)
  objs+=("$o")
done
fail=0
for p in "${pids[@]}"; do
  wait "$p" || fail=1
done
if [[ $fail -ne 0 ]]; then
  echo "build failed" >&2
  exit 1
fi
hipcc --offload-arch=gfx950 -shared -fPIC "${objs[@]}" -o "$out/libdpnegf_amd.so.tmp"
mv "$out/libdpnegf_amd.so.tmp" "$out/libdpnegf_amd.so"
echo "built $out/libdpnegf_amd.so"
