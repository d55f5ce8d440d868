#!/bin/bash
# Builds libcovasim_amd.so for gfx950 from the HIP sources in this folder: one compile per source, all in
# parallel (objects under _build/ beside them), then one link.
set -euo pipefail
here="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
out="${here}/../libcovasim_amd.so"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
# hash of every source and header that goes into the library (reported by cvk_source_hash)
src_hash="0x$(cat "${here}"/*.hip "${here}"/*.h "${here}/../../include/covasim_amd.h" | sha256sum | cut -c1-16)ull"
flags=(--offload-arch=gfx950 -O3 -std=c++17 -fPIC
       -ffp-contract=off -fhip-fp32-correctly-rounded-divide-sqrt -fno-gpu-rdc
       -Xclang -target-feature -Xclang -packed-fp32-ops
       -Wall -Wno-unused-function -DCVK_SOURCE_HASH="${src_hash}")
sources=(cvk_mirror cvk_day cvk_batch cvk_relax cvk_popgen cvk_results cvk_clip cvk_diag cvk_sources)
objdir="${here}/_build"
mkdir -p "${objdir}"
pids=()
objs=()
for s in "${sources[@]}"; do
  rm -f "${objdir}/${s}.o"
  "${HIPCC}" "${flags[@]}" -c "${here}/${s}.hip" -o "${objdir}/${s}.o" &
  pids+=("$!")
  objs+=("${objdir}/${s}.o")
done
failed=0
for p in "${pids[@]}"; do
  wait "${p}" || failed=1
done
if [ "${failed}" -ne 0 ]; then
  echo "build failed: a source did not compile" >&2
  exit 1
fi
"${HIPCC}" --offload-arch=gfx950 -shared -fPIC -fno-gpu-rdc "${objs[@]}" -o "${out}.tmp"
mv "${out}.tmp" "${out}"
echo "built ${out}"
