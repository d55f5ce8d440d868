#!/bin/bash
# Builds libcovasim_amd.so for gfx950 from the HIP sources in this folder.
set -euo pipefail
here="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
out="${here}/../libcovasim_amd.so"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
# hash of every source and header that goes into the library (reported by cvk_source_hash)
src_hash="0x$(cat "${here}"/*.hip "${here}"/*.h "${here}/../../include/covasim_amd.h" | sha256sum | cut -c1-16)ull"
"${HIPCC}" --offload-arch=gfx950 -O3 -std=c++17 -fPIC -shared \
  -ffp-contract=off -fhip-fp32-correctly-rounded-divide-sqrt -fno-gpu-rdc \
  -Xclang -target-feature -Xclang -packed-fp32-ops \
  -Wall -Wno-unused-function -DCVK_SOURCE_HASH="${src_hash}" \
  "${here}/cvk_mirror.hip" "${here}/cvk_day.hip" "${here}/cvk_batch.hip" "${here}/cvk_relax.hip" "${here}/cvk_popgen.hip" "${here}/cvk_results.hip" "${here}/cvk_clip.hip" "${here}/cvk_diag.hip" "${here}/cvk_sources.hip" \
  -o "${out}.tmp"
mv "${out}.tmp" "${out}"
echo "built ${out}"
