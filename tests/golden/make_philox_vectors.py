"""Writes philox_vectors.json: Philox4x32-10 outputs of rocRAND's engine.

rocRAND's engine (ROCm headers) is an implementation independent of this
repository; its state for (seed, subsequence, offset = 4*block) is key = seed and
counter = (block_lo, block_hi, subsequence_lo, subsequence_hi). Run:
    python tests/golden/make_philox_vectors.py
"""
import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, 'kat')
        subprocess.run(['/opt/rocm/bin/hipcc', '-x', 'hip', '--offload-arch=gfx950', '-O1',
                        os.path.join(HERE, 'philox_rocrand.cpp'), '-o', exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    rows = []
    for line in out.strip().splitlines():
        s, q, b, *w = (int(x) for x in line.split())
        rows.append(dict(key=[s & 0xFFFFFFFF, s >> 32], counter=[b & 0xFFFFFFFF, b >> 32, q & 0xFFFFFFFF, q >> 32],
                         out=w))
    with open(os.path.join(HERE, 'philox_vectors.json'), 'w') as f:
        json.dump(dict(source='rocRAND philox4x32_10_engine (ROCm 7.2 headers)', vectors=rows), f, indent=0)
    print(f'wrote {len(rows)} vectors')


if __name__ == '__main__':
    main()
