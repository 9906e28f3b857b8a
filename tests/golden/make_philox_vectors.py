"""Write tests/golden/philox_vectors.json: Philox4x32-10 outputs for random
counters and keys, computed by two independent implementations compiled on
the host from their source headers:

  * rocRAND  /opt/rocm/include/rocrand/rocrand_philox4x32_10.h
    (philox4x32_10_engine::ten_rounds, reached through a subclass);
  * ATen     torch/include/ATen/core/PhiloxRNGEngine.h
    (philox_engine with reset_state(seed, subsequence) and set_offset(offset),
    whose first four draws are the Philox block of counter
    (offset_lo, offset_hi, subsequence_lo, subsequence_hi) and key = seed).

The two must agree with each other; the test suite then pins the oracle's
Philox twin and the library's host/device Philox against the file.
Run: python tests/golden/make_philox_vectors.py  (needs hipcc and torch).
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, 'philox_vectors.json')

ROCRAND_SRC = r'''
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
struct probe : rocrand_device::philox4x32_10_engine {
  uint4 run(uint4 c, uint2 k) { return this->ten_rounds(c, k); }
};
int main() {
  unsigned c0, c1, c2, c3, k0, k1;
  probe p;
  while (std::scanf("%u %u %u %u %u %u", &c0, &c1, &c2, &c3, &k0, &k1) == 6) {
    uint4 r = p.run(uint4{c0, c1, c2, c3}, uint2{k0, k1});
    std::printf("%u %u %u %u\n", r.x, r.y, r.z, r.w);
  }
  return 0;
}
'''

ATEN_SRC = r'''
#include <ATen/core/PhiloxRNGEngine.h>
#include <cstdio>
int main() {
  unsigned c0, c1, c2, c3, k0, k1;
  while (std::scanf("%u %u %u %u %u %u", &c0, &c1, &c2, &c3, &k0, &k1) == 6) {
    at::philox_engine e(((uint64_t)k1 << 32) | k0, ((uint64_t)c3 << 32) | c2, 0);
    e.set_offset(((uint64_t)c1 << 32) | c0);
    unsigned a = e(), b = e(), c = e(), d = e();
    std::printf("%u %u %u %u\n", a, b, c, d);
  }
  return 0;
}
'''


def _build_and_run(src, cmd_prefix, name, stdin):
    with tempfile.TemporaryDirectory() as td:
        cpp = os.path.join(td, name + '.cpp')
        exe = os.path.join(td, name)
        with open(cpp, 'w') as f:
            f.write(src)
        subprocess.run(cmd_prefix + [cpp, '-o', exe], check=True,
                       capture_output=True)
        res = subprocess.run([exe], input=stdin, capture_output=True,
                             text=True, check=True)
    return [[int(v) for v in line.split()] for line in res.stdout.splitlines()]


def main():
    import torch
    tinc = os.path.join(os.path.dirname(torch.__file__), 'include')
    rng = np.random.default_rng(20261202)
    ctr = rng.integers(0, 2 ** 32, size=(256, 4), dtype=np.uint64)
    key = rng.integers(0, 2 ** 32, size=(256, 2), dtype=np.uint64)
    # include the counter layout the kernels use: (block, path, stream, epoch)
    ctr[:16, 0] = np.arange(16)
    ctr[:16, 1] = 12345
    ctr[:16, 2] = 0
    ctr[:16, 3] = 0
    stdin = ''.join('{} {} {} {} {} {}\n'.format(*c, *k) for c, k in zip(ctr, key))
    roc = _build_and_run(ROCRAND_SRC, ['hipcc', '-std=c++17', '-O1',
                                       '--offload-arch=gfx950'], 'roc', stdin)
    aten = _build_and_run(ATEN_SRC, ['g++', '-std=c++17', '-O1', '-I', tinc],
                          'aten', stdin)
    if roc != aten:
        sys.exit('rocRAND and ATen Philox disagree')
    with open(OUT, 'w') as f:
        json.dump({'generator': 'tests/golden/make_philox_vectors.py',
                   'sources': ['rocrand/rocrand_philox4x32_10.h',
                               'ATen/core/PhiloxRNGEngine.h'],
                   'counter': ctr.tolist(), 'key': key.tolist(),
                   'output': roc}, f)
    print('wrote', OUT, len(roc), 'vectors')


if __name__ == '__main__':
    main()
