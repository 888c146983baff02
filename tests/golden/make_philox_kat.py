"""Write philox4x32_10_kat.npz: Philox4x32-10 block outputs for a set of
(counter, key) pairs, computed by rocRAND's own implementation
(/opt/rocm/include/rocrand/rocrand_philox4x32_10.h, compiled for the host
with hipcc) -- an implementation independent of this package's.

    python tests/golden/make_philox_kat.py
"""
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

SRC = r"""
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
int main() {
    unsigned int c[4], k[2];
    while (scanf("%u %u %u %u %u %u", &c[0], &c[1], &c[2], &c[3], &k[0], &k[1]) == 6) {
        unsigned long long seed = (unsigned long long)k[0] | ((unsigned long long)k[1] << 32);
        unsigned long long sub = (unsigned long long)c[2] | ((unsigned long long)c[3] << 32);
        unsigned long long off = ((unsigned long long)c[0] | ((unsigned long long)c[1] << 32)) * 4ull;
        rocrand_device::philox4x32_10_engine e(seed, sub, off);
        uint4 r = e.next4();
        printf("%u %u %u %u\n", r.x, r.y, r.z, r.w);
    }
    return 0;
}
"""


def cases():
    rng = np.random.default_rng(20260101)
    # the engine's offset is 4*(c0 + 2^32 c1) numbers, so c1 < 2^30 keeps it
    # from wrapping; counters beyond that are covered by the published
    # Random123 vectors in the test itself
    fixed = [(0, 0, 0, 0, 0, 0), (1, 0, 0, 0, 0, 0), (0, 0, 7, 0, 42, 0), (12345, 0, 3, 0, 42, 0),
             (0xFFFFFFFF, 0x3FFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF)]
    rnd = []
    for _ in range(59):
        r = [int(x) for x in rng.integers(0, 2 ** 32, 6, dtype=np.uint64)]
        r[1] &= 0x3FFFFFFF
        rnd.append(tuple(r))
    return np.array(fixed + rnd, dtype=np.uint32)


def main():
    cs = cases()
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "kat.cpp")
        exe = os.path.join(d, "kat")
        open(src, "w").write(SRC)
        subprocess.run(["hipcc", "-O1", "-o", exe, src], check=True)
        inp = "\n".join(" ".join(str(int(x)) for x in row) for row in cs) + "\n"
        out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout
    res = np.array([[int(x) for x in ln.split()] for ln in out.strip().split("\n")], dtype=np.uint32)
    np.savez(os.path.join(HERE, "philox4x32_10_kat.npz"), counter=cs[:, :4], key=cs[:, 4:], out=res)
    print(res[:3])


if __name__ == "__main__":
    main()
