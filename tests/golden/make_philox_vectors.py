"""Write tests/golden/philox_rocrand.npz: Philox4x32-10 outputs computed by
rocRAND's own implementation (host-callable, /opt/rocm/include/rocrand), an
independent implementation of Salmon et al. (SC'11), for a set of counters
and keys including the ones the bond kinetics use (replica, spring, step, 0).

Usage: python tests/golden/make_philox_vectors.py   (needs hipcc; no GPU)
"""
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.default_rng(123)
    rows = [(0, 0, 0, 0, 0, 0), (0xffffffff,) * 6,
            (0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344, 0xa4093822, 0x299f31d0)]
    for rep in range(3):
        for spr in range(5):
            for step in range(4):
                rows.append((rep, spr, step, 0, 5, 0))
    rows += [tuple(int(v) for v in rng.integers(0, 2 ** 32, 6, dtype=np.uint64)) for _ in range(40)]
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "philox")
        subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O2", os.path.join(HERE, "philox_rocrand.cpp"), "-o", exe],
                       check=True)
        inp = "\n".join(" ".join(str(v) for v in r) for r in rows) + "\n"
        out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout
    o = np.array([[int(v) for v in l.split()] for l in out.strip().split("\n")], dtype=np.uint64)
    np.savez_compressed(os.path.join(HERE, "philox_rocrand.npz"), inputs=np.array(rows, dtype=np.uint64), outputs=o)
    print("wrote philox_rocrand.npz", o.shape)


if __name__ == "__main__":
    main()
