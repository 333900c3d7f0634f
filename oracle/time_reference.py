"""Per-call time of the reference's own per-step residual work, in the build
container (the reference does not travel to the GPU host).

Runs oracle/_ref/refharness_<variant> (the reference's UL_growth_explicit.cpp
and nondestructivetrimesh.cpp compiled from their sources by `make -C oracle
ref`) in its `time` mode on the reference's shipped 483-node / 905-triangle
disk (tests/golden/mesh_disk483.npz) against the 50 x 50 substrate grid of
the one-cell workload: calculateAreaTot + calculateRes (cell law),
calculateForce_FA and updateFA, one thread, as the reference runs them once
per explicit step (main.cpp:356-363).  These are explicit-step costs, not
Newton steps.

    python oracle/time_reference.py [--reps 400]
"""
import argparse
import json
import os
import platform
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, REPO)


def main():
    from wholecellfe_amd import mesh
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=400)
    a = ap.parse_args()
    m = np.load(os.path.join(REPO, "tests", "golden", "mesh_disk483.npz"), allow_pickle=False)
    N = len(m["X"])
    X = np.zeros((N, 3))
    X[:, :2] = m["X"]
    rng = np.random.default_rng(7)
    x = X.copy()
    x[:, :2] += 0.01 * rng.standard_normal((N, 2))
    SX2, ST = mesh.substrate_grid(50, 20.0)
    SX = np.zeros((len(SX2), 3))
    SX[:, :2] = SX2
    ea = np.zeros((N, 6))
    ea[:, [0, 1, 3]] = 0.01 * rng.standard_normal((N, 3))
    inputs = dict(X=X, x=x, tris=m["tris"].astype(np.int64), ea=ea, cA=np.full(N, 100.0), cC=rng.uniform(0.2, 1, N),
                  fau=np.pad(0.004 * rng.standard_normal((N, 2)), ((0, 0), (0, 1))), SX=SX, Sx=SX.copy(),
                  Stris=ST.astype(np.int64))
    meta = dict(dt=0.005, time=1.0, kint=1000.0, reps=a.reps)
    out = {"cpu": platform.processor() or platform.machine(), "threads": 1, "nodes": N, "triangles": len(m["tris"]),
           "substrate_nodes": len(SX), "substrate_triangles": len(ST), "reps": a.reps}
    for variant, kint in (("k1", 1000.0), ("kMD", 0.0)):
        d = tempfile.mkdtemp(prefix="wcfe_time_")
        try:
            for k, v in inputs.items():
                np.ascontiguousarray(v).tofile(os.path.join(d, k + ".bin"))
            with open(os.path.join(d, "meta.txt"), "w") as f:
                for k, v in dict(meta, kint=kint).items():
                    f.write("%s %.17g\n" % (k, v))
            subprocess.run([os.path.join(HERE, "_ref", "refharness_" + variant), "time", d], check=True)
            t = np.fromfile(os.path.join(d, "out_time.bin"), dtype=np.float64)
        finally:
            shutil.rmtree(d)
        out[variant] = {"calculateRes_ms": t[0] * 1e3, "calculateForce_FA_ms": t[1] * 1e3, "updateFA_ms": t[2] * 1e3,
                        "sum_ms": (t[0] + t[1] + t[2]) * 1e3}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
