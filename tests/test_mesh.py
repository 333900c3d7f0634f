"""The synthetic workloads' mesh helpers: vectorised point location equals a
plain per-point scan of the same buckets, bit for bit."""
import math

import numpy as np

from wholecellfe_amd import mesh


def locate_scalar(X, T, P):
    a = X[T]
    lo, hi = a.min(axis=1), a.max(axis=1)
    xmin, ymin = X.min(axis=0) - 1e-9
    xmax, ymax = X.max(axis=0) + 1e-9
    nb = max(1, int(math.sqrt(len(T) / 2)))
    sx, sy = (xmax - xmin) / nb, (ymax - ymin) / nb
    buckets = {}
    ix0 = np.floor((lo[:, 0] - xmin) / sx).astype(int)
    ix1 = np.floor((hi[:, 0] - xmin) / sx).astype(int)
    iy0 = np.floor((lo[:, 1] - ymin) / sy).astype(int)
    iy1 = np.floor((hi[:, 1] - ymin) / sy).astype(int)
    for t in range(len(T)):
        for bx in range(ix0[t], ix1[t] + 1):
            for by in range(iy0[t], iy1[t] + 1):
                buckets.setdefault((bx, by), []).append(t)
    tri = np.full(len(P), -1, dtype=np.int32)
    bary = np.zeros((len(P), 3))
    for k, p in enumerate(P):
        key = (int((p[0] - xmin) // sx), int((p[1] - ymin) // sy))
        for t in buckets.get(key, ()):
            p0, p1, p2 = a[t]
            det = (p1[0] - p0[0]) * (p2[1] - p0[1]) - (p2[0] - p0[0]) * (p1[1] - p0[1])
            l1 = ((p[0] - p0[0]) * (p2[1] - p0[1]) - (p2[0] - p0[0]) * (p[1] - p0[1])) / det
            l2 = ((p1[0] - p0[0]) * (p[1] - p0[1]) - (p[0] - p0[0]) * (p1[1] - p0[1])) / det
            l0 = 1.0 - l1 - l2
            if l0 >= -1e-12 and l1 >= -1e-12 and l2 >= -1e-12:
                tri[k] = t
                bary[k] = (l0, l1, l2)
                break
    return tri, bary


def test_vectorised_locate_equals_scalar_scan():
    rng = np.random.default_rng(0)
    for K in (2, 5, 13, 30):
        X, T, b = mesh.disk_rings(K, 10.0)
        X = mesh.jitter(X, T, b, 0.05, rng)
        P = rng.uniform(-11, 11, (2000, 2))
        # points on nodes and edge midpoints: ties between triangles
        P = np.concatenate([P, X[:50], 0.5 * (X[T[:40, 0]] + X[T[:40, 1]])])
        t1, b1 = locate_scalar(X, T, P)
        t2, b2 = mesh.locate(X, T, P)
        assert np.array_equal(t1, t2) and np.array_equal(b1, b2)
        assert (t2 >= 0).sum() > 1000
    SX, ST = mesh.substrate_grid(50, 20.0)
    P = rng.uniform(-21, 21, (2000, 2))
    t1, b1 = locate_scalar(SX, ST, P)
    t2, b2 = mesh.locate(SX, ST, P)
    assert np.array_equal(t1, t2) and np.array_equal(b1, b2)
