"""Synthetic meshes and spring layouts for the batch configurations.

Disk cells are meshed with concentric rings (ring k has 6k nodes, 6k^2
triangles in total for K rings, all counter-clockwise like the reference's
shipped meshes), optionally with interior node jitter.  Springs are placed at
uniform random points of the disk, located in their host triangle and given
barycentric coordinates; with rest length 0 each anchor is the same point on
the (rigid) substrate.
"""
import math

import numpy as np


def disk_rings(K, radius=10.0):
    """Concentric-ring disk: returns X (N, 2), tris (T, 3) int32, boundary mask."""
    pts = [(0.0, 0.0)]
    rings = [[0]]
    for k in range(1, K + 1):
        r = radius * k / K
        n = 6 * k
        idx = []
        for m in range(n):
            a = 2.0 * math.pi * m / n
            idx.append(len(pts))
            pts.append((r * math.cos(a), r * math.sin(a)))
        rings.append(idx)
    tris = []
    for j in range(6):
        tris.append((0, rings[1][j], rings[1][(j + 1) % 6]))
    for k in range(2, K + 1):
        inner, outer = rings[k - 1], rings[k]
        ni, no = len(inner), len(outer)
        i = j = 0
        while i < ni or j < no:
            ai = 2.0 * math.pi * (i + 1) / ni
            ao = 2.0 * math.pi * (j + 1) / no
            if j < no and (i >= ni or ao <= ai):
                tris.append((inner[i % ni], outer[j % no], outer[(j + 1) % no]))
                j += 1
            else:
                tris.append((inner[i % ni], outer[j % no], inner[(i + 1) % ni]))
                i += 1
    X = np.array(pts, dtype=np.float64)
    T = np.array(tris, dtype=np.int32)
    bnd = np.zeros(len(X), dtype=bool)
    bnd[rings[K]] = True
    return X, T, bnd


def rings_for_triangles(T_target):
    return max(2, int(round(math.sqrt(T_target / 6.0))))


def signed_areas(X, T):
    a = X[T]
    return 0.5 * ((a[:, 1, 0] - a[:, 0, 0]) * (a[:, 2, 1] - a[:, 0, 1]) -
                  (a[:, 1, 1] - a[:, 0, 1]) * (a[:, 2, 0] - a[:, 0, 0]))


def jitter(X, T, bnd, frac, rng):
    """Move interior nodes by a vector uniform in a disk of radius frac * mean edge."""
    e = np.concatenate([X[T[:, 1]] - X[T[:, 0]], X[T[:, 2]] - X[T[:, 1]], X[T[:, 0]] - X[T[:, 2]]])
    h = np.mean(np.linalg.norm(e, axis=1))
    n = len(X)
    r = frac * h * np.sqrt(rng.uniform(0, 1, n))
    th = rng.uniform(0, 2 * np.pi, n)
    d = np.stack([r * np.cos(th), r * np.sin(th)], axis=1)
    d[bnd] = 0.0
    Y = X + d
    assert (signed_areas(Y, T) > 0).all()
    return Y


def locate(X, T, P, cell=None):
    """Host triangle and barycentric coordinates of points P: uniform-grid
    buckets (about two triangles each) listing the triangles whose bounding
    box meets them in increasing index; a point takes the first triangle of
    its bucket that contains it (barycentric coordinates >= -1e-12).
    Vectorised over the points, one candidate slot of every bucket at a time."""
    X = np.asarray(X, dtype=np.float64)
    T = np.asarray(T)
    P = np.asarray(P, dtype=np.float64).reshape(-1, 2)
    a = X[T]
    lo, hi = a.min(axis=1), a.max(axis=1)
    xmin, ymin = X.min(axis=0) - 1e-9
    xmax, ymax = X.max(axis=0) + 1e-9
    nb = max(1, int(math.sqrt(len(T) / 2)))
    sx, sy = (xmax - xmin) / nb, (ymax - ymin) / nb
    ix0 = np.floor((lo[:, 0] - xmin) / sx).astype(np.int64)
    ix1 = np.floor((hi[:, 0] - xmin) / sx).astype(np.int64)
    iy0 = np.floor((lo[:, 1] - ymin) / sy).astype(np.int64)
    iy1 = np.floor((hi[:, 1] - ymin) / sy).astype(np.int64)
    # (bucket, triangle) pairs; the bucket key packs (bx, by) with an offset
    # wide enough for every coordinate a point or box can produce
    span = nb + 4
    keys, tids = [], []
    tri_ids = np.arange(len(T), dtype=np.int64)
    for dx in range(int((ix1 - ix0).max()) + 1):
        for dy in range(int((iy1 - iy0).max()) + 1):
            m = (ix0 + dx <= ix1) & (iy0 + dy <= iy1)
            keys.append((ix0[m] + dx + 2) * span + (iy0[m] + dy + 2))
            tids.append(tri_ids[m])
    keys = np.concatenate(keys)
    tids = np.concatenate(tids)
    order = np.lexsort((tids, keys))  # by bucket, then triangle index
    keys, tids = keys[order], tids[order]
    ukeys, start, count = np.unique(keys, return_index=True, return_counts=True)
    pk = (np.floor_divide(P[:, 0] - xmin, sx).astype(np.int64) + 2) * span + \
        (np.floor_divide(P[:, 1] - ymin, sy).astype(np.int64) + 2)
    slot = np.searchsorted(ukeys, pk)
    has = (slot < len(ukeys)) & (ukeys[np.minimum(slot, len(ukeys) - 1)] == pk)
    b0 = np.where(has, start[np.minimum(slot, len(ukeys) - 1)], 0)
    nc = np.where(has, count[np.minimum(slot, len(ukeys) - 1)], 0)
    tri = np.full(len(P), -1, dtype=np.int32)
    bary = np.zeros((len(P), 3))
    todo = np.nonzero(nc > 0)[0]
    k = 0
    while len(todo):
        t = tids[b0[todo] + k]
        p0, p1, p2 = a[t, 0], a[t, 1], a[t, 2]
        q = P[todo]
        det = (p1[:, 0] - p0[:, 0]) * (p2[:, 1] - p0[:, 1]) - (p2[:, 0] - p0[:, 0]) * (p1[:, 1] - p0[:, 1])
        l1 = ((q[:, 0] - p0[:, 0]) * (p2[:, 1] - p0[:, 1]) - (p2[:, 0] - p0[:, 0]) * (q[:, 1] - p0[:, 1])) / det
        l2 = ((p1[:, 0] - p0[:, 0]) * (q[:, 1] - p0[:, 1]) - (q[:, 0] - p0[:, 0]) * (p1[:, 1] - p0[:, 1])) / det
        l0 = 1.0 - l1 - l2
        hit = (l0 >= -1e-12) & (l1 >= -1e-12) & (l2 >= -1e-12)
        idx = todo[hit]
        tri[idx] = t[hit]
        bary[idx] = np.stack([l0[hit], l1[hit], l2[hit]], axis=1)
        k += 1
        todo = todo[~hit]
        todo = todo[nc[todo] > k]
    return tri, bary


def random_springs(X, T, n, radius, rng):
    """n springs at uniform random points of the disk (rejection sampling),
    anchored at the same point (rest length 0)."""
    P = np.zeros((n, 2))
    k = 0
    while k < n:
        q = rng.uniform(-radius, radius, (2 * (n - k) + 8, 2))
        q = q[(q ** 2).sum(axis=1) < (0.98 * radius) ** 2]
        take = min(len(q), n - k)
        P[k:k + take] = q[:take]
        k += take
    tri, bary = locate(X, T, P)
    ok = tri >= 0
    tri, bary, P = tri[ok], bary[ok], P[ok]
    anchor = (bary[:, :, None] * X[T[tri]]).sum(axis=1)
    return tri, bary, anchor


def stiffening_table(n=64, k0=1000.0, k1=31000.0, d_max=0.01):
    """Monotone stiffening force table in the reference's (px, py, k) form:
    k rises linearly from k0 to k1 pN/um over extensions 0..d_max um; py is
    continuous (py_{j+1} = py_j + k_j (px_{j+1} - px_j))."""
    px = np.linspace(0.0, d_max, n)
    k = np.linspace(k0, k1, n)
    py = np.zeros(n)
    for j in range(1, n):
        py[j] = py[j - 1] + k[j - 1] * (px[j] - px[j - 1])
    return px, py, k


def substrate_grid(n_side=50, half=20.0):
    """Regular triangulated square [-half, half]^2 with n_side^2 nodes, split
    along one diagonal per quad (the create_sheet pattern of
    geometryinit.cpp:208-240, laid in the XY plane)."""
    g = np.linspace(-half, half, n_side)
    xx, yy = np.meshgrid(g, g, indexing="xy")
    X = np.stack([xx.ravel(), yy.ravel()], axis=1)
    tris = []
    for j in range(n_side - 1):
        for i in range(n_side - 1):
            v0 = j * n_side + i
            v1, v2, v3 = v0 + 1, v0 + n_side, v0 + n_side + 1
            tris.append((v0, v1, v3))
            tris.append((v0, v3, v2))
    return X, np.array(tris, dtype=np.int32)
