"""Synthetic batch configurations (cell meshes, materials, springs).

Each builder returns (cells, tables): cells is a list of dicts with the
arguments of Context.add_cell, tables a dict table_id -> (px, py, k).
Units: um, pN, s, Pa.
"""
import math

import numpy as np

from . import mesh

TABLE_ID = 0


def _disk_cell(T_target, rng, radius=10.0, jitter=0.05):
    K = mesh.rings_for_triangles(T_target)
    X, T, bnd = mesh.disk_rings(K, radius)
    if jitter > 0:
        X = mesh.jitter(X, T, bnd, jitter, rng)
    return X, T


def single_cell(T_target=905, n_spr=300, E=1000.0, nu=0.3, thick=0.2, spring="linear", k_lin=1000.0, seed=1,
                kinetics=0, solver=0):
    """One disk cell (config families 1, 2, 4): linear plane-stress CST,
    springs at uniform random points with rest length 0."""
    rng = np.random.default_rng(seed)
    X, T = _disk_cell(T_target, rng)
    tri, bary, anchor = mesh.random_springs(X, T, n_spr, 10.0, rng)
    cell = dict(X=X, tris=T, law=1, E=E, nu=nu, thick=thick, spr_tri=tri, spr_bary=bary, spr_anchor=anchor,
                spring_law=0 if spring == "linear" else 1, k_lin=k_lin, table_id=TABLE_ID, kinetics=kinetics,
                solver=solver)
    return [cell], {TABLE_ID: mesh.stiffening_table()}


# the named workloads: family 1 (one 905-triangle-class cell with 300 linear
# integrins on a rigid 50 x 50 substrate grid, 200 load steps, FP64), family 2
# (one 20k-triangle cell, dense FP32 LU), family 3 (256 uneven cells, the
# headline batch), family 4 (one 500k-triangle cell, block-CSR PCG), family 5
# (1024 Bell replicas)
WORKLOADS = ("single905", "batched256", "single20k", "single500k", "replicas1024")

# family 1: rigid substrate, a regular 50 x 50 node grid over [-20, 20]^2 um
C1_SUBSTRATE = dict(n_side=50, half=20.0)


def c1_substrate():
    """The rigid substrate of family 1 (nodes, triangles).  Rigid, it only
    holds the anchors: each spring's anchor is the substrate point under its
    attachment point (rest length 0), located in the grid's triangles."""
    return mesh.substrate_grid(**C1_SUBSTRATE)


def workload(name):
    if name == "single905":
        # 300 integrins, k = 1 pN/nm, anchored on the rigid substrate grid
        return single_cell(T_target=905, n_spr=300, spring="linear", k_lin=1000.0, seed=1, solver=1)
    if name == "batched256":
        return batched_cells(256)
    if name == "single20k":
        return single_cell(T_target=20000, n_spr=5000, spring="table", seed=2, solver=1)
    if name == "single500k":
        return single_cell(T_target=500000, n_spr=100000, spring="table", seed=4, solver=2)
    if name == "replicas1024":
        return replicas(1024)
    raise KeyError(name)


def batched_cells(n_cells=256, t_min=200, t_max=5000, seed=3, spr_per_tri=1.0 / 3.0, kinetics=0):
    """Config family 3: independent cells, T log-uniform in [t_min, t_max],
    E uniform in [0.5, 5] kPa, spring law drawn from {1 pN/nm, 31 pN/nm,
    64-knot stiffening table}, T/3 springs per cell."""
    rng = np.random.default_rng(seed)
    cells = []
    for _ in range(n_cells):
        Tn = int(round(math.exp(rng.uniform(math.log(t_min), math.log(t_max)))))
        X, T = _disk_cell(Tn, rng)
        E = rng.uniform(500.0, 5000.0)
        law = int(rng.integers(3))
        n_spr = max(3, int(len(T) * spr_per_tri))
        tri, bary, anchor = mesh.random_springs(X, T, n_spr, 10.0, rng)
        cells.append(dict(X=X, tris=T, law=1, E=E, nu=0.3, thick=0.2, spr_tri=tri, spr_bary=bary,
                          spr_anchor=anchor, spring_law=1 if law == 2 else 0, k_lin=1000.0 if law == 0 else 31000.0,
                          table_id=TABLE_ID, kinetics=kinetics))
    return cells, {TABLE_ID: mesh.stiffening_table()}


def replicas(n_rep=1024, t_min=1000, t_max=10000, seed=5):
    """Config family 5: stochastic replicas with Bell kinetics."""
    rng = np.random.default_rng(seed)
    cells = []
    for _ in range(n_rep):
        Tn = int(rng.integers(t_min, t_max + 1))
        X, T = _disk_cell(Tn, rng)
        n_spr = max(3, len(T) // 3)
        tri, bary, anchor = mesh.random_springs(X, T, n_spr, 10.0, rng)
        cells.append(dict(X=X, tris=T, law=1, E=1000.0, nu=0.3, thick=0.2, spr_tri=tri, spr_bary=bary,
                          spr_anchor=anchor, spring_law=0, k_lin=1000.0, table_id=TABLE_ID, kinetics=2))
    return cells, {TABLE_ID: mesh.stiffening_table()}


def load(ctx, cells, tables):
    """Register tables and cells on a Context and finalize it."""
    for tid, (px, py, k) in tables.items():
        ctx.set_table(tid, px, py, k)
    for c in cells:
        ctx.add_cell(**c)
    ctx.finalize()


def lu_flops(cells, symmetric=False):
    """Dense factorisation + one solve per cell, n = 2 N: (2/3) n^3 + 2 n^2
    for LU, (1/3) n^3 + 2 n^2 for a symmetric (LDL^T-shaped) factorisation."""
    f = 0.0
    for c in cells:
        n = 2 * len(c["X"])
        f += (1.0 if symmetric else 2.0) / 3.0 * n ** 3 + 2.0 * n ** 2
    return f
