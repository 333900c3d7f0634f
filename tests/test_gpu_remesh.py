"""Remesh -> device sync (wcfe_remesh_cell): a cell's new triangles after the
host's remeshing (edge splits with El Topo's midpoint interpolation,
edgesplitter.cpp:563-577, then a renumbering like its defragmentation,
surftrack.cpp:487-600), its state mapped on the device; the next load step
against the oracle on the new mesh; the unchanged-topology fast path; the
cost of one remesh in the 256-cell batch against one Newton step."""
import time

import numpy as np
import pytest

import cells
import oracle_py as O
import wholecellfe_amd as W
from wholecellfe_amd import configs, mesh

pytestmark = pytest.mark.gpu


def split_edges(T, n_nodes, n_split, rng):
    """Split n_split interior edges whose triangles are disjoint: returns the
    new triangles (counter-clockwise kept), and per new node (src0, src1, w)."""
    T = [list(t) for t in np.asarray(T).reshape(-1, 3)]
    edge_tris = {}
    for ti, t in enumerate(T):
        for a in range(3):
            p, q = t[a], t[(a + 1) % 3]
            edge_tris.setdefault((min(p, q), max(p, q)), []).append(ti)
    used, picks = set(), []
    keys = [k for k, v in edge_tris.items() if len(v) == 2]
    for idx in rng.permutation(len(keys)):
        k = keys[idx]
        t0, t1 = edge_tris[k]
        if t0 in used or t1 in used:
            continue
        used.update((t0, t1))
        picks.append(k)
        if len(picks) == n_split:
            break
    src0, src1, w = list(range(n_nodes)), list(range(n_nodes)), [1.0] * n_nodes
    for (a, b) in picks:
        e = len(src0)
        src0.append(a)
        src1.append(b)
        w.append(0.5)
        for ti in edge_tris[(a, b)]:
            t = T[ti]
            i = next(j for j in range(3) if {t[j], t[(j + 1) % 3]} == {a, b})
            p, q, r = t[i], t[(i + 1) % 3], t[(i + 2) % 3]
            T[ti] = [p, e, r]
            T.append([e, q, r])
    return np.array(T, dtype=np.int32), np.array(src0), np.array(src1), np.array(w)


def renumber(T, src0, src1, w, rng):
    """A renumbering of the nodes (defragmentation-like): new index j is old-side
    index perm[j]."""
    perm = rng.permutation(len(src0))
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    return inv[T].astype(np.int32), src0[perm], src1[perm], w[perm]


def relocate_springs(X_old, T_old, tri, bary, X_new, T_new):
    p = (bary.reshape(-1, 3)[:, :, None] * X_old[T_old[tri]]).sum(axis=1)
    t2, b2 = mesh.locate(X_new, T_new, p)
    assert np.all(t2 >= 0)
    return t2, b2


def test_split_renumber_then_load_step_vs_oracle():
    rng = np.random.default_rng(11)
    d = cells.cst_cell(K=7, n_spr=80, law_table=True, seed=5, time=1.0)
    other = cells.cst_cell(K=5, n_spr=40, seed=6, time=1.0)
    dt = 0.02
    ctx = W.Context(precision=W.PREC_MIXED, t0=1.0, dt=dt, newton_rtol=1e-11, newton_atol=0.0, max_refine=12)
    cells.add_to_context(ctx, d)
    cells.add_to_context(ctx, other)
    ctx.finalize()
    ctx.load_step()
    N = len(d["X"]) // 2
    X_old = ctx.ul_state(0)[0]
    x_old = ctx.positions(0)
    info0 = ctx.cell_info(0)
    other_x, other_info = ctx.positions(1), ctx.cell_info(1)
    T_old = d["tris"].reshape(-1, 3)
    T_new, s0, s1, w = split_edges(T_old, N, 6, rng)
    T_new, s0, s1, w = renumber(T_new, s0, s1, w, rng)
    X_map = w[:, None] * X_old.reshape(-1, 2)[s0] + (1 - w[:, None]) * X_old.reshape(-1, 2)[s1]
    x_map = w[:, None] * x_old.reshape(-1, 2)[s0] + (1 - w[:, None]) * x_old.reshape(-1, 2)[s1]
    spr_tri, spr_bary = relocate_springs(X_old.reshape(-1, 2), T_old, d["spr_tri"], d["spr_bary"], X_map, T_new)
    st_before, an_before = ctx.springs(0)
    print("case: %d -> %d nodes, %d -> %d triangles" % (N, len(s0), len(T_old), len(T_new)), flush=True)
    relaid = ctx.remesh_cell(0, T_new, s0, s1, w, spr_tri=spr_tri, spr_bary=spr_bary)
    assert relaid == 1
    # the mapped state, bit for bit
    X1, _, _, _ = ctx.ul_state(0)
    assert np.array_equal(X1, X_map.reshape(-1))
    assert np.array_equal(ctx.positions(0), x_map.reshape(-1))
    st1, an1 = ctx.springs(0)
    assert np.array_equal(st1, st_before) and np.array_equal(an1, an_before)
    info1 = ctx.cell_info(0)
    assert info1["step"] == info0["step"] and info1["time"] == info0["time"]
    # the other cell keeps its state
    assert np.array_equal(ctx.positions(1), other_x)
    assert ctx.cell_info(1)["step"] == other_info["step"]
    # the next load step on the new mesh against the oracle
    ctx.load_step()
    x2 = ctx.positions(0)
    ctx.close()
    px, py, kk = mesh.stiffening_table()
    dn = dict(X=X_map.reshape(-1), x=x_map.reshape(-1), tris=T_new.reshape(-1), law=1, E=1000.0, nu=0.3, thick=0.2,
              time=info0["time"], spr_tri=spr_tri, spr_bary=spr_bary.reshape(-1), spr_anchor=an_before,
              spr_state=st_before, spring_law=1, k_lin=1000.0, tab_px=px, tab_py=py, tab_k=kk)
    # with fixed anchors the CST equilibrium at the load time does not depend
    # on the start: solved from the rest positions (||G0|| the whole load)
    x_o, it, _ = O.newton_solve(O.make_cell(dn), X_map.reshape(-1), rtol=1e-11, atol=0.0, maxit=20)
    assert it >= 0
    err = np.max(np.abs(x2 - x_o)) / np.max(np.abs(x_o - X_map.reshape(-1)))
    print("remeshed cell, next load step: |x - x_oracle| / |u| = %.2e" % err, flush=True)
    assert err <= 1e-8


def test_node_state_maps_for_the_reference_model():
    """UL law with node-attached springs: strain, bound fraction and spring
    extension follow the node maps (bit for bit)."""
    rng = np.random.default_rng(3)
    d = cells.ul_cell(K=5, table=False, time=1.3, deform=0.0)
    d["x"] = d["X"].copy()
    ctx = W.Context(precision=W.PREC_MIXED, t0=1.3, newton_rtol=1e-10, newton_atol=1e-9)
    cells.add_to_context(ctx, d, kinetics=W.KIN_MEANFIELD)
    ctx.finalize()
    ctx.load_step()
    N = len(d["X"]) // 2
    X0, ea0, cC0, fau0 = ctx.ul_state(0)
    T_new, s0, s1, w = split_edges(d["tris"].reshape(-1, 3), N, 4, rng)
    T_new, s0, s1, w = renumber(T_new, s0, s1, w, rng)
    assert ctx.remesh_cell(0, T_new, s0, s1, w) == 1
    X1, ea1, cC1, fau1 = ctx.ul_state(0)
    mix = lambda v, k: (w[:, None] * v.reshape(-1, k)[s0] + (1 - w[:, None]) * v.reshape(-1, k)[s1]).reshape(-1)
    assert np.array_equal(X1, mix(X0, 2))
    assert np.array_equal(ea1, mix(ea0, 3))
    assert np.array_equal(cC1, mix(cC0, 1))
    assert np.array_equal(fau1, mix(fau0, 2))
    st = ctx.load_step()
    assert st["max_lin_relres"] <= 1e-8
    ctx.close()


def test_unchanged_topology_is_not_laid_out_again():
    d = cells.cst_cell(K=6, n_spr=50, seed=9, time=1.0)
    ctx = W.Context(precision=W.PREC_MIXED, t0=1.0)
    cells.add_to_context(ctx, d)
    ctx.finalize()
    ctx.load_step()
    x = ctx.positions(0)
    x_new = x + 1e-4 * np.random.default_rng(1).standard_normal(len(x))
    assert ctx.remesh_cell(0, d["tris"], x_new=x_new) == 0
    assert np.array_equal(ctx.positions(0), x_new)
    N = len(x) // 2
    ident = np.arange(N)
    assert ctx.remesh_cell(0, d["tris"], ident, ident, np.ones(N)) == 0
    st = ctx.load_step()
    assert st["max_lin_relres"] <= 1e-8
    ctx.close()


def test_remesh_cost_in_the_256_cell_batch():
    """One remesh (a split edge in the largest cell) of the 256-cell batch and
    the unchanged-topology path, timed against one Newton step."""
    cs, tables = configs.batched_cells(256)
    ctx = W.Context(precision=W.PREC_MIXED, t0=0.5, dt=0.005)
    configs.load(ctx, cs, tables)
    for _ in range(3):
        ctx.newton_step(stats=False)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        ctx.newton_step(stats=False)
    ctx.synchronize()
    t_step = (time.perf_counter() - t0) / 5
    big = max(range(len(cs)), key=lambda i: len(cs[i]["tris"]))
    c = cs[big]
    N = len(c["X"])
    T_new, s0, s1, w = split_edges(c["tris"], N, 1, np.random.default_rng(0))
    X_map = w[:, None] * c["X"][s0] + (1 - w[:, None]) * c["X"][s1]
    spr_tri, spr_bary = relocate_springs(c["X"], c["tris"], c["spr_tri"], c["spr_bary"], X_map, T_new)
    t0 = time.perf_counter()
    assert ctx.remesh_cell(big, T_new, s0, s1, w, spr_tri=spr_tri, spr_bary=spr_bary) == 1
    t_remesh = time.perf_counter() - t0
    t0 = time.perf_counter()
    assert ctx.remesh_cell(big, T_new) == 0
    t_fast = time.perf_counter() - t0
    st = ctx.newton_step()
    assert st["max_lin_relres"] <= 1e-8
    ctx.close()
    print("256-cell batch: Newton step %.3f ms; remesh with a topology change %.1f ms (%.0f steps); unchanged "
          "topology %.3f ms" % (t_step * 1e3, t_remesh * 1e3, t_remesh / t_step, t_fast * 1e3), flush=True)
