"""Parity at the sizes of the five benchmark workload families, against the
FP64 oracle on the same inputs:

  family 1  one cell (864 triangles, 300 linear integrins on a rigid grid),
            200 load steps, FP64 and mixed precision
  family 2  one 20k-triangle cell, 5,000 table springs, dense FP32 LU
  family 3  the full 256-cell batch (387,828 triangles), one load step
  family 4  one 500k-triangle cell, 100,000 table springs, block-CSR PCG
  family 5  16 Bell replicas (2.6k-9.6k triangles), 50 load steps

Positions are checked through the oracle's FP64 residual at the GPU's
positions (size independent: ||G(x_gpu)|| <= 1e-8 ||G(x_start)||) and,
wherever the oracle can solve the same problem in seconds (dense LU up to
1k unknowns, Jacobi-PCG Newton beyond), against the oracle's solution:
max |x_gpu - x_oracle| <= 1e-8 max |x_oracle - X|.  Bond states and anchors
are bit-exact.
"""
import time

import numpy as np
import pytest

import oracle_py as O
import wholecellfe_amd as W
from wholecellfe_amd import configs, mesh

pytestmark = pytest.mark.gpu

T0, DT = 0.5, 0.005


def oracle_dict(c, tables, x=None, time=T0, state=None, anchor=None):
    """The oracle's description of a workload cell at positions x."""
    px, py, kk = tables[c.get("table_id", 0)]
    n_spr = len(c["spr_tri"])
    X = np.asarray(c["X"], dtype=np.float64).reshape(-1)
    return dict(X=X, x=X.copy() if x is None else np.asarray(x, dtype=np.float64).reshape(-1),
                tris=np.asarray(c["tris"]).reshape(-1), law=c["law"], E=c["E"], nu=c["nu"], thick=c["thick"],
                time=time, spr_tri=c["spr_tri"], spr_bary=np.asarray(c["spr_bary"]).reshape(-1),
                spr_anchor=np.asarray(c["spr_anchor"] if anchor is None else anchor).reshape(-1),
                spr_state=np.ones(n_spr) if state is None else state, spring_law=c["spring_law"],
                k_lin=c["k_lin"], tab_px=px, tab_py=py, tab_k=kk)


def residual(d):
    G, _ = O.assemble(O.make_cell(d), want_K=False)
    return G


def log(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("prec", [W.PREC_FP64, W.PREC_MIXED])
def test_family1_200_load_steps(prec):
    cells, tables = configs.workload("single905")
    c = cells[0]
    SX, ST = configs.c1_substrate()
    tri, _ = mesh.locate(SX, ST, np.asarray(c["spr_anchor"]).reshape(-1, 2))
    assert np.all(tri >= 0)  # every anchor lies on the rigid substrate grid
    log("case: family 1, prec %d, %d nodes, %d triangles, %d springs" % (prec, len(c["X"]), len(c["tris"]),
                                                                        len(c["spr_tri"])))
    ctx = W.Context(precision=prec, t0=T0, dt=DT, newton_rtol=1e-10, newton_atol=1e-12, max_refine=12)
    configs.load(ctx, cells, tables)
    X = np.asarray(c["X"]).reshape(-1)
    x_prev = X.copy()
    worst, iters = 0.0, []
    for step in range(200):
        st = ctx.load_step()
        x = ctx.positions(0)
        info = ctx.cell_info(0)
        t = T0 + step * DT
        assert info["step"] == step + 1 and abs(info["time"] - (t + DT)) < 1e-12
        assert st["max_lin_relres"] <= 1e-8
        G = residual(oracle_dict(c, tables, x=x, time=t))
        G0 = residual(oracle_dict(c, tables, x=x_prev, time=t))
        rel = np.linalg.norm(G) / np.linalg.norm(G0)
        worst = max(worst, rel)
        assert rel <= 1e-8, (step, rel)
        iters.append(st["iterations"])
        if step in (0, 199):
            # with fixed anchors the equilibrium at load time t does not depend
            # on the path: the oracle solves it from the rest positions (where
            # ||G0|| is the whole load, far above the FP64 residual floor)
            x_o, it, _ = O.newton_solve(O.make_cell(oracle_dict(c, tables, time=t)), X, rtol=1e-12, atol=0.0,
                                        maxit=10)
            assert it >= 0
            err = np.max(np.abs(x - x_o)) / np.max(np.abs(x_o - X))
            log("step %d: |x - x_oracle| / |u| = %.2e" % (step + 1, err))
            assert err <= 1e-8
        x_prev = x
    ctx.close()
    log("family 1: 200 load steps, Newton iterations %d..%d, worst ||G|| / ||G0|| = %.2e"
        % (min(iters), max(iters), worst))


def test_family2_20k_dense_vs_oracle():
    cells, tables = configs.workload("single20k")
    c = cells[0]
    log("case: family 2, %d nodes, %d triangles, %d table springs" % (len(c["X"]), len(c["tris"]), len(c["spr_tri"])))
    # the FP64 residual floor of this cell is about 1.5e-12 of ||G0||
    ctx = W.Context(precision=W.PREC_MIXED, t0=T0, dt=DT, newton_rtol=1e-11, newton_atol=0.0, max_refine=12)
    configs.load(ctx, cells, tables)
    st = ctx.load_step()
    x = ctx.positions(0)
    ctx.close()
    assert st["max_lin_relres"] <= 1e-8
    X = np.asarray(c["X"]).reshape(-1)
    G = residual(oracle_dict(c, tables, x=x))
    G0 = residual(oracle_dict(c, tables, x=X))
    rel = np.linalg.norm(G) / np.linalg.norm(G0)
    t1 = time.time()
    x_o, it, hist = O.newton_solve_pcg(O.make_cell(oracle_dict(c, tables, x=X)), X, rtol=1e-11, atol=0.0, maxit=30)
    err = np.max(np.abs(x - x_o)) / np.max(np.abs(x_o - X))
    log("family 2: %d Newton iterations, ||G(x_gpu)|| / ||G0|| = %.2e; oracle PCG Newton %d iterations (%.1f s), "
        "|x - x_oracle| / |u| = %.2e" % (st["iterations"], rel, it, time.time() - t1, err))
    assert rel <= 1e-8
    assert it >= 0
    assert err <= 1e-8


def test_family3_full_batch_one_load_step():
    cells, tables = configs.batched_cells(256)
    ntri = sum(len(c["tris"]) for c in cells)
    log("case: family 3, %d cells, %d triangles" % (len(cells), ntri))
    ctx = W.Context(precision=W.PREC_MIXED, t0=T0, dt=DT, newton_rtol=1e-11, newton_atol=0.0, max_refine=12)
    configs.load(ctx, cells, tables)
    # cells that converge first start their next load step while the others
    # iterate: each cell's positions are read right after the Newton iteration
    # that committed its first load step (the solution at load time T0)
    xs = [None] * len(cells)
    iters, lin = 0, 0.0
    while any(x is None for x in xs):
        st = ctx.newton_step()
        iters += 1
        lin = max(lin, st["max_lin_relres"])
        assert iters <= 30
        for i in range(len(cells)):
            if xs[i] is None:
                s = ctx.cell_info(i)["step"]
                if s:
                    assert s == 1
                    xs[i] = ctx.positions(i)
    ts = [T0] * len(cells)
    ctx.close()
    assert lin <= 1e-8
    st = dict(iterations=iters)
    worst = 0.0
    for i, c in enumerate(cells):
        G = residual(oracle_dict(c, tables, x=xs[i], time=ts[i]))
        G0 = residual(oracle_dict(c, tables, time=ts[i]))
        rel = np.linalg.norm(G) / np.linalg.norm(G0)
        worst = max(worst, rel)
        assert rel <= 1e-8, (i, rel)
    # full oracle solves on 16 cells spread over the size range
    order = sorted(range(len(cells)), key=lambda i: len(cells[i]["tris"]))
    pick = [order[int(round(k * (len(order) - 1) / 15))] for k in range(16)]
    worst_x = 0.0
    for i in pick:
        c = cells[i]
        X = np.asarray(c["X"]).reshape(-1)
        x_o, it, _ = O.newton_solve_pcg(O.make_cell(oracle_dict(c, tables, time=ts[i])), X, rtol=1e-11, atol=0.0,
                                        maxit=30)
        assert it >= 0, i
        err = np.max(np.abs(xs[i] - x_o)) / np.max(np.abs(x_o - X))
        worst_x = max(worst_x, err)
        assert err <= 1e-8, (i, len(c["tris"]), err)
    log("family 3: %d Newton iterations; worst ||G(x_gpu)|| / ||G0|| over 256 cells %.2e; worst |x - x_oracle| / |u| "
        "over 16 cells (%d..%d triangles) %.2e" % (st["iterations"], worst, len(cells[pick[0]]["tris"]),
                                                     len(cells[pick[-1]]["tris"]), worst_x))


def test_family4_500k_csr_residual():
    cells, tables = configs.workload("single500k")
    c = cells[0]
    log("case: family 4, %d nodes, %d triangles, %d table springs" % (len(c["X"]), len(c["tris"]),
                                                                     len(c["spr_tri"])))
    ctx = W.Context(precision=W.PREC_MIXED, t0=T0, dt=DT, newton_rtol=1e-10, newton_atol=0.0, max_refine=12)
    configs.load(ctx, cells, tables)
    st = ctx.load_step()
    x = ctx.positions(0)
    ctx.close()
    assert st["max_lin_relres"] <= 1e-8
    G = residual(oracle_dict(c, tables, x=x))
    G0 = residual(oracle_dict(c, tables))
    rel = np.linalg.norm(G) / np.linalg.norm(G0)
    log("family 4: %d Newton iterations, PCG iterations per sweep up to %d, ||G(x_gpu)|| / ||G0|| = %.2e"
        % (st["iterations"], st["max_cg_iters"], rel))
    assert rel <= 1e-8


def test_family5_bell_replicas_50_steps():
    cells, tables = configs.replicas(16)
    dt, k_on, k_off0, f_b, key = 0.05, 1.0, 0.1, 4.0, (5, 0)
    log("case: family 5, %d replicas, %d..%d triangles, 50 Bell load steps" % (
        len(cells), min(len(c["tris"]) for c in cells), max(len(c["tris"]) for c in cells)))
    ctx = W.Context(precision=W.PREC_MIXED, t0=T0, dt=dt, newton_rtol=1e-10, newton_atol=1e-12, max_refine=12,
                    seed_lo=key[0], seed_hi=key[1], k_on=k_on, k_off0=k_off0, f_bell=f_b)
    configs.load(ctx, cells, tables)
    transitions, margin = 0, 1.0
    before = [ctx.springs(i) for i in range(len(cells))]
    steps = [0] * len(cells)
    # a cell that converges first starts its next load step while the others
    # iterate: each cell is checked right after the Newton iteration that
    # committed its step (its state then is that step's)
    it = 0
    while min(steps) < 50:
        ctx.newton_step()
        it += 1
        assert it < 50 * 20
        for i, c in enumerate(cells):
            s = ctx.cell_info(i)["step"]
            if s == steps[i]:
                continue
            assert s == steps[i] + 1
            step = steps[i]
            st0, an0 = before[i]
            x = ctx.positions(i)
            st1, an1 = ctx.springs(i)
            d = oracle_dict(c, tables, x=x, time=T0 + step * dt, state=st0, anchor=an0)
            st_o, an_o, nt, mm = O.bell_update(O.make_cell(d), st0, an0, dt, k_on, k_off0, f_b, key, i, step)
            assert np.array_equal(st1, st_o), (step, i)
            assert np.array_equal(an1, an_o), (step, i)
            transitions += nt
            margin = min(margin, mm)
            before[i] = (st1, an1)
            steps[i] = s
    ctx.close()
    log("family 5: %d transitions over 50 steps x 16 replicas, min |u - p| = %.3e" % (transitions, margin))
    assert transitions > 0
    # no decision within rounding distance of its probability
    assert margin > 1e-13
