"""Pins for the oracle's sparse forms (used to check the configurations too
large for a dense FP64 solve): the CSR assembly equals the dense assembly,
the PCG Newton solve equals the dense Newton solve, the band LU step equals a
dense LU step, and the reverse Cuthill-McKee numbering is a permutation that
narrows the band."""
import numpy as np
import pytest

import cells
import oracle_py as O
from wholecellfe_amd import mesh


def _dense(rowptr, col, val, n):
    K = np.zeros((n, n))
    for r in range(n):
        K[r, col[rowptr[r]:rowptr[r + 1]]] = val[rowptr[r]:rowptr[r + 1]]
    return K


@pytest.mark.parametrize("table", [False, True])
def test_csr_assembly_equals_dense(table):
    d = cells.cst_cell(K=5, n_spr=60, law_table=table, seed=4, time=1.1)
    d["x"] = d["X"] + 0.002 * np.random.default_rng(1).standard_normal(len(d["X"]))
    c = O.make_cell(d)
    Gd, Kd = O.assemble(c)
    rowptr, col = O.csr_pattern(c.n_nodes, d["tris"])
    G, val = O.assemble_csr(c, rowptr, col)
    assert np.array_equal(G, Gd)
    K = _dense(rowptr, col, val, len(G))
    assert np.array_equal(K, Kd)
    # every nonzero of the dense tangent lies in the pattern
    mask = np.zeros_like(Kd, dtype=bool)
    for r in range(len(G)):
        mask[r, col[rowptr[r]:rowptr[r + 1]]] = True
    assert not np.any(Kd[~mask])


def test_pcg_newton_equals_dense_newton():
    d = cells.cst_cell(K=6, n_spr=80, law_table=True, seed=6, time=1.4)
    xd, itd, _ = O.newton_solve(O.make_cell(d), d["x"], rtol=1e-11, atol=0.0, maxit=20)
    xp, itp, hist = O.newton_solve_pcg(O.make_cell(d), d["x"], rtol=1e-11, atol=0.0, maxit=20)
    assert itd >= 0 and itp >= 0
    u = np.max(np.abs(xd - d["X"]))
    assert np.max(np.abs(xp - xd)) <= 1e-11 * u


def test_band_newton_step_equals_dense_step():
    for K, table in ((4, False), (7, True)):
        d = cells.cst_cell(K=K, n_spr=12 * K, law_table=table, seed=K, time=0.9)
        c = O.make_cell(d)
        G, Kd = O.assemble(c)
        x_dense = d["x"] + np.linalg.solve(Kd, -G)
        x_band, rc = O.band_newton_step(c, d["x"])
        assert rc == 0
        assert np.max(np.abs(x_band - x_dense)) <= 1e-12 * np.max(np.abs(x_dense - d["x"]))


def test_rcm_order_narrows_the_band():
    rng = np.random.default_rng(0)
    X, T, b = mesh.disk_rings(12, 10.0)
    # scramble the node numbering: the natural ring order is already banded
    perm = rng.permutation(len(X))
    inv = np.argsort(perm)
    Ts = inv[T]
    w, order = O.band_width(len(X), Ts)
    assert sorted(order.tolist()) == list(range(len(X)))
    nat = max(abs(int(a) - int(c)) for t in Ts for a in t for c in t)
    assert w < nat / 5
    # a disk of K rings needs about one ring's worth of nodes, at most a couple of rings
    assert w <= 2 * 6 * 12 + 6


def test_batch_band_step_matches_single_steps():
    ds = [cells.cst_cell(K=k, n_spr=8 * k, law_table=(k % 2 == 0), seed=k, time=1.2) for k in (3, 5, 8)]
    cs = [O.make_cell(d) for d in ds]
    xs = [d["x"].copy() for d in ds]
    t, err = O.batch_band_newton_step(cs, xs, n_threads=2)
    assert err == 0 and t > 0
    for d, c, x in zip(ds, cs, xs):
        x1, rc = O.band_newton_step(O.make_cell(d), d["x"])
        assert rc == 0 and np.array_equal(x, x1)
