"""Sparse path (block-CSR assembly + block-Jacobi PCG inside the FP64
refinement) against the oracle and against the dense path, on the GPU."""
import numpy as np
import pytest

import cells
import oracle_py as O
import wholecellfe_amd as W

pytestmark = pytest.mark.gpu


def _ctx(ds, solvers, **opt):
    c = W.Context(**opt)
    for d, s in zip(ds, solvers):
        cells.add_to_context(c, d, solver=s)
    c.finalize()
    return c


@pytest.mark.parametrize("prec", [W.PREC_MIXED, W.PREC_FP64])
@pytest.mark.parametrize("table", [False, True])
def test_csr_assembled_system_equals_dense(prec, table):
    """The block-CSR values, densified, equal the dense assembly bit for bit
    (same element systems, same colour-ordered sums)."""
    print("case prec", prec, "table", table, flush=True)
    d = cells.cst_cell(K=6, n_spr=50, law_table=table, seed=5, time=1.0)
    d["x"] = d["X"] + 0.002 * np.random.default_rng(3).standard_normal(len(d["X"]))
    out = []
    for solver in (W.SOLVER_DENSE, W.SOLVER_CSR):
        c = _ctx([d], [solver], precision=prec, t0=1.0)
        out.append(c.assembled_system(0))
        c.close()
    (Kd, Gd), (Ks, Gs) = out
    assert np.array_equal(Gd, Gs)
    assert np.array_equal(Kd, Ks)


@pytest.mark.parametrize("prec", [W.PREC_MIXED, W.PREC_FP64])
@pytest.mark.parametrize("table", [False, True])
def test_csr_newton_load_step_vs_oracle(prec, table):
    print("case prec", prec, "table", table, flush=True)
    d = cells.cst_cell(K=9, n_spr=120, law_table=table, time=1.2, seed=23)
    c = _ctx([d], [W.SOLVER_CSR], precision=prec, t0=d["time"], newton_rtol=1e-10, newton_atol=1e-10, max_refine=12)
    st = c.load_step()
    x_gpu = c.positions(0)
    c.close()
    assert st["max_lin_relres"] <= 1e-8
    assert st["max_cg_iters"] > 0
    assert st["n_cg_breakdown"] == 0
    x_o, it, _ = O.newton_solve(O.make_cell(d), d["x"], rtol=1e-12, atol=1e-12, maxit=20)
    assert it >= 0
    assert np.max(np.abs(x_gpu - x_o)) <= 1e-8 * np.max(np.abs(x_o - d["X"]))


def test_mixed_dense_and_csr_batch_matches_all_dense():
    """A batch whose middle cell is sparse: every cell ends its load step where
    the all-dense batch does (to far below the Newton tolerance)."""
    ds = [cells.cst_cell(K=k, n_spr=20 + 10 * k, law_table=(k % 2 == 0), seed=k, time=1.0) for k in (4, 8, 6)]
    res = []
    for solvers in ([W.SOLVER_DENSE] * 3, [W.SOLVER_DENSE, W.SOLVER_CSR, W.SOLVER_DENSE]):
        c = _ctx(ds, solvers, precision=W.PREC_MIXED, t0=1.0, newton_rtol=1e-11, newton_atol=1e-11, max_refine=12)
        st = c.load_step()
        assert st["max_lin_relres"] <= 1e-8
        res.append([c.positions(i) for i in range(3)])
        c.close()
    for d, a, b in zip(ds, res[0], res[1]):
        assert np.max(np.abs(a - b)) <= 1e-9 * np.max(np.abs(a - d["X"]))


def test_csr_pcg_deterministic_and_auto_threshold():
    """Two identical runs give identical positions (fixed-order reductions);
    dense_max_dof routes the larger cell of an AUTO batch to the sparse path."""
    ds = [cells.cst_cell(K=4, n_spr=30, seed=1, time=1.0), cells.cst_cell(K=10, n_spr=90, seed=2, time=1.0)]
    runs = []
    for _ in range(2):
        c = W.Context(precision=W.PREC_MIXED, t0=1.0, dense_max_dof=200)
        for d in ds:
            cells.add_to_context(c, d)
        c.finalize()
        for _ in range(3):
            st = c.newton_step()
        runs.append([c.positions(i) for i in range(2)])
        c.close()
    assert st["max_cg_iters"] > 0  # the K=10 cell (> 200 unknowns) went through PCG
    for a, b in zip(runs[0], runs[1]):
        assert np.array_equal(a, b)
