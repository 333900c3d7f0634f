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


def test_pcg_lanes_agree_and_are_deterministic():
    """Lanes per block row of the SpMV (4, 8, 16): a batch of two sparse cells
    of several 256-row chunks each and a dense cell between them reaches the
    oracle's FP64 Newton solution with every lane count; a repeated run is
    bit for bit the same.  The per-cell scalars come from each cell's own
    chunk partials, so a wrong chunk range or parity shows as a wrong
    solution of one of the two sparse cells."""
    ds = [cells.cst_cell(K=k, n_spr=12 * k, law_table=(k % 2 == 0), seed=70 + k, time=1.1) for k in (14, 5, 18)]
    solvers = [W.SOLVER_CSR, W.SOLVER_DENSE, W.SOLVER_CSR]
    ref = [O.newton_solve_pcg(O.make_cell(d), d["x"], rtol=1e-12)[0] for d in ds]
    for lanes in (4, 8, 16):
        runs = []
        for _ in range(2):
            c = _ctx(ds, solvers, precision=W.PREC_MIXED, t0=1.1, newton_rtol=1e-11, newton_atol=1e-11,
                     max_refine=12, pcg_lanes=lanes)
            # each cell's positions right after the iteration that commits its
            # first load step (a cell that converges first goes on to the next)
            xs = [None] * len(ds)
            for _ in range(20):
                st = c.newton_step()
                assert st["max_lin_relres"] <= 1e-8 and st["n_cg_breakdown"] == 0
                for i in range(len(ds)):
                    if xs[i] is None and c.cell_info(i)["step"] >= 1:
                        xs[i] = c.positions(i)
                if all(x is not None for x in xs):
                    break
            c.close()
            assert all(x is not None for x in xs)
            runs.append(xs)
        for a, b in zip(runs[0], runs[1]):
            assert np.array_equal(a, b), lanes
        for d, x, xo in zip(ds, runs[0], ref):
            err = np.max(np.abs(x - xo)) / np.max(np.abs(xo - d["X"]))
            print("lanes %d, cell of %d triangles: |x - x_oracle| / |u| = %.2e" % (lanes, len(d["tris"]) // 3, err),
                  flush=True)
            assert err <= 1e-8


def test_autotune_picks_pcg_lanes():
    ds = [cells.cst_cell(K=16, n_spr=200, law_table=True, seed=81, time=1.0)]
    c = _ctx(ds, [W.SOLVER_CSR], precision=W.PREC_MIXED, t0=1.0, newton_rtol=1e-11, newton_atol=1e-11, max_refine=12)
    ch = c.autotune(flags=W.TUNE_PCG, steps=1)
    assert ch["pcg_lanes"] in (4, 8, 16) and ch["ms_per_step"] > 0
    st = c.load_step()
    x = c.positions(0)
    c.close()
    assert st["max_lin_relres"] <= 1e-8
    xo = O.newton_solve_pcg(O.make_cell(ds[0]), ds[0]["x"], rtol=1e-12)[0]
    assert np.max(np.abs(x - xo)) <= 1e-8 * np.max(np.abs(xo - ds[0]["X"]))


def test_pcg_lanes_out_of_range_is_an_argument_error():
    d = cells.cst_cell(K=4, n_spr=20, seed=3, time=1.0)
    with pytest.raises(W.WcfeError) as e:
        c = _ctx([d], [W.SOLVER_CSR], pcg_lanes=5)
        c.close()
    assert e.value.code == W.ERR_ARG
