"""Two-chain band LU (wcfe_options.twist): cells whose chain of block steps
is longer than the batch needs are factored top down and bottom up at once,
meeting in a middle window of bw blocks; the triangular sweeps follow the
same two chains.  Checked against the one-chain factorization, the
full-matrix LU and the oracle; the fused step forms stay bit for bit equal."""
import numpy as np
import pytest

import cells
import oracle_py as O
import wholecellfe_amd as W

pytestmark = pytest.mark.gpu

KS = (3, 9, 16, 22, 27)


def _batch():
    return [cells.cst_cell(K=k, n_spr=20 + 10 * k, law_table=(k % 2 == 1), seed=40 + k, time=1.0) for k in KS]


def _ctx(ds, **opt):
    c = W.Context(precision=W.PREC_MIXED, t0=1.0, **opt)
    for d in ds:
        cells.add_to_context(c, d)
    c.finalize()
    return c


def test_layout_splits_only_the_long_cells():
    ds = _batch()
    c = _ctx(ds)
    lay = [c.cell_layout(i) for i in range(len(ds))]
    c.close()
    print("layouts:", [(l["nblk"], l["bw"], l["chain_top"], l["chain_bottom"]) for l in lay], flush=True)
    chain = max(l["chain_top"] + l["bw"] if l["chain_top"] else l["nblk"] for l in lay)
    for l in lay:
        if l["chain_top"]:
            m, nb, n, bw = l["chain_top"], l["chain_bottom"], l["nblk"], l["bw"]
            assert m + bw + nb == n and 1 <= nb <= m - bw
            assert n > chain
        else:
            assert l["nblk"] <= chain
    assert lay[-1]["chain_top"] > 0 and lay[0]["chain_top"] == 0
    c = _ctx(ds, twist=0)
    assert all(c.cell_layout(i)["chain_top"] == 0 for i in range(len(ds)))
    c.close()


@pytest.mark.parametrize("sym", [1, 0])
def test_two_chains_match_one_chain(sym):
    """One batched Newton step with and without the split: the same update to
    the refinement tolerance; each split cell alone through a whole load
    step against the oracle's FP64 Newton solve (from the rest positions: with
    fixed anchors the equilibrium at the load time does not depend on the
    start)."""
    ds = _batch()
    res = {}
    for tw in (1, 0):
        c = _ctx(ds, twist=tw, symmetric=sym, max_refine=12)
        st = c.newton_step()
        assert st["max_lin_relres"] <= 1e-8
        res[tw] = [c.positions(i) for i in range(len(ds))]
        c.close()
    for d, a, b in zip(ds, res[1], res[0]):
        assert np.max(np.abs(a - b)) <= 1e-7 * np.max(np.abs(b - d["x"]))
    for d in ds[-2:]:
        c = _ctx([d], symmetric=sym, newton_rtol=1e-11, newton_atol=0.0, max_refine=12)
        assert c.cell_layout(0)["chain_top"] > 0
        st = c.load_step()
        x = c.positions(0)
        c.close()
        assert st["max_lin_relres"] <= 1e-8
        # the oracle's Jacobi-PCG Newton (pinned against its dense Newton)
        x_o, it, _ = O.newton_solve_pcg(O.make_cell(d), d["X"], rtol=1e-11)
        assert it >= 0
        err = np.max(np.abs(x - x_o)) / np.max(np.abs(x_o - d["X"]))
        print("split cell of %d triangles: |x - x_oracle| / |u| = %.2e" % (len(d["tris"]) // 3, err), flush=True)
        assert err <= 1e-8


def test_two_chain_first_solve_accuracy():
    """One Newton iteration with one solve sweep (max_refine = 1): the split
    cells' FP32 factors and sweeps give a first solve as accurate as the
    one-chain factorization's (a wrong block, inverse or sweep term leaves an
    O(1) residual)."""
    ds = _batch()
    deltas = {}
    for tw in (1, 0):
        c = _ctx(ds, twist=tw, max_refine=1)
        st = c.newton_step(stats=True, check=False)
        assert st["error"] in (W.OK, W.ERR_NOT_CONVERGED)
        deltas[tw] = [c.positions(i) - d["x"] for i, d in enumerate(ds)]
        c.close()
    for i, d in enumerate(ds):
        G, K = O.assemble(O.make_cell(d))
        nb = np.linalg.norm(G)
        r = {tw: np.linalg.norm(-G - K @ deltas[tw][i]) / nb for tw in (1, 0)}
        print("cell %d: first-solve relres split %.2e, one chain %.2e" % (i, r[1], r[0]), flush=True)
        assert r[1] <= 10.0 * r[0] + 1e-6
        assert r[1] < 1e-3


def test_fused_forms_agree_bit_for_bit_with_two_chains():
    """L items, panel items, split panel items, the per-launch choice and one
    workgroup per cell run the same products in the same order on the split
    cells too."""
    ds = _batch()
    ref = None
    for fused in (1, 2, 3, 4, 5):
        c = _ctx(ds, fused_step=fused, max_refine=3)
        assert c.cell_layout(len(ds) - 1)["chain_top"] > 0
        c.newton_step(stats=True, check=False)
        xs = [c.positions(i) for i in range(len(ds))]
        c.close()
        if ref is None:
            ref = xs
        else:
            for a, b in zip(xs, ref):
                assert np.array_equal(a, b), fused
