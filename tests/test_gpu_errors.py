"""Failure modes reported as error codes, not NaN: a linear solve that cannot
meet lin_rtol, an exactly singular column under partial pivoting, checkpoint
files the library must refuse; and the first solve of the fused band-LU block
steps (whose forward sweep runs inside the factorization) against the
unfused factorization."""
import numpy as np
import pytest

import cells
import oracle_py as O
import wholecellfe_amd as W

pytestmark = pytest.mark.gpu


def test_inconsistent_singular_cell_is_not_converged():
    """A cell with no springs has rigid-body modes; a net external force is
    outside the tangent's range, so no refinement can meet lin_rtol: the step
    returns NOT_CONVERGED with the statistics filled in and finite positions."""
    d = cells.cst_cell(K=4, n_spr=0, seed=2, time=1.0)
    for k in ("spr_tri", "spr_bary", "spr_anchor", "spr_state"):
        d.pop(k)
    N = len(d["X"]) // 2
    f_ext = np.tile([5.0, 0.0], N)  # 5 pN on every node along x: no equilibrium
    print("case: spring-free cell, %d nodes, net force %.1f pN" % (N, 5.0 * N), flush=True)
    g = W.Context(precision=W.PREC_MIXED, t0=1.0, max_refine=6)
    g.add_cell(X=d["X"], tris=d["tris"], law=W.LAW_CST, E=1000.0, nu=0.3, thick=0.2, f_ext=f_ext)
    g.finalize()
    st = g.newton_step(stats=True, check=False)
    x = g.positions(0)
    g.close()
    assert st["error"] == W.ERR_NOT_CONVERGED
    assert st["n_failed"] == 1
    assert st["max_lin_relres"] > 1e-8
    assert np.all(np.isfinite(x))
    # the same call with checking raises with that code
    g = W.Context(precision=W.PREC_MIXED, t0=1.0, max_refine=6)
    g.add_cell(X=d["X"], tris=d["tris"], law=W.LAW_CST, E=1000.0, nu=0.3, thick=0.2, f_ext=f_ext)
    g.finalize()
    with pytest.raises(W.WcfeError) as e:
        g.newton_step()
    g.close()
    assert e.value.code == W.ERR_NOT_CONVERGED
    assert "lin_rtol" in str(e.value)


def test_zero_column_partial_pivoting_is_zero_pivot():
    """A node that belongs to no triangle and holds no spring gives two exactly
    zero columns of K: partial pivoting reports ZERO_PIVOT."""
    d = cells.cst_cell(K=4, n_spr=20, seed=5, time=1.0)
    X = np.concatenate([d["X"], [30.0, 30.0]])  # isolated node, last
    print("case: isolated node, %d nodes, partial pivoting" % (len(X) // 2), flush=True)
    g = W.Context(precision=W.PREC_MIXED, pivoting=W.PIVOT_PARTIAL, t0=1.0)
    g.add_cell(X=X, tris=d["tris"], law=W.LAW_CST, E=1000.0, nu=0.3, thick=0.2, spr_tri=d["spr_tri"],
               spr_bary=d["spr_bary"], spr_anchor=d["spr_anchor"])
    g.finalize()
    with pytest.raises(W.WcfeError) as e:
        g.newton_step()
    g.close()
    assert e.value.code == W.ERR_ZERO_PIVOT


def _save(tmp_path):
    d = cells.cst_cell(K=3, n_spr=10, seed=1, time=1.0)
    path = tmp_path / "c.ckp"
    c = W.Context(precision=W.PREC_MIXED, t0=1.0)
    cells.add_to_context(c, d)
    c.finalize()
    c.checkpoint_save(path)
    c.close()
    return path, bytearray(path.read_bytes())


def test_checkpoint_refuses_bad_files(tmp_path):
    path, raw = _save(tmp_path)
    # header: magic 8, version 4, state record size 4, cells 4, tables 4 (no
    # table here); then cell 0: n_nodes, n_tris, law, ...
    assert int.from_bytes(raw[8:12], "little") == 4
    cases = {"version": (8, 3), "record size": (12, 7), "law": (32, 7), "n_tris": (28, 0)}
    for name, (off, val) in cases.items():
        bad = bytearray(raw)
        bad[off:off + 4] = int(val).to_bytes(4, "little", signed=True)
        p = tmp_path / ("bad_%s.ckp" % name.replace(" ", "_"))
        p.write_bytes(bytes(bad))
        r = W.Context(precision=W.PREC_MIXED)
        with pytest.raises(W.WcfeError) as e:
            r.checkpoint_load(p)
        assert e.value.code == W.ERR_ARG, name
        r.close()
    # truncated file
    p = tmp_path / "short.ckp"
    p.write_bytes(bytes(raw[:len(raw) // 2]))
    r = W.Context(precision=W.PREC_MIXED)
    with pytest.raises(W.WcfeError) as e:
        r.checkpoint_load(p)
    assert e.value.code == W.ERR_ARG
    r.close()
    # the intact file loads, with the library's sizes
    r = W.Context(precision=W.PREC_MIXED)
    r.checkpoint_load(path)
    assert len(r.cells) == 1 and r.cells[0]["n_tris"] == 6 * 9
    r.close()


def _first_solve(ds, fused, sym):
    """x after one Newton iteration with one solve sweep (max_refine = 1):
    x1 - x0 is the first solve's delta."""
    g = W.Context(precision=W.PREC_MIXED, t0=1.0, max_refine=1, symmetric=sym, fused_step=fused, lu_depth=1)
    for d in ds:
        cells.add_to_context(g, d)
    g.finalize()
    sched = g.schedule()
    st = g.newton_step(stats=True, check=False)
    xs = [g.positions(i) for i in range(len(ds))]
    g.close()
    assert st["error"] in (W.OK, W.ERR_NOT_CONVERGED)
    return xs, sched


@pytest.mark.parametrize("sym", [0, 1])
def test_fused_forward_block_first_solve(sym):
    """The first solve's forward sweep runs inside the fused block steps
    (every form: per-launch choice, L items, panel items, one workgroup per
    cell for the whole factorization, split panel items).  With one solve
    sweep and no refinement, its delta must be as accurate as the unfused
    factorization's.  The batch mixes one-block cells (no block step),
    two-block cells (a last step without trailing tiles) and wider cells, and
    is large enough (about 1,000 items in the first launches) for the
    per-launch choice to take the panel form."""
    ks = [1, 2, 3, 4, 5, 6, 8, 10, 12]
    ds = [cells.cst_cell(K=ks[i % len(ks)], n_spr=10 + 3 * ks[i % len(ks)] ** 2 // 2,
                         law_table=(i % 3 == 0), seed=100 + i, time=1.0) for i in range(270)]
    print("case: sym %d, %d cells" % (sym, len(ds)), flush=True)
    res = {}
    for fused in (0, 1, 2, 3, 4, 5):
        xs, sched = _first_solve(ds, fused, sym)
        assert sched["fused_step"] == (2 if fused == 4 else 1 if fused else 0)
        res[fused] = xs
    # every fused form runs the same products in the same order (the split
    # panel form of the large first launches included)
    for f in (2, 3, 4, 5):
        for a, b in zip(res[1], res[f]):
            assert np.array_equal(a, b), f
    worst = {f: 0.0 for f in (1, 2, 3, 4, 5)}
    for i, d in enumerate(ds):
        G, K = O.assemble(O.make_cell(d))
        nb = np.linalg.norm(G)
        if nb == 0.0:
            continue
        d_exact = np.linalg.solve(K, -G)
        r0 = np.linalg.norm(-G - K @ (res[0][i] - d["x"])) / nb
        e0 = np.linalg.norm(res[0][i] - d["x"] - d_exact) / np.linalg.norm(d_exact)
        for f in (1, 2, 3, 4, 5):
            df = res[f][i] - d["x"]
            rf = np.linalg.norm(-G - K @ df) / nb
            ef = np.linalg.norm(df - d_exact) / np.linalg.norm(d_exact)
            # FP32 first solves: same accuracy class as the unfused one; a
            # wrong forward block leaves an O(1) residual
            assert rf <= 10.0 * r0 + 1e-6, (i, f, rf, r0)
            assert ef <= 10.0 * e0 + 1e-6, (i, f, ef, e0)
            worst[f] = max(worst[f], rf)
    assert max(worst.values()) < 1e-3
