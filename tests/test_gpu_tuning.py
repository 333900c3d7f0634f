"""Tunable LU schedule (delayed-update depth, concurrent cell groups), the
autotuner and the state restore, on the GPU."""
import numpy as np
import pytest

import cells
import wholecellfe_amd as W

pytestmark = pytest.mark.gpu


def _batch():
    return [cells.cst_cell(K=k, n_spr=20 + 10 * k, law_table=(k % 2 == 1), seed=k, time=1.0) for k in (3, 7, 13, 10)]


def _run(ds, **opt):
    c = W.Context(precision=W.PREC_MIXED, t0=1.0, newton_rtol=1e-11, newton_atol=1e-11, max_refine=12, **opt)
    for d in ds:
        cells.add_to_context(c, d)
    c.finalize()
    st = c.load_step()
    xs = [c.positions(i) for i in range(len(ds))]
    c.close()
    assert st["max_lin_relres"] <= 1e-8
    return xs


@pytest.mark.parametrize("sym,depth,fused", [(1, 4, 1), (0, 4, 1), (1, 1, 1), (0, 1, 1), (1, 1, 0), (0, 1, 0),
                                             (0, 2, 1), (1, 3, 1), (1, 1, 2), (0, 1, 2), (1, 1, 3), (0, 1, 3),
                                             (1, 1, 4), (0, 1, 4)])
def test_band_lu_matches_full_matrix(sym, depth, fused):
    """Band LU in the reverse Cuthill-McKee numbering (the default) against the
    full-matrix LU in the natural numbering, for symmetric and full
    factorisations and delayed-update depths below and above the block
    bandwidth: one batched Newton step (every cell's first, so all are in the
    same phase) gives the same update to the refinement tolerance, and a
    whole load step of each cell alone the same converged positions.  (In a
    batch, cells that converge early start their next load step while the
    others iterate, so batch-final positions depend on the iteration counts.)"""
    print("case sym", sym, "depth", depth, "fused", fused, flush=True)
    ds = [cells.cst_cell(K=k, n_spr=20 + 10 * k, law_table=(k % 2 == 1), seed=k, time=1.0) for k in (3, 9, 16, 22)]
    steps = []
    for band in (1, 0):
        c = W.Context(precision=W.PREC_MIXED, t0=1.0, max_refine=12, symmetric=sym, lu_depth=depth, band=band,
                      fused_step=fused)
        for d in ds:
            cells.add_to_context(c, d)
        c.finalize()
        st = c.newton_step()
        assert st["max_lin_relres"] <= 1e-8
        steps.append([c.positions(i) for i in range(len(ds))])
        c.close()
    for d, a, b in zip(ds, steps[0], steps[1]):
        assert np.max(np.abs(a - b)) <= 1e-7 * np.max(np.abs(b - d["x"]))
    for d in ds[2:]:
        band = _run([d], symmetric=sym, lu_depth=depth, fused_step=fused)[0]
        full = _run([d], symmetric=sym, lu_depth=depth, band=0)[0]
        assert np.max(np.abs(band - full)) <= 1e-9 * np.max(np.abs(full - d["X"]))


@pytest.mark.parametrize("sym", [1, 0])
def test_fused_step_forms_agree_bit_for_bit(sym):
    """The fused block step's forms -- L items that each invert the diagonal
    block, one panel item per cell that inverts it once and solves every L
    block, and the whole factorization in one launch with each cell's steps
    in one workgroup (the first step in L items, the others in panel form) --
    run the same products in the same order: one Newton step gives identical
    positions and statistics."""
    print("case sym", sym, flush=True)
    ds = [cells.cst_cell(K=k, n_spr=20 + 10 * k, law_table=(k % 2 == 1), seed=k, time=1.0) for k in (1, 2, 9, 16, 22)]
    out = []
    for fused in (2, 3, 4):
        c = W.Context(precision=W.PREC_MIXED, t0=1.0, max_refine=12, symmetric=sym, lu_depth=1, fused_step=fused)
        for d in ds:
            cells.add_to_context(c, d)
        c.finalize()
        st = c.newton_step()
        assert c.schedule()["fused_step"] == (2 if fused == 4 else 1)
        out.append(([c.positions(i) for i in range(len(ds))], st["max_lin_relres"]))
        c.close()
    for o in out[1:]:
        for a, b in zip(out[0][0], o[0]):
            assert np.array_equal(a, b)
        assert out[0][1] == o[1]


def test_band_assembled_system_is_natural_order():
    """The assembled system read back is in the natural numbering whatever
    numbering the band LU uses."""
    d = cells.cst_cell(K=16, n_spr=60, law_table=True, seed=3, time=1.0)
    d["x"] = d["X"] + 0.002 * np.random.default_rng(3).standard_normal(len(d["X"]))
    out = []
    for band in (1, 0):
        c = W.Context(precision=W.PREC_FP64, t0=1.0, band=band)
        cells.add_to_context(c, d)
        c.finalize()
        out.append(c.assembled_system(0))
        c.close()
    assert np.array_equal(out[0][1], out[1][1])
    assert np.max(np.abs(out[0][0] - out[1][0])) <= 1e-12 * np.max(np.abs(out[1][0]))


@pytest.mark.parametrize("depth,groups", [(1, 1), (2, 3), (3, 2), (6, 4), (8, 1)])
def test_lu_depth_and_groups_same_solution(depth, groups):
    """Every delayed-update depth and group count factors the same systems:
    the converged positions agree with the default schedule far below the
    Newton tolerance."""
    print("case depth", depth, "groups", groups, flush=True)
    ds = _batch()
    ref = _run(ds)
    got = _run(ds, lu_depth=depth, lu_groups=groups)
    for d, a, b in zip(ds, ref, got):
        assert np.max(np.abs(a - b)) <= 1e-9 * np.max(np.abs(a - d["X"]))


def test_autotune_picks_valid_settings_and_keeps_results():
    ds = _batch()
    ref = _run(ds)
    c = W.Context(precision=W.PREC_MIXED, t0=1.0, newton_rtol=1e-11, newton_atol=1e-11, max_refine=12)
    for d in ds:
        cells.add_to_context(c, d)
    c.finalize()
    ch = c.autotune(steps=1)
    assert ch["lu_depth"] in (1, 2, 4, 6) and ch["lu_groups"] in (1, 2, 4) and ch["ms_per_step"] > 0
    assert ch["fused_step"] in (0, 1, 4, 5)
    # the reported two-chain count is the batch's layout after the tuning
    assert ch["two_chain_cells"] == sum(c.cell_layout(i)["chain_top"] > 0 for i in range(len(ds)))
    # the batch restarts from the state the cells were added with
    st = c.load_step()
    assert st["max_lin_relres"] <= 1e-8
    for i, d in enumerate(ds):
        x = c.positions(i)
        assert np.max(np.abs(x - ref[i])) <= 1e-9 * np.max(np.abs(ref[i] - d["X"]))
    with pytest.raises(W.WcfeError) as e:
        c.autotune(steps=1)
    assert e.value.code == W.ERR_STATE
    c.close()


def test_set_ul_state_redoes_a_load_step_bitwise():
    """Reference law + node springs + mean-field kinetics: save the state after
    one load step, take another, restore, take it again: identical results."""
    d = cells.ul_cell(K=6, table=False, time=1.3, deform=0.0)
    d["x"] = d["X"].copy()
    c = W.Context(precision=W.PREC_MIXED, t0=d["time"], newton_rtol=1e-10, newton_atol=1e-9)
    cells.add_to_context(c, d, kinetics=W.KIN_MEANFIELD)
    c.finalize()
    c.load_step()
    X, ea, cC, fau = c.ul_state(0)
    x1 = c.positions(0)
    t1 = c.cell_info(0)["time"]
    c.load_step()
    x2 = c.positions(0)
    s2 = c.ul_state(0)
    c.set_ul_state(0, X=X, ea=ea, cC=cC, fau=fau, time=t1)
    c.set_positions(0, x1)
    c.load_step()
    assert np.array_equal(c.positions(0), x2)
    for a, b in zip(c.ul_state(0), s2):
        assert np.array_equal(a, b)
    c.close()


def _kw(d, kinetics=0):
    return dict(x=d.get("x"), law=d["law"], mu=d.get("mu", 1000.0), E=d.get("E", 1000.0), nu=d.get("nu", 0.3),
                thick=d.get("thick", 0.2), ea=d.get("ea"), node_springs=d.get("node_springs", 0), cC=d.get("cC"),
                fau=d.get("fau0"), rho_max=d.get("rho_max", 100.0), spr_tri=d.get("spr_tri"),
                spr_bary=d.get("spr_bary"), spr_anchor=d.get("spr_anchor"), spr_state=d.get("spr_state"),
                spring_law=d.get("spring_law", 0), k_lin=d.get("k_lin", 1000.0), kinetics=kinetics)


def test_replace_cell_keeps_the_other_cells():
    """Remesh one cell of a batch mid-run: the other cells go on exactly as in
    a batch without the replaced cell (bitwise, after the same number of
    Newton iterations), and the new cell converges."""
    ds = _batch()[:3]
    new = cells.cst_cell(K=9, n_spr=70, seed=42, time=1.0)
    opt = dict(precision=W.PREC_MIXED, t0=1.0, newton_rtol=1e-10, newton_atol=1e-9)
    c = W.Context(**opt)
    for d in ds:
        cells.add_to_context(c, d)
    c.finalize()
    for _ in range(4):
        c.newton_step()
    step1 = c.cell_info(1)["step"]
    c.replace_cell(1, new["X"], new["tris"], **_kw(new))
    for _ in range(5):
        st = c.newton_step()
    assert st["max_lin_relres"] <= 1e-8
    assert c.cell_info(1)["step"] > step1  # the new mesh committed load steps
    got = [c.positions(i) for i in (0, 2)]
    c.close()
    r = W.Context(**opt)
    for d in (ds[0], ds[2]):
        cells.add_to_context(r, d)
    r.finalize()
    for _ in range(9):
        r.newton_step()
    for i, g in enumerate(got):
        assert np.array_equal(r.positions(i), g)
    r.close()


def test_checkpoint_round_trip_bitwise(tmp_path):
    """Save after one load step, continue; reload into a fresh context and
    continue: identical positions, strains and bond states (Bell stream)."""
    d0 = cells.cst_cell(K=6, n_spr=60, seed=3, time=1.0)
    d1 = cells.ul_cell(K=5, table=False, time=1.3, deform=0.0)
    d1["x"] = d1["X"].copy()
    path = tmp_path / "state.ckp"
    c = W.Context(precision=W.PREC_MIXED, t0=1.0, newton_rtol=1e-10, newton_atol=1e-9, seed_lo=11)
    cells.add_to_context(c, d0, kinetics=W.KIN_BELL)
    cells.add_to_context(c, d1, kinetics=W.KIN_MEANFIELD)
    c.finalize()
    c.load_step()
    c.checkpoint_save(path)
    steps = [c.cell_info(i)["step"] for i in range(2)]
    sizes = [(x["n_nodes"], x["n_tris"], x["n_spr"]) for x in c.cells]
    c.load_step()
    ref = [(c.positions(i), c.ul_state(i), c.springs(i)) for i in range(2)]
    c.close()
    r = W.Context(precision=W.PREC_MIXED, t0=1.0, newton_rtol=1e-10, newton_atol=1e-9, seed_lo=11)
    r.checkpoint_load(path)
    assert [(x["n_nodes"], x["n_tris"], x["n_spr"]) for x in r.cells] == sizes
    r.finalize()
    assert [r.cell_info(i)["step"] for i in range(2)] == steps
    r.load_step()
    for i in range(2):
        x, ul, sp = r.positions(i), r.ul_state(i), r.springs(i)
        assert np.array_equal(x, ref[i][0])
        for a, b in zip(ul, ref[i][1]):
            assert np.array_equal(a, b)
        for a, b in zip(sp, ref[i][2]):
            assert np.array_equal(a, b)
    r.close()
