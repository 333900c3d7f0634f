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
    assert ch["lu_depth"] in (2, 4, 6) and ch["lu_groups"] in (1, 2, 4) and ch["ms_per_step"] > 0
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
