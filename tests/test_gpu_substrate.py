"""GPU parity of the deformable substrate: its law, the cell's integrin forces
on it, and the substrate as a passive cell of the implicit batch.

Reference (wholeCellFE_kMD/cellBP/UL_growth_explicit.cpp): the substrate's
updateForces_Explicit :429-504 (calculateRes with mat = 2 at :451, mu = 1e6
Pa :727 and t_act = 0 :763; calculateForce_FA_subs :1305-1475 at :467).

* Reference-shaped calls against the compiled reference's golden vectors:
  1e-10 (residual) and 1e-12 (lumped mass) of the array's largest entry.
* The substrate as a cell of the batch (passive law, constant nodal load,
  drag) against the oracle's planar Newton model: system 1e-10 relative (FP64
  tangent; the mixed form stores K in FP32: 1e-6); converged positions to
  1e-6 of the displacement (both solves stop at ||G|| <= 1e-11 ||G0||, and
  the softest mode, drag / dt = 1e4 pN/um, turns that into a bound 1e3 times
  tighter).
"""
import numpy as np
import pytest

import oracle_py as O
from conftest import golden

pytestmark = pytest.mark.gpu

W = pytest.importorskip("wholecellfe_amd")

DT = 0.01
DRAG = 100.0  # pN s/um: drag / dt = 1e4 pN/um holds the substrate's rigid-body modes


def relerr(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module")
def ctx():
    c = W.Context()
    yield c
    c.close()


def test_calculate_res_substrate_vs_reference(ctx):
    g = golden("calcres_subs")
    Res, Rea, MM = ctx.calculate_res(g["tris"], g["X"], g["x"], g["ea"], float(g["time"]), mat=2)
    assert relerr(Res, g["res_Res"]) < 1e-10
    assert relerr(Rea, g["res_Res_ea"]) < 1e-10
    assert relerr(MM, g["res_MM"]) < 1e-12


def test_calculate_res_rejects_unknown_material(ctx):
    g = golden("calcres_subs")
    with pytest.raises(W.WcfeError) as e:
        ctx.calculate_res(g["tris"], g["X"], g["x"], g["ea"], 1.0, mat=3)
    assert e.value.code == 1


def substrate_model(seed=0, scale=0.1):
    """The fixture's substrate (320 nodes) loaded by the cell's integrin
    forces (calculateForce_FA_subs of the kMD fixture, in-plane part, times
    `scale`: bound fractions a tenth of the fixture's keep the substrate's
    elements far from inversion)."""
    g = golden("forcesubs_kMD")
    SX = g["SX"][:, :2].copy()
    M = len(SX)
    F = (scale * g["F_subs"].reshape(-1, 3)[:, :2]).reshape(-1)
    rng = np.random.default_rng(seed)
    x0 = SX + 1e-3 * rng.standard_normal(SX.shape)
    return dict(X=SX.reshape(-1), x=x0.reshape(-1), tris=g["Stris"].reshape(-1).astype(np.int32), law=0,
                mu=1.0e6, thick=1.0, time=1.0, ea=np.zeros(3 * M), passive=1, f_ext=F, drag=DRAG, dt=DT)


def add_substrate(g, d, solver=0):
    return g.add_cell(X=d["X"], x=d["x"], tris=d["tris"], law=0, mu=d["mu"], thick=d["thick"], ea=d["ea"],
                      passive=d["passive"], f_ext=d["f_ext"], drag=d["drag"], solver=solver)


@pytest.mark.parametrize("prec,solver", [(W.PREC_FP64, W.SOLVER_DENSE), (W.PREC_MIXED, W.SOLVER_DENSE),
                                         (W.PREC_MIXED, W.SOLVER_CSR)])
def test_substrate_system_vs_oracle(prec, solver):
    print("case prec", prec, "solver", solver, flush=True)
    d = substrate_model()
    g = W.Context(precision=prec, t0=d["time"], dt=DT)
    add_substrate(g, d, solver)
    g.finalize()
    K, G = g.assembled_system(0)
    g.close()
    Go, Ko = O.assemble(O.make_cell(d))
    assert relerr(G, Go) < 1e-10
    assert relerr(K, Ko) < (1e-10 if prec == W.PREC_FP64 else 1e-6)
    # the passive law and the load are in effect: an active cell without the
    # load has another residual
    Ga, _ = O.assemble(O.make_cell(dict(d, passive=0)))
    Gn, _ = O.assemble(O.make_cell(dict(d, f_ext=None)))
    assert relerr(Ga, Go) > 1e-8 and relerr(Gn, Go) > 1e-3


@pytest.mark.parametrize("prec,solver", [(W.PREC_FP64, W.SOLVER_DENSE), (W.PREC_MIXED, W.SOLVER_DENSE),
                                         (W.PREC_MIXED, W.SOLVER_CSR)])
def test_substrate_load_step_vs_oracle(prec, solver):
    print("case prec", prec, "solver", solver, flush=True)
    d = substrate_model()
    g = W.Context(precision=prec, t0=d["time"], dt=DT, newton_rtol=1e-11, newton_atol=1e-6, max_refine=16)
    add_substrate(g, d, solver)
    g.finalize()
    st = g.load_step()
    x_gpu = g.positions(0)
    g.close()
    assert st["max_lin_relres"] <= 1e-8
    x_o, it, _ = O.newton_solve(O.make_cell(d), d["x"], rtol=1e-12, atol=1e-6, maxit=25)
    assert it >= 0
    u = np.max(np.abs(x_o - d["X"]))
    assert u > 1e-3
    assert np.max(np.abs(x_gpu - x_o)) <= 1e-6 * u


def test_set_external_forces_and_checkpoint(tmp_path):
    """A new load enters the next load step; a checkpoint keeps the passive
    flag and the load, and a reloaded batch resumes bit for bit."""
    d = substrate_model()
    g = W.Context(precision=W.PREC_MIXED, t0=1.0, dt=DT, newton_rtol=1e-11, newton_atol=1e-6, max_refine=16)
    add_substrate(g, d)
    g.finalize()
    g.load_step()
    X1, ea1, _, _ = g.ul_state(0)
    F2 = -0.5 * d["f_ext"]
    g.set_external_forces(0, F2)
    path = str(tmp_path / "s.ckp")
    g.checkpoint_save(path)
    g.load_step()
    x2 = g.positions(0)
    g.close()
    m = dict(d, X=X1, x=X1, ea=ea1, f_ext=F2, time=1.0 + DT)
    x_o, it, _ = O.newton_solve(O.make_cell(m), X1, rtol=1e-12, atol=1e-6, maxit=25)
    assert it >= 0
    assert np.max(np.abs(x2 - x_o)) <= 1e-6 * np.max(np.abs(x_o - X1))
    h = W.Context(precision=W.PREC_MIXED, t0=1.0, dt=DT, newton_rtol=1e-11, newton_atol=1e-6, max_refine=16)
    h.checkpoint_load(path)
    h.finalize()
    h.load_step()
    assert np.array_equal(h.positions(0), x2)
    h.close()
    # a cell added without a load has none to replace
    k = W.Context(precision=W.PREC_MIXED, t0=1.0, dt=DT)
    try:
        k.add_cell(X=d["X"], tris=d["tris"], law=0, mu=1e6, thick=1.0, passive=1, drag=DRAG)
        k.finalize()
        with pytest.raises(W.WcfeError) as e:
            k.set_external_forces(0, F2)
        assert e.value.code == 1
    finally:
        k.close()


def test_cell_and_substrate_in_one_batch():
    """One staggered coupling step: the cell's integrin forces on the
    substrate (wcfe_force_fa_subs) become the substrate's load, and one batch
    advances the anchored cell and the substrate together.  Each matches its
    own oracle solve; the flags stay per cell (the cell keeps its active
    stress, the substrate its load).  A cell that converges first starts its
    next load step while the other iterates, so each cell's first committed
    positions are read right after the Newton iteration that commits them."""
    gsub = golden("forcesubs_kMD")
    N = len(gsub["X"])
    rng = np.random.default_rng(4)
    Xc = gsub["X"][:, :2].copy()
    xc = Xc + 0.01 * rng.standard_normal(Xc.shape)
    cC = 0.1 * gsub["cC"]
    fau3 = gsub["fau"].copy()
    cell = dict(X=Xc.reshape(-1), x=xc.reshape(-1), tris=gsub["tris"].reshape(-1).astype(np.int32), law=0,
                mu=1000.0, thick=1.0, time=1.0, ea=np.zeros(3 * N), node_springs=1, cC=cC,
                fau0=fau3[:, :2].reshape(-1), rho_max=100.0, spring_law=0, k_lin=31000.0, dt=DT)
    with W.Context() as c0:
        x3 = np.zeros((N, 3))
        x3[:, :2] = Xc
        F = c0.force_fa_subs(gsub["tris"], x3, cC, fau3, gsub["Stris"], gsub["Sx"], 0, 31000.0)
    sub = substrate_model()
    sub["f_ext"] = F.reshape(-1, 3)[:, :2].reshape(-1)
    g = W.Context(precision=W.PREC_MIXED, t0=1.0, dt=DT, newton_rtol=1e-11, newton_atol=1e-6, max_refine=16)
    g.add_cell(X=cell["X"], x=cell["x"], tris=cell["tris"], law=0, mu=1000.0, thick=1.0, ea=cell["ea"],
               node_springs=1, cC=cC, fau=cell["fau0"], rho_max=100.0, spring_law=W.SPRING_LINEAR, k_lin=31000.0)
    add_substrate(g, sub)
    g.finalize()
    xs = {}
    for _ in range(40):
        g.newton_step()
        for c in (0, 1):
            if c not in xs and g.cell_info(c)["step"] >= 1:
                xs[c] = g.ul_state(c)[0]
        if len(xs) == 2:
            break
    g.close()
    assert len(xs) == 2
    xs = [xs[0], xs[1]]
    for x_gpu, m in zip(xs, (cell, sub)):
        x_o, it, _ = O.newton_solve(O.make_cell(m), m["x"], rtol=1e-12, atol=1e-6, maxit=25)
        assert it >= 0
        assert np.max(np.abs(x_gpu - x_o)) <= 1e-6 * np.max(np.abs(x_o - m["X"]))
