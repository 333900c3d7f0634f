"""GPU parity tests: the HIP path (through the C-ABI) against the reference's
golden vectors and the FP64 oracle.

Tolerances (FP64 kernels vs the FP64 oracle/reference): 1e-10 relative to the
array's max magnitude for residuals, forces and tangents (different but
exact-arithmetic-equivalent formulations: analytic planar tangent vs the
oracle's dual-number derivative of the 3D restatement).  Linear solves: FP64
relative residual <= 1e-8 (the refinement target).  Integer state (colours,
pivots, bond states): bit-exact.
"""
import numpy as np
import pytest

import cells
import oracle_py as O
from conftest import golden

pytestmark = pytest.mark.gpu

W = pytest.importorskip("wholecellfe_amd")


def relerr(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module")
def ctx():
    c = W.Context()
    yield c
    c.close()


def test_device_is_gfx950(ctx):
    assert "gfx950" in ctx.arch()


# ---------------------------------------------------------------------------
# reference-shaped calls vs the compiled reference's golden vectors
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("case", ["ul_disk296_t1p3", "ul_disk483_t3"])
def test_calculate_res_vs_reference(ctx, case):
    g = golden(case)
    print("case", case, flush=True)
    Res, Rea, MM = ctx.calculate_res(g["tris"], g["X"], g["x"], g["ea"], float(g["time"]))
    assert relerr(Res, g["res_Res"]) < 1e-10
    assert relerr(Rea, g["res_Res_ea"]) < 1e-10
    assert relerr(MM, g["res_MM"]) < 1e-12


@pytest.mark.parametrize("case,law", [("forcefa_k1", 0), ("forcefa_k31", 0), ("forcefa_kMD", 1),
                                      ("forcefa_kMD_3d", 1)])
def test_force_fa_vs_reference(ctx, case, law):
    g = golden(case)
    print("case", case, flush=True)
    F = ctx.force_fa(g["tris"], g["x"], g["cC"], g["fau"], law, float(g["kint"]))
    assert relerr(F, g["F_FA"]) < 1e-12


@pytest.mark.parametrize("case,law", [("updatefa_k1_rigid", 0), ("updatefa_k31_moving", 0),
                                      ("updatefa_kMD_rigid", 1), ("updatefa_kMD_moving", 1)])
def test_update_fa_vs_reference(ctx, case, law):
    g = golden(case)
    print("case", case, flush=True)
    cA, cC, fau, un = ctx.update_fa(g["X"], g["x"], g["cA"], g["cC"], g["fau"], float(g["dt"]), g["Stris"], g["SX"],
                                    g["Sx"], law, float(g["kint"]))
    assert un == 0
    assert relerr(cA, g["out_cA"]) < 1e-12
    assert relerr(cC, g["out_cC"]) < 1e-12
    assert relerr(fau, g["out_fau"]) < 1e-11


@pytest.mark.parametrize("case,law", [("forcesubs_k31", 0), ("forcesubs_kMD", 1)])
def test_force_fa_subs_vs_reference(ctx, case, law):
    g = golden(case)
    print("case", case, flush=True)
    F = ctx.force_fa_subs(g["tris"], g["x"], g["cC"], g["fau"], g["Stris"], g["Sx"], law, float(g["kint"]))
    assert relerr(F, g["F_subs"]) < 1e-12


def _substrate_variant(g, kind):
    """The fixture's substrate as given, with every triangle clockwise (the
    containment test then accepts the triangle's reflection through its first
    vertex), or lifted off z = 0 (point location falls back to every
    triangle)."""
    Stris, SX, Sx = g["Stris"].copy(), g["SX"].copy(), g["Sx"].copy()
    if kind == "clockwise":
        Stris = Stris[:, [0, 2, 1]]
    elif kind == "lifted":
        SX[:, 2] = 1e-3 * np.sin(SX[:, 0])
        Sx[:, 2] = SX[:, 2]
    return Stris, SX, Sx


@pytest.mark.parametrize("kind", ["as_given", "clockwise", "lifted"])
@pytest.mark.parametrize("law", [0, 1])
def test_force_fa_subs_and_update_fa_vs_oracle(ctx, kind, law):
    g = golden("forcesubs_kMD")
    print("substrate", kind, "law", law, flush=True)
    Stris, SX, Sx = _substrate_variant(g, kind)
    kint = 31000.0 if law == 0 else -1.0
    F = ctx.force_fa_subs(g["tris"], g["x"], g["cC"], g["fau"], Stris, Sx, law, kint)
    Fo = O.force_fa_subs(g["tris"], g["x"], g["cC"], g["fau"], Stris, Sx, law, kint)
    np.testing.assert_array_equal(np.abs(F) > 0, np.abs(Fo) > 0)
    assert relerr(F, Fo) < 1e-13
    cA = np.full(len(g["X"]), 50.0)
    out = ctx.update_fa(g["X"], g["x"], cA, g["cC"], g["fau"], 0.005, Stris, SX, Sx, law, kint)
    ref = O.update_fa(g["tris"], g["X"], g["x"], cA, g["cC"], g["fau"], 0.005, Stris, SX, Sx, law, kint)
    assert out[3] == ref[3]
    for a, b in zip(out[:3], ref[:3]):
        assert relerr(a, b) < 1e-12


def test_force_fa_subs_large_substrate_vs_oracle(ctx):
    # 20k-triangle substrate grid, 2k anchors, some on vertices and edges
    n = 101
    gx, gy = np.meshgrid(np.linspace(-50, 50, n), np.linspace(-50, 50, n))
    Sx = np.zeros((n * n, 3))
    Sx[:, 0], Sx[:, 1] = gx.ravel(), gy.ravel()
    q = np.arange(n * n).reshape(n, n)
    a, b, c, d = q[:-1, :-1].ravel(), q[:-1, 1:].ravel(), q[1:, 1:].ravel(), q[1:, :-1].ravel()
    Stris = np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)])
    g = golden("forcesubs_kMD")
    rng = np.random.default_rng(5)
    reps = 7
    x = np.concatenate([g["x"] + [rng.uniform(-40, 40), rng.uniform(-40, 40), 0] for _ in range(reps)])
    tris = np.concatenate([g["tris"] + k * len(g["x"]) for k in range(reps)])
    fau = rng.uniform(-0.01, 0.01, x.shape)
    fau[:, 2] = 0
    fau[:40] = x[:40] - Sx[rng.integers(0, n * n, 40)]                      # on vertices
    fau[40:80] = x[40:80] - (Sx[Stris[:40, 0]] + Sx[Stris[:40, 2]]) / 2  # on diagonals
    cC = rng.uniform(0, 1, len(x))
    F = ctx.force_fa_subs(tris, x, cC, fau, Stris, Sx, 0, 1000.0)
    Fo = O.force_fa_subs(tris, x, cC, fau, Stris, Sx, 0, 1000.0)
    assert relerr(F, Fo) < 1e-13


def test_calculate_res_rejects_out_of_plane(ctx):
    g = golden("ul_disk296_t1p3")
    x = g["x"].copy()
    x[0, 2] = 1e-3
    with pytest.raises(W.WcfeError) as e:
        ctx.calculate_res(g["tris"], g["X"], x, g["ea"], 1.0)
    assert e.value.code == 1


# ---------------------------------------------------------------------------
# element systems and assembly vs the oracle
# ---------------------------------------------------------------------------
def _gpu_cell(d, precision=W.PREC_FP64, kinetics=0, **opt):
    c = W.Context(precision=precision, **opt)
    cells.add_to_context(c, d, kinetics=kinetics)
    c.finalize()
    return c


@pytest.mark.parametrize("kind", ["cst_linear", "cst_table", "ul_table", "ul_linear"])
def test_element_systems_vs_oracle(kind):
    print("case", kind, flush=True)
    if kind.startswith("cst"):
        d = cells.cst_cell(K=5, n_spr=60, law_table=kind.endswith("table"), seed=11)
        d["x"] = d["X"] + 0.003 * np.random.default_rng(2).standard_normal(len(d["X"]))
    else:
        d = cells.ul_cell(K=5, table=kind.endswith("table"))
    g = _gpu_cell(d, t0=d["time"])
    Ge, Ke = g.element_systems(0)
    oc = O.make_cell(d)
    T = len(d["tris"]) // 3
    Go = np.zeros((T, 6))
    Ko = np.zeros((T, 6, 6))
    for e in range(T):
        Go[e], Ko[e] = O.element_system(oc, e)
    g.close()
    assert relerr(Ge, Go) < 1e-10
    assert relerr(Ke, Ko) < 1e-10


@pytest.mark.parametrize("prec", [W.PREC_FP64, W.PREC_MIXED])
def test_assembled_system_and_colours_vs_oracle(prec):
    d = cells.cst_cell(K=6, n_spr=50, law_table=True, seed=5)
    d["x"] = d["X"] + 0.002 * np.random.default_rng(3).standard_normal(len(d["X"]))
    g = _gpu_cell(d, precision=prec, t0=d["time"])
    K, G = g.assembled_system(0)
    col = g.colors(0)
    g.close()
    oc = O.make_cell(d)
    ocol, nc = O.color_elements(oc.n_nodes, d["tris"])
    assert np.array_equal(col, ocol)
    Go, Ko = O.assemble(oc, ocol, nc)
    assert relerr(G, Go) < 1e-10
    tol = 1e-10 if prec == W.PREC_FP64 else 1e-6  # FP32 storage of K in the mixed form
    assert relerr(K, Ko) < tol


def test_residual_fp64_vs_oracle():
    d = cells.ul_cell(K=6, table=True)
    g = _gpu_cell(d, t0=d["time"])
    G = g.residual(0)
    g.close()
    Go, _ = O.assemble(O.make_cell(d), want_K=False)
    assert relerr(G, Go) < 1e-10


# ---------------------------------------------------------------------------
# dense LU + FP64 refinement
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("prec,piv", [(W.PREC_MIXED, W.PIVOT_STATIC), (W.PREC_MIXED, W.PIVOT_PARTIAL),
                                      (W.PREC_FP64, W.PIVOT_STATIC), (W.PREC_FP64, W.PIVOT_PARTIAL)])
def test_dense_solve_batched_uneven(prec, piv):
    print("case prec", prec, "piv", piv, flush=True)
    rng = np.random.default_rng(7)
    mats, rhss = [], []
    for n in [1, 5, 63, 64, 65, 130, 300]:
        if piv == W.PIVOT_PARTIAL:
            A = rng.standard_normal((n, n)) + 2 * np.sqrt(n) * np.eye(n) * rng.choice([-1, 1], n)
        else:
            M = rng.standard_normal((n, n))
            A = M @ M.T / n + np.eye(n)
        mats.append(A)
        rhss.append(rng.standard_normal(n))
    c = W.Context(precision=prec, pivoting=piv, max_refine=10)
    xs, rel, it, pv = c.dense_solve(mats, rhss)
    c.close()
    for A, b, x, r in zip(mats, rhss, xs, rel):
        res = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
        assert res <= 1e-8, (A.shape, res)
        assert r <= 1e-8
        assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-6, atol=1e-9)
    assert it.max() <= 10


@pytest.mark.parametrize("prec", [W.PREC_MIXED, W.PREC_FP64])
def test_dense_solve_symmetric_path(prec):
    """Exactly symmetric systems (SPD and indefinite) take the lower-triangle
    factorisation with U = D L^T; both it and the full LU reach 1e-8."""
    print("case symmetric prec", prec, flush=True)
    rng = np.random.default_rng(11)
    mats, rhss = [], []
    for n in [5, 64, 65, 130, 300, 611]:
        T = rng.standard_normal((n, n))
        S = T + T.T  # exactly symmetric
        if n % 2:
            A = S / np.sqrt(n) + 3.0 * np.eye(n) * rng.choice([-1, 1], n)  # indefinite
        else:
            A = S @ S / n + np.eye(n)
            A = 0.5 * (A + A.T)
        assert np.array_equal(A, A.T)
        mats.append(A)
        rhss.append(rng.standard_normal(n))
    out = {}
    for sym in (1, 0):
        c = W.Context(precision=prec, pivoting=W.PIVOT_STATIC, max_refine=10, symmetric=sym)
        xs, rel, it, _ = c.dense_solve(mats, rhss)
        c.close()
        for A, b, x, r in zip(mats, rhss, xs, rel):
            res = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
            assert res <= 1e-8, (sym, A.shape, res)
            assert r <= 1e-8
            assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-6, atol=1e-9)
        out[sym] = xs
    # multi-block systems went through a different factorisation
    assert any(not np.array_equal(a, b) for a, b in zip(out[0][3:], out[1][3:]))


def test_partial_pivots_match_oracle_fp32():
    """Pivot rows of the GPU panel factorisation equal the FP32 oracle LU's on
    a random matrix (well-separated column maxima)."""
    rng = np.random.default_rng(17)
    n = 200
    A = rng.standard_normal((n, n))
    b = rng.standard_normal(n)
    c = W.Context(precision=W.PREC_MIXED, pivoting=W.PIVOT_PARTIAL)
    _, _, _, pv = c.dense_solve([A], [b])
    c.close()
    _, piv_o, _ = O.lu_factor(A, pivoting=1, f32=True)
    assert np.array_equal(pv, piv_o)


# ---------------------------------------------------------------------------
# Newton step vs the oracle's FP64 Newton solve
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("prec", [W.PREC_MIXED, W.PREC_FP64])
@pytest.mark.parametrize("table", [False, True])
def test_newton_load_step_vs_oracle(prec, table):
    print("case prec", prec, "table", table, flush=True)
    d = cells.cst_cell(K=7, n_spr=80, law_table=table, time=1.2, seed=21)
    g = _gpu_cell(d, precision=prec, t0=d["time"], newton_rtol=1e-10, newton_atol=1e-10, max_refine=12)
    st = g.load_step()
    x_gpu = g.positions(0)
    info = g.cell_info(0)
    g.close()
    assert info["step"] == 1
    assert st["max_lin_relres"] <= 1e-8
    x_o, it, hist = O.newton_solve(O.make_cell(d), d["x"], rtol=1e-12, atol=1e-12, maxit=20)
    assert it >= 0
    u_o = x_o - d["X"]
    assert np.max(np.abs(x_gpu - x_o)) <= 1e-8 * np.max(np.abs(u_o))
    # the GPU solution's own FP64 residual, evaluated by the oracle
    dd = dict(d)
    dd["x"] = x_gpu
    G, _ = O.assemble(O.make_cell(dd), want_K=False)
    G0, _ = O.assemble(O.make_cell(d), want_K=False)
    assert np.linalg.norm(G) <= 1e-8 * np.linalg.norm(G0)


def test_symmetric_factorisation_matches_full_lu():
    """A batch of uneven cells (up to 15 LU blocks) advanced one load step with
    the symmetric (lower-triangle) factorisation and with the full LU: the
    same converged positions to well below the Newton tolerance."""
    ds = [cells.cst_cell(K=k, n_spr=30 + 10 * k, law_table=(k % 2 == 1), seed=k, time=1.0) for k in (5, 9, 13)]
    outs = []
    for sym in (1, 0):
        g = W.Context(precision=W.PREC_MIXED, t0=1.0, newton_rtol=1e-11, newton_atol=1e-11, max_refine=12,
                      symmetric=sym)
        for d in ds:
            cells.add_to_context(g, d)
        g.finalize()
        st = g.load_step()
        assert st["max_lin_relres"] <= 1e-8
        outs.append([g.positions(i) for i in range(len(ds))])
        g.close()
    for d, a, b in zip(ds, outs[0], outs[1]):
        assert np.max(np.abs(a - b)) <= 1e-9 * np.max(np.abs(b - d["X"]))


def test_ul_md_table_one_newton_iteration_vs_oracle():
    """Reference law + node springs with the kMD force table: one Newton
    iteration against the oracle's.  (A full load step is not asserted: the
    table is discontinuous with a negative-stiffness segment and Newton cycles
    on it, in the FP64 oracle as well.)"""
    d = cells.ul_cell(K=6, table=True, time=1.3, deform=0.0)
    d["x"] = d["X"].copy()
    g = _gpu_cell(d, t0=d["time"])
    g.newton_step()
    x_gpu = g.positions(0)
    g.close()
    x_o, _, _ = O.newton_solve(O.make_cell(d), d["x"], rtol=1e-300, atol=0.0, maxit=1)
    assert np.max(np.abs(x_gpu - x_o)) <= 1e-8 * np.max(np.abs(x_o - d["X"]))


def test_ul_reference_model_load_step_vs_oracle():
    """Reference law + node springs (linear k = 1 pN/nm, the k1 variant) +
    mean-field kinetics: one load step; positions vs the oracle Newton solve,
    then the commit (strain update UL:607-612 and updateFA kinetics
    UL1:2046-2109) vs the oracle restatement."""
    d = cells.ul_cell(K=6, table=False, time=1.3, deform=0.0)
    d["x"] = d["X"].copy()
    g = _gpu_cell(d, t0=d["time"], newton_rtol=1e-10, newton_atol=1e-9, kinetics=W.KIN_MEANFIELD)
    st = g.load_step()
    x_gpu = g.positions(0)
    X_new, ea_new, cC_new, fau_new = g.ul_state(0)
    g.close()
    x_o, it, _ = O.newton_solve(O.make_cell(d), d["x"], rtol=1e-12, atol=1e-11, maxit=25)
    assert it >= 0
    assert np.max(np.abs(x_gpu - x_o)) <= 1e-8 * max(np.max(np.abs(x_o - d["X"])), 1e-12)
    # commit: strain update with the GPU's converged x
    dd = dict(d)
    dd["x"] = x_gpu
    ea_o = O.commit_ul_strain(O.make_cell(dd))
    assert relerr(ea_new, ea_o) < 1e-10
    assert np.array_equal(X_new, x_gpu)
    # kinetics: the reference updateFA restatement (rigid substrate) on 3D arrays
    N = len(d["X"]) // 2
    X3 = np.zeros((N, 3)); X3[:, :2] = d["X"].reshape(-1, 2)
    x3 = np.zeros((N, 3)); x3[:, :2] = x_gpu.reshape(-1, 2)
    fau3 = np.zeros((N, 3)); fau3[:, :2] = d["fau0"].reshape(-1, 2)
    SX = np.array([[-50.0, -50, 0], [50, -50, 0], [50, 50, 0], [-50, 50, 0]])
    ST = np.array([[0, 1, 2], [0, 2, 3]])
    _, cC_o, fau_o, un = O.update_fa(d["tris"].reshape(-1, 3), X3, x3, np.full(N, 100.0), d["cC"], fau3, 0.005,
                                     ST, SX, SX, 0, 1000.0)
    assert un == 0
    assert relerr(cC_new, cC_o) < 1e-12
    assert relerr(fau_new, fau_o[:, :2]) < 1e-10


def test_batched_graph_equals_per_cell_bitwise():
    """The batched, graph-captured launch sequence and the one-cell-at-a-time
    sequence run the same per-cell arithmetic in the same order."""
    ds = [cells.cst_cell(K=k, n_spr=20 + 5 * k, law_table=(k % 2 == 0), seed=k, time=1.0) for k in (3, 6, 9)]
    outs = []
    for batched in (1, 0):
        g = W.Context(precision=W.PREC_MIXED, batched=batched, use_graph=batched, t0=1.0)
        for d in ds:
            cells.add_to_context(g, d)
        g.finalize()
        for _ in range(4):
            g.newton_step(stats=False)
        g.synchronize()
        outs.append([g.positions(i) for i in range(len(ds))])
        g.close()
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)


def test_graph_replays_match_eager():
    d = cells.cst_cell(K=6, n_spr=40, law_table=True, seed=4, time=1.0)
    res = []
    for use_graph in (1, 0):
        g = _gpu_cell(d, precision=W.PREC_MIXED, use_graph=use_graph, t0=1.0)
        xs = []
        for _ in range(5):
            g.newton_step(stats=False)
            g.synchronize()
            xs.append(g.positions(0))
        res.append(xs)
        g.close()
    for a, b in zip(res[0], res[1]):
        assert np.array_equal(a, b)


def test_bell_kinetics_bit_exact():
    """Bond states and anchors after each committed load step equal the
    oracle's Bell update (same Philox stream) applied to the GPU's converged
    positions: bit-exact."""
    d = cells.cst_cell(K=6, n_spr=120, law_table=False, seed=8, time=1.0)
    g = _gpu_cell(d, precision=W.PREC_MIXED, kinetics=W.KIN_BELL, t0=1.0, dt=0.2, seed_lo=5, seed_hi=0,
                  newton_rtol=1e-10, newton_atol=1e-10)
    oc_d = dict(d)
    transitions = 0
    for step in range(6):
        st0, an0 = g.springs(0)
        g.load_step()
        x = g.positions(0)
        st1, an1 = g.springs(0)
        oc_d["x"] = x
        oc_d["spr_state"] = st0
        oc_d["spr_anchor"] = an0
        st_o, an_o, nt, margin = O.bell_update(O.make_cell(oc_d), st0, an0, 0.2, 1.0, 0.1, 4.0, (5, 0), 0, step)
        transitions += nt
        assert np.array_equal(st1, st_o), step
        assert np.array_equal(an1, an_o), step
        # no decision within rounding distance of its probability: the states
        # would agree even with an exp that differed in the last bits
        assert margin > 1e-13, (step, margin)
    g.close()
    assert transitions > 0


def test_gpu_exp_equals_cpu_exp_bitwise():
    """The kinetics' exp on the GPU and the same function compiled for the CPU
    (oracle) round identically, over the kinetics' arguments and the whole
    range."""
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(-50.0, 50.0, 200000), rng.uniform(-745.0, 709.7, 200000),
                        np.array([0.0, -0.0, 1e-300, 710.5, -746.5, np.inf, -np.inf])])
    with W.Context() as g:
        y = g.eval_exp(x)
    y_cpu = np.zeros_like(x)
    O.lib().orc_exp_array(len(x), O.dp(x), O.dp(y_cpu))
    assert np.array_equal(y.view(np.int64), y_cpu.view(np.int64))


def test_degenerate_element_is_reported():
    d = cells.cst_cell(K=3, n_spr=5, seed=1)
    X = d["X"].copy()
    T = d["tris"].reshape(-1, 3)
    X[2 * T[0, 1]: 2 * T[0, 1] + 2] = X[2 * T[0, 0]: 2 * T[0, 0] + 2]  # collapse element 0
    d["X"] = X
    d["x"] = X.copy()
    g = W.Context(precision=W.PREC_FP64)
    cells.add_to_context(g, d)
    g.finalize()
    with pytest.raises(W.WcfeError) as e:
        g.newton_step()
    g.close()
    assert e.value.code == 4
