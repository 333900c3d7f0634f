"""Pins for the oracle's new components (the planar Newton model the
reference lacks).  Each check is independent of the formula it tests:
finite differences, a textbook closed form, LAPACK via scipy, rocRAND's
Philox and the Random123 known-answer vectors, or the golden-pinned Part 1
restatement of the reference.
"""
import numpy as np
import pytest
import scipy.linalg as sla

import cells
import oracle_py as O
from conftest import golden


def fd_tangent(d, h=1e-7):
    c = O.make_cell(d)
    G0, K = O.assemble(c)
    n = len(G0)
    Kfd = np.zeros((n, n))
    x0 = d["x"].copy()
    for j in range(n):
        for s in (+1, -1):
            dd = dict(d)
            xx = x0.copy()
            xx[j] += s * h
            dd["x"] = xx
            G, _ = O.assemble(O.make_cell(dd), want_K=False)
            Kfd[:, j] += s * G / (2 * h)
    return K, Kfd


@pytest.mark.parametrize("kind", ["cst_linear", "cst_table", "ul_table", "ul_linear"])
def test_tangent_is_derivative_of_residual(kind):
    if kind.startswith("cst"):
        d = cells.cst_cell(K=3, n_spr=25, law_table=kind.endswith("table"), seed=3)
        d["x"] = d["X"] + 0.003 * np.random.default_rng(0).standard_normal(len(d["X"]))
    else:
        d = cells.ul_cell(K=3, table=kind.endswith("table"))
    K, Kfd = fd_tangent(d)
    err = np.max(np.abs(K - Kfd)) / np.max(np.abs(K))
    assert err < 2e-6, err


def test_ul_planar_residual_equals_reference_restatement():
    """Planar model (mu = 1000, thick = 1, no springs) == Part 1 calculateRes
    x/y components on the golden inputs (which Part 1 matches to 1e-16)."""
    g = golden("ul_disk296_t1p3")
    X, x, ea = g["X"], g["x"], g["ea"]
    N = len(X)
    d = dict(X=X[:, :2].reshape(-1), x=x[:, :2].reshape(-1), tris=g["tris"].reshape(-1), law=O.LAW_UL, mu=1000.0,
             thick=1.0, time=float(g["time"]), ea=ea[:, [0, 1, 3]].reshape(-1))
    G, _ = O.assemble(O.make_cell(d), want_K=False)
    ref = g["res_Res"].reshape(N, 3)[:, :2].reshape(-1)
    assert np.max(np.abs(G - ref)) <= 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("case,law", [("forcefa_k1", 0), ("forcefa_kMD", 1)])
def test_node_springs_equal_reference_force_fa(case, law):
    """Planar node springs: G_spring = -F_FA (UL:1198, golden pinned).  The
    reference force depends on x (one-ring areas) and fa_u only, so the
    planar model is evaluated with X = x, which makes its extension
    fau0 + (x - X) exactly fa_u (the fixture places some extensions exactly on
    the table's discontinuous knots)."""
    g = golden(case)
    x, fau, cC, tris = g["x"], g["fau"], g["cC"], g["tris"]
    N = len(x)
    d = dict(X=x[:, :2].reshape(-1), x=x[:, :2].reshape(-1), tris=tris.reshape(-1), law=O.LAW_UL, mu=0.0,
             thick=1.0, time=0.0, ea=np.zeros(3 * N), node_springs=1, cC=cC,
             fau0=fau[:, :2].reshape(-1), rho_max=100.0,
             spring_law=O.SPRING_LINEAR if law == 0 else O.SPRING_TABLE, k_lin=float(g["kint"]))
    if law == 1:
        u = cells.ul_cell()
        d.update(tab_px=u["tab_px"], tab_py=u["tab_py"], tab_k=u["tab_k"])
    G, _ = O.assemble(O.make_cell(d), want_K=False)
    ref = -g["F_FA"].reshape(N, 3)[:, :2].reshape(-1)
    # D == 0 nodes: the reference force is 0 there and so is ours (k0 * u = 0)
    assert np.max(np.abs(G - ref)) <= 1e-11 * np.max(np.abs(ref))


def test_ul_small_strain_tangent_is_incompressible_plane_stress():
    """At F = I, ea = 0, t_act = 0 the UL membrane tangent equals linear
    plane-stress CST with E = 3 mu, nu = 1/2 (incompressible limit)."""
    d = cells.cst_cell(K=3, n_spr=0)
    mu, th = 1000.0, 1.0
    ul = dict(X=d["X"], x=d["X"].copy(), tris=d["tris"], law=O.LAW_UL, mu=mu, thick=th, time=0.0,
              ea=np.zeros(3 * (len(d["X"]) // 2)))
    cst = dict(X=d["X"], x=d["X"].copy(), tris=d["tris"], law=O.LAW_CST, E=3 * mu, nu=0.5, thick=th, time=0.0)
    G1, K1 = O.assemble(O.make_cell(ul))
    G2, K2 = O.assemble(O.make_cell(cst))
    assert np.max(np.abs(G1)) < 1e-9
    assert np.max(np.abs(K1 - K2)) <= 1e-10 * np.max(np.abs(K2))


def test_cst_textbook_right_triangle():
    """K_e of the unit right triangle (0,0),(1,0),(0,1), plane stress,
    thickness t: t/2 B^T D B with B the constant-strain matrix (Zienkiewicz)."""
    E, nu, t = 1000.0, 0.3, 0.2
    d = dict(X=np.array([0, 0, 1, 0, 0, 1.0]), x=np.array([0, 0, 1, 0, 0, 1.0]), tris=np.array([0, 1, 2]),
             law=O.LAW_CST, E=E, nu=nu, thick=t, time=0.0)
    _, K = O.assemble(O.make_cell(d))
    B = np.array([[-1, 0, 1, 0, 0, 0], [0, -1, 0, 0, 0, 1], [-1, -1, 0, 1, 1, 0]], dtype=float)
    Dm = E / (1 - nu ** 2) * np.array([[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu) / 2]])
    Kref = 0.5 * t * B.T @ Dm @ B
    assert np.allclose(K, Kref, rtol=1e-13, atol=1e-10)


def test_cst_rigid_modes_and_patch():
    d = cells.cst_cell(K=4, n_spr=0)
    _, K = O.assemble(O.make_cell(d))
    X = d["X"].reshape(-1, 2)
    tx = np.tile([1.0, 0.0], len(X))
    ty = np.tile([0.0, 1.0], len(X))
    rot = np.stack([-X[:, 1], X[:, 0]], axis=1).reshape(-1)
    for r in (tx, ty, rot):
        assert np.max(np.abs(K @ r)) < 1e-9 * np.max(np.abs(K))
    # patch test: a uniform strain field leaves interior nodes force free
    eps = np.array([[0.01, 0.004], [0.004, -0.006]])
    u = (X @ eps.T).reshape(-1)
    f = (K @ u).reshape(-1, 2)
    _, _, bnd = __import__("wholecellfe_amd.mesh", fromlist=["disk_rings"]).disk_rings(4)
    assert np.max(np.abs(f[~bnd])) < 1e-9 * np.max(np.abs(f))


def test_coloring_is_valid_and_greedy():
    d = cells.cst_cell(K=8, n_spr=0)
    tris = d["tris"].reshape(-1, 3)
    N = len(d["X"]) // 2
    col, nc = O.color_elements(N, tris)
    assert 3 <= nc <= 12
    for c in range(nc):
        nodes = tris[col == c].reshape(-1)
        assert len(nodes) == len(np.unique(nodes))
    # first-fit: every element's colour is the lowest not used by an earlier neighbour
    used = [set() for _ in range(N)]
    for e, t in enumerate(tris):
        forb = used[t[0]] | used[t[1]] | used[t[2]]
        assert col[e] == min(set(range(64)) - forb)
        for v in t:
            used[v].add(col[e])


def test_colour_order_assembly_equals_element_order():
    d = cells.cst_cell(K=5, n_spr=30, law_table=True)
    d["x"] = d["X"] + 0.002 * np.random.default_rng(1).standard_normal(len(d["X"]))
    c = O.make_cell(d)
    col, nc = O.color_elements(c.n_nodes, d["tris"])
    G1, K1 = O.assemble(c)
    G2, K2 = O.assemble(c, col, nc)
    assert np.allclose(G1, G2, rtol=0, atol=1e-12 * np.max(np.abs(G1)))
    assert np.allclose(K1, K2, rtol=0, atol=1e-12 * np.max(np.abs(K1)))


@pytest.mark.parametrize("n", [1, 3, 17, 64, 150])
def test_lu_partial_pivoting_matches_lapack(n):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    b = rng.standard_normal(n)
    LU, piv, info = O.lu_factor(A, pivoting=1)
    assert info == 0
    lu_ref, piv_ref = sla.lu_factor(A)
    assert np.array_equal(piv, piv_ref)
    assert np.allclose(LU, lu_ref, rtol=1e-10, atol=1e-12)
    x = O.lu_solve(LU, piv, b, pivoting=1)
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-9, atol=1e-12)


def test_lu_without_pivoting_on_spd_and_zero_pivot():
    rng = np.random.default_rng(4)
    M = rng.standard_normal((40, 40))
    A = M @ M.T + 40 * np.eye(40)
    b = rng.standard_normal(40)
    LU, piv, info = O.lu_factor(A, pivoting=0)
    assert info == 0
    assert np.allclose(O.lu_solve(LU, piv, b, pivoting=0), np.linalg.solve(A, b), rtol=1e-10)
    Z = np.array([[0.0, 1.0], [1.0, 0.0]])
    assert O.lu_factor(Z, pivoting=0)[2] == 1
    assert O.lu_factor(Z, pivoting=1)[2] == 0


def test_pcg_on_laplacian():
    n = 30
    T = 2 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    A = np.kron(T, np.eye(n)) + np.kron(np.eye(n), T)
    rows, cols = np.nonzero(A)
    rowptr = np.searchsorted(rows, np.arange(n * n + 1))
    b = np.random.default_rng(0).standard_normal(n * n)
    x, it = O.pcg_csr(rowptr, cols, A[rows, cols], b, rtol=1e-10)
    assert it > 0
    assert np.linalg.norm(A @ x - b) <= 1.1e-10 * np.linalg.norm(b)
    # CG on the 2D Laplacian: iterations ~ sqrt(cond) ln(2/tol)/2, cond ~ (2(n+1)/pi)^2
    cond = (2 * (n + 1) / np.pi) ** 2
    assert it <= np.sqrt(cond) * np.log(2 / 1e-10) / 2 + 10


def test_philox_known_answers_and_rocrand():
    g = golden("philox_rocrand")
    kat = [[0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8], [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd],
           [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]]  # Random123 kat_vectors, philox4x32 R=10
    for i, row in enumerate(g["inputs"]):
        out = O.philox([int(v) for v in row[:4]], [int(v) for v in row[4:]])
        assert out == [int(v) for v in g["outputs"][i]]
        if i < 3:
            assert out == kat[i]


def test_newton_converges_quadratically_on_table_springs():
    d = cells.cst_cell(K=5, n_spr=60, law_table=True, time=1.5)
    c = O.make_cell(d)
    x, it, hist = O.newton_solve(c, d["x"], rtol=1e-12, atol=1e-14, maxit=12)
    assert it > 0
    h = hist[: it + 1]
    assert h[-1] <= 1e-12 * h[0]
    # superlinear tail
    assert h[-1] < 1e-3 * h[-2] or h[-2] < 1e-8 * h[0]


def test_bell_update_statistics():
    """Bound fraction after many steps approaches the mean-field fixed point
    k_on / (k_on + k_off) for unloaded springs (anchors at the attachment)."""
    d = cells.cst_cell(K=6, n_spr=400, seed=9)
    c = O.make_cell(d)
    st = np.ones(len(d["spr_tri"]))
    an = d["spr_anchor"].copy()
    fr = []
    for step in range(400):
        st, an, _, _ = O.bell_update(c, st, an, 0.05, 1.0, 0.1, 4.0, (5, 0), 0, step)
        fr.append(st.mean())
    assert abs(np.mean(fr[200:]) - 1.0 / 1.1) < 0.03
