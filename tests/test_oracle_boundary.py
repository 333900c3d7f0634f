"""Oracle boundary forces (bending, area constraint, actin polymerisation)
against the compiled reference's golden vectors and closed forms.

Reference: UL_growth_explicit.cpp (wholeCellFE_kMD/cellBP) calculateForce_AP
:1477-1568, calculateBendingForce2D :1571-1660, calculateAreaConstraintForce
:1691-1784; boundary vertices by El Topo's edge-valence rule
(eltopo3d/nondestructivetrimesh.cpp:527-543).
"""
import math
import os

import numpy as np
import pytest

import oracle_py as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
AREA_0 = 10 * 10 * 3.1416  # UL:1777


def gold(name):
    return np.load(os.path.join(GOLD, name + ".npz"))


def fan_polygon(n, r=9.7131, rot=0.3):
    """Regular n-gon (counter-clockwise) fanned from its centre: node 0 is the
    centre, nodes 1..n the boundary."""
    th = rot + 2 * np.pi * np.arange(n) / n
    X = np.zeros((n + 1, 3))
    X[1:, 0], X[1:, 1] = r * np.cos(th), r * np.sin(th)
    T = np.array([[0, 1 + i, 1 + (i + 1) % n] for i in range(n)], dtype=np.int64)
    return X, T


@pytest.mark.parametrize("name", ["bound_disk296", "bound_disk483", "bound_disk296_notch"])
def test_bending_and_area_match_reference(name):
    z = gold(name)
    x, T = z["x"], z["tris"]
    Fb, nopen = O.bending_force_2d(T, x, 20.0)
    assert nopen == 0
    np.testing.assert_allclose(Fb, z["F_bend"], rtol=0, atol=1e-13 * np.abs(z["F_bend"]).max())
    A = O.area_tot(T, x)
    np.testing.assert_allclose(A, z["areatot"], rtol=1e-14)
    Fa, _ = O.area_constraint_force(T, x, A[0], 1.0, AREA_0)
    np.testing.assert_allclose(Fa, z["F_area"], rtol=0, atol=1e-13 * np.abs(z["F_area"]).max())


def test_notched_mesh_has_multi_neighbour_vertices():
    # the fixture exercises the reference's "last candidate wins" rule
    z = gold("bound_disk296_notch")
    isb, st, cand = O.boundary_candidates(len(z["x"]), z["tris"])
    assert np.diff(st).max() >= 4


@pytest.mark.parametrize("name", ["bound_disk296", "bound_disk483"])
def test_ap_matches_reference_realisation(name):
    # calculateForce_AP draws from rand() seeded by the clock: the fixture is
    # one realisation, so each boundary vertex must be either 0 or the fired
    # value, and both outcomes must occur at p = 1/e
    z = gold(name)
    x, T, dt = z["x"], z["tris"], float(z["ap_dt"])
    n = len(x)
    isb, _, _ = O.boundary_candidates(n, T)
    fired, _ = O.force_ap(T, x, dt, np.zeros(n))
    idle, _ = O.force_ap(T, x, dt, np.full(n, 0.999999))
    g = z["F_ap"].reshape(-1, 3)
    f = fired.reshape(-1, 3)
    on = np.abs(g).max(axis=1) > 0
    assert 0 < on.sum() < isb.sum()
    np.testing.assert_array_equal(g[on], f[on])
    assert not np.any(g[~on])
    assert not np.any(g[isb == 0])
    assert not np.any(idle)


def test_candidates_two_triangle_square():
    # square 0-1-2-3 split along 0-2: every vertex is on the boundary; vertex
    # 0 meets 1 and 2 in triangle 0, 2 and 3 in triangle 1
    X = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], float)
    T = np.array([[0, 1, 2], [0, 2, 3]])
    isb, st, cand = O.boundary_candidates(4, T)
    assert isb.tolist() == [1, 1, 1, 1]
    assert cand[st[0]:st[1]].tolist() == [[1, 2], [2, 1], [2, 3], [3, 2]]
    assert cand[st[1]:st[2]].tolist() == [[0, 2], [2, 0]]


@pytest.mark.parametrize("n", [6, 17, 64])
def test_regular_polygon_closed_forms(n):
    X, T = fan_polygon(n)
    r, kb, ka = 9.7131, 20.0, 1.0  # 100 s is not near an integer for these n
    s = 2 * r * math.sin(math.pi / n)            # side
    rhat = X[1:, :2] / r                          # outward unit normals
    # bending: kappa = (cos(2 pi/n) - 1) * 2 / (2 s), normal_i = cos(pi/n) rhat
    Fb, nopen = O.bending_force_2d(T, X, kb)
    assert nopen == 0
    mag = kb * (math.cos(2 * math.pi / n) - 1) / s * math.cos(math.pi / n)
    np.testing.assert_allclose(Fb.reshape(-1, 3)[1:, :2], mag * rhat, rtol=1e-12, atol=1e-12)
    assert not np.any(Fb.reshape(-1, 3)[0])
    # area: A = n/2 r^2 sin(2 pi / n); F_i = -k (A - A0)/2 * |x_{i+1} - x_{i-1}| rhat
    A = O.area_tot(T, X)[0]
    assert A == pytest.approx(n / 2 * r * r * math.sin(2 * math.pi / n), rel=1e-13)
    Fa, _ = O.area_constraint_force(T, X, A, ka, AREA_0)
    mag = -ka * (A - AREA_0) / 2 * 2 * r * math.sin(2 * math.pi / n)
    np.testing.assert_allclose(Fa.reshape(-1, 3)[1:, :2], mag * rhat, rtol=1e-11, atol=1e-12)
    # actin polymerisation: quantum * ceil(density * 2 s / 2) along normal_i
    # when u < rate dt exp(-rate dt)
    dt = 0.02
    p = 10 * dt * math.exp(-10 * dt)
    u = np.full(n + 1, p * (1 - 1e-12))
    u[1::2] = p * (1 + 1e-12)
    Fp, _ = O.force_ap(T, X, dt, u)
    Fp = Fp.reshape(-1, 3)[1:, :2]
    mag = 5 * math.ceil(100 * 2 * s / 2) * math.cos(math.pi / n)
    np.testing.assert_allclose(Fp[1::2], mag * rhat[1::2], rtol=1e-12)
    assert not np.any(Fp[0::2])


def test_area_force_is_area_penalty_gradient():
    # F_area = -k (A - A0) dA/dx on the boundary, with dA/dx by central
    # differences of calculateAreaTot (interior vertices: dA/dx = 0)
    z = gold("bound_disk483")
    x, T = z["x"].copy(), z["tris"]
    A = O.area_tot(T, x)[0]
    Fa, _ = O.area_constraint_force(T, x, A, 1.0, AREA_0)
    isb, _, _ = O.boundary_candidates(len(x), T)
    h = 1e-6
    for i in list(np.flatnonzero(isb)[:12]) + list(np.flatnonzero(isb == 0)[:4]):
        for c in range(2):
            xp, xm = x.copy(), x.copy()
            xp[i, c] += h
            xm[i, c] -= h
            dA = (O.area_tot(T, xp)[0] - O.area_tot(T, xm)[0]) / (2 * h)
            assert Fa[3 * i + c] == pytest.approx(-1.0 * (A - AREA_0) * dA, rel=1e-6, abs=1e-7)


# ---------------- implicit form (planar Newton model) ----------------

def ul_cell(name="bound_disk296", **kw):
    z = gold(name)
    X = z["X"][:, :2].copy()
    N = len(X)
    rng = np.random.default_rng(7)
    x = X @ np.array([[1.03, 0.02], [-0.01, 0.98]]).T + 0.01 * rng.standard_normal(X.shape)
    m = dict(X=X.reshape(-1), x=x.reshape(-1), tris=z["tris"].reshape(-1), law=O.LAW_UL, mu=1000.0, thick=1.0,
             time=1.0, ea=np.zeros(3 * N), k_bend=20.0, k_area=1.0, drag=0.001, dt=0.01,
             ap_u=rng.uniform(size=N))
    m.update(kw)
    return m


def test_node_terms_residual_is_reference_forces():
    # G(with terms) - G(without) = drag (x - X)/dt - F_bend(x) - F_area(x) - F_AP(X)
    m = ul_cell()
    N = len(m["X"]) // 2
    G1, _ = O.assemble(O.make_cell(m), want_K=False)
    G0, _ = O.assemble(O.make_cell(dict(m, k_bend=0.0, k_area=0.0, drag=0.0, ap_u=None)), want_K=False)
    x3 = np.zeros((N, 3)); x3[:, :2] = m["x"].reshape(-1, 2)
    X3 = np.zeros((N, 3)); X3[:, :2] = m["X"].reshape(-1, 2)
    T = m["tris"].reshape(-1, 3)
    Fb, _ = O.bending_force_2d(T, x3, 20.0)
    Fa, _ = O.area_constraint_force(T, x3, O.area_tot(T, x3)[0], 1.0, AREA_0)
    Fp, _ = O.force_ap(T, X3, 0.01, m["ap_u"])
    want = 0.001 / 0.01 * (m["x"] - m["X"]) - (Fb + Fa + Fp).reshape(-1, 3)[:, :2].reshape(-1)
    np.testing.assert_allclose(G1 - G0, want, rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("scale", [1.0, 100.0])
def test_node_terms_tangent_matches_finite_differences(scale):
    # scale amplifies the boundary terms against the membrane stiffness
    m = ul_cell(mu=1000.0 / scale, k_bend=20.0 * scale, k_area=1.0 * scale)
    G, K = O.assemble(O.make_cell(m))
    N = len(m["X"]) // 2
    isb, _, _ = O.boundary_candidates(N, m["tris"].reshape(-1, 3))
    cols = [2 * i + q for i in np.flatnonzero(isb)[:10] for q in range(2)] + [2 * i for i in range(0, N, 29)]
    h = 1e-6
    for j in cols:
        xp, xm = m["x"].copy(), m["x"].copy()
        xp[j] += h
        xm[j] -= h
        Gp, _ = O.assemble(O.make_cell(dict(m, x=xp)), want_K=False)
        Gm, _ = O.assemble(O.make_cell(dict(m, x=xm)), want_K=False)
        np.testing.assert_allclose(K[:, j], (Gp - Gm) / (2 * h), rtol=0, atol=1e-7 * np.abs(K).max())


def test_node_terms_newton_converges_quadratically():
    m = ul_cell(k_bend=200.0, k_area=5.0)
    x, it, hist = O.newton_solve(O.make_cell(m), m["x"], rtol=1e-13, atol=1e-8, maxit=12)
    assert 0 < it <= 8
    h = hist[:it + 1]
    # quadratic phase: some step squares the relative residual
    rel = h / h[0]
    assert any(rel[k + 1] < 10 * rel[k] ** 2 for k in range(1, it) if rel[k] < 1e-3)


# ---------------- substrate form (passive law, constant nodal load) ----------------

def test_passive_cell_is_the_cell_law_without_active_stress():
    # t_act(0) = 0: a passive cell at t = 1.3 equals an active one at t = 0
    m = ul_cell(k_bend=0.0, k_area=0.0, drag=0.0, ap_u=None, mu=1.0e6)
    Gp, Kp = O.assemble(O.make_cell(dict(m, passive=1, time=1.3)))
    G0, K0 = O.assemble(O.make_cell(dict(m, passive=0, time=0.0)))
    Ga, _ = O.assemble(O.make_cell(dict(m, passive=0, time=1.3)))
    np.testing.assert_array_equal(Gp, G0)
    np.testing.assert_array_equal(Kp, K0)
    assert np.abs(Ga - Gp).max() > 1e-6 * np.abs(Gp).max()


def test_external_load_enters_the_residual_only():
    m = ul_cell(mu=1.0e6)
    N = len(m["X"]) // 2
    f = np.random.default_rng(11).standard_normal(2 * N) * 1e3
    G1, K1 = O.assemble(O.make_cell(dict(m, passive=1, f_ext=f)))
    G0, K0 = O.assemble(O.make_cell(dict(m, passive=1)))
    np.testing.assert_allclose(G1 - G0, -f, rtol=0, atol=1e-12 * np.abs(G0).max())
    np.testing.assert_array_equal(K1, K0)
