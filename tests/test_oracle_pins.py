"""Pins for the oracle pieces that the golden vectors and the finite-difference
tests leave open: the isotropic active stress of the CST law (the benchmark's
only load), the barycentric spring residual, the Bell off-rate under load,
and the shared FP64 exp of the kinetics.  Each check comes from mathematics
that does not restate the oracle's formula: the divergence theorem, a
brute-force table scan in Python, the closed-form transition probability and
the law of large numbers, and the C library's exp.
"""
import ctypes as C
import math

import numpy as np
import pytest

import cells
import oracle_py as O
from wholecellfe_amd import mesh


def _boundary_half_normals(X, T):
    """For each node, half the sum over its boundary edges of L * n (outward
    unit normal n): the closed form of sum_e A_e grad N_a (divergence
    theorem; zero at interior nodes)."""
    N = len(X)
    edges = {}
    for t in T:
        for a in range(3):
            p, q = int(t[a]), int(t[(a + 1) % 3])
            key = (min(p, q), max(p, q))
            edges.setdefault(key, []).append((p, q))
    out = np.zeros((N, 2))
    for key, occ in edges.items():
        if len(occ) != 1:
            continue
        p, q = occ[0]  # counter-clockwise triangle: the domain is on the left of p -> q
        dx, dy = X[q] - X[p]
        out[p] += 0.5 * np.array([dy, -dx])
        out[q] += 0.5 * np.array([dy, -dx])
    return out


@pytest.mark.parametrize("law,time", [(O.LAW_CST, 1.3), (O.LAW_CST, 2.5), (O.LAW_UL, 0.7), (O.LAW_UL, 3.0)])
def test_active_stress_is_uniform_tension(law, time):
    """At the reference configuration (no strain) the only stress is the
    isotropic active stress s = kappa t_act, kappa = 0.5, t_act = 200 t below
    t = 2 and 400 beyond (UL:743-757, 872-878).  A uniform stress s I in a
    membrane of thickness h gives G_a = h s sum_e A_e grad N_a: zero at
    interior nodes and h s (L1 n1 + L2 n2) / 2 at a boundary node."""
    rng = np.random.default_rng(7)
    X, T, b = mesh.disk_rings(6, 10.0)
    X = mesh.jitter(X, T, b, 0.08, rng)
    thick = 0.2
    d = dict(X=X.reshape(-1), x=X.reshape(-1).copy(), tris=T.reshape(-1), law=law, E=2300.0, nu=0.3, mu=1000.0,
             thick=thick, time=time)
    if law == O.LAW_UL:
        d["ea"] = np.zeros(3 * len(X))
    G, _ = O.assemble(O.make_cell(d), want_K=False)
    t_act = 200.0 * time if time < 2.0 else 400.0
    expect = thick * 0.5 * t_act * _boundary_half_normals(X, T)
    G = G.reshape(-1, 2)
    assert np.max(np.abs(G - expect)) <= 1e-12 * np.max(np.abs(expect))
    assert np.all(np.abs(G[~b]) <= 1e-12 * np.max(np.abs(expect)))
    # the substrate form (passive) carries no active stress
    G0, _ = O.assemble(O.make_cell(dict(d, passive=1)), want_K=False)
    assert np.max(np.abs(G0)) <= 1e-12 * np.max(np.abs(expect))


def _table_force_bruteforce(px, py, k, D):
    """F(D) of the (px, py, k) table by a plain scan: segment j with
    px_j <= D < px_{j+1}, the last segment extrapolating."""
    j = 0
    for i in range(len(px)):
        if px[i] <= D:
            j = i
    return py[j] + k[j] * (D - px[j])


@pytest.mark.parametrize("law", ["linear", "table"])
def test_barycentric_spring_residual(law):
    """One triangle at its rest positions (CST, t = 0: no stress), one spring
    at barycentric weights (0.2, 0.3, 0.5) with extension u: the spring force
    F(|u|) u/|u| pulls the point back, so G_a = b_a F(|u|) u/|u| -- each node
    with its own weight (a transposed or shifted weight fails)."""
    px, py, kk = mesh.stiffening_table()
    X = np.array([[0.0, 0.0], [1.0, 0.1], [0.3, 0.9]])
    bary = np.array([0.2, 0.3, 0.5])
    p = bary @ X
    for u0 in ([0.004, -0.003], [0.0123, 0.004], [-1e-4, 2e-4]):
        u0 = np.array(u0)
        d = dict(X=X.reshape(-1), x=X.reshape(-1).copy(), tris=np.array([0, 1, 2]), law=O.LAW_CST, E=1000.0,
                 nu=0.3, thick=0.2, time=0.0, spr_tri=np.array([0]), spr_bary=bary, spr_anchor=p - u0,
                 spr_state=np.ones(1), spring_law=0 if law == "linear" else 1, k_lin=1000.0)
        if law == "table":
            d.update(tab_px=px, tab_py=py, tab_k=kk)
        G, _ = O.assemble(O.make_cell(d), want_K=False)
        D = np.linalg.norm(u0)
        F = 1000.0 * D if law == "linear" else _table_force_bruteforce(px, py, kk, D)
        expect = np.outer(bary, F * u0 / D).reshape(-1)
        assert np.max(np.abs(G - expect)) <= 1e-12 * np.max(np.abs(expect)), (u0, G, expect)
        # an unbound spring exerts nothing
        G0, _ = O.assemble(O.make_cell(dict(d, spr_state=np.zeros(1))), want_K=False)
        assert np.max(np.abs(G0)) == 0.0


def test_stiffening_table_law_matches_bruteforce_and_is_continuous():
    """The 64-knot table: the oracle's F(D) equals a brute-force scan at
    random extensions (inside and beyond the table) and at every knot; F is
    continuous across the knots (py_{j+1} = py_j + k_j dx) and increasing."""
    px, py, kk = mesh.stiffening_table()
    d = cells.cst_cell(K=2, n_spr=3, law_table=True)
    c = O.make_cell(d)
    rng = np.random.default_rng(3)
    for D in np.concatenate([rng.uniform(0.0, 0.013, 400), px]):
        f, _ = O.spring_law(c, float(D))
        ref = _table_force_bruteforce(px, py, kk, float(D))
        assert abs(f - ref) <= 1e-13 * max(abs(ref), 1.0)
    eps = 1e-12
    kmax = kk.max()
    for j in range(1, len(px)):
        fl, _ = O.spring_law(c, float(px[j] - eps))
        fr, _ = O.spring_law(c, float(px[j] + eps))
        assert abs(fr - fl) <= 2.5 * kmax * eps + 1e-12
    Ds = np.linspace(0.0, 0.012, 997)
    Fs = [O.spring_law(c, float(D))[0] for D in Ds]
    assert np.all(np.diff(Fs) > 0)
    # the stiffness rises from 1 to 31 pN/nm over 0-10 nm (in pN/um)
    assert kk[0] == 1000.0 and kk[-1] == 31000.0 and px[-1] == pytest.approx(0.01)


def _loaded_bond_cell(n, u0, k_lin=1000.0):
    """n bound springs on one triangle, all at the same attachment point and
    extension u0 (linear law): every spring carries |F| = k |u0|."""
    X = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
    bary = np.tile([0.25, 0.25, 0.5], n)
    p = np.array([0.25, 0.5])
    anchor = np.tile(p - np.asarray(u0), n)
    return dict(X=X.reshape(-1), x=X.reshape(-1).copy(), tris=np.array([0, 1, 2]), law=O.LAW_CST, E=1000.0, nu=0.3,
                thick=0.2, time=0.0, spr_tri=np.zeros(n, dtype=np.int32), spr_bary=bary, spr_anchor=anchor,
                spr_state=np.ones(n), spring_law=0, k_lin=k_lin)


@pytest.mark.parametrize("F0", [0.0, 4.0, 8.0])
def test_bell_off_rate_under_load(F0):
    """Bell model at a fixed load F0: p_off = 1 - exp(-k_off0 e^{F0/f_b} dt).
    Every bond unbinds exactly when its Philox uniform u(key; replica, bond,
    step) is below p_off, and over 20,000 bonds the unbound fraction is p_off
    within 5 standard deviations (a dropped or sign-flipped load factor, a
    wrong dt or a swapped state fails both)."""
    n, dt, k_off0, f_b, k_on = 20000, 1.0, 0.1, 4.0, 1.0
    key, replica, step = (12345, 678), 3, 9
    u0 = np.array([0.6, -0.8]) * (F0 / 1000.0)
    d = _loaded_bond_cell(n, u0)
    st, an, nt, margin = O.bell_update(O.make_cell(d), d["spr_state"], d["spr_anchor"], dt, k_on, k_off0, f_b, key,
                                       replica, step)
    p = 1.0 - math.exp(-k_off0 * math.exp(F0 / f_b) * dt)
    u = np.array([O.uniform(key[0], key[1], replica, s, step) for s in range(n)])
    decided = np.abs(u - p) > 1e-12
    assert np.array_equal((st == 0.0)[decided], (u < p)[decided])
    assert nt == int(np.sum(st == 0.0))
    frac = np.mean(st == 0.0)
    assert abs(frac - p) <= 5.0 * math.sqrt(p * (1 - p) / n)
    assert np.array_equal(an, d["spr_anchor"])  # unbinding keeps the anchor
    assert margin == pytest.approx(np.min(np.abs(u - p)), abs=1e-12)  # F = k |u0| up to rounding of u0


def test_bell_binding_reanchors_at_the_current_point():
    """Unbound bonds bind with p_on = 1 - exp(-k_on dt) and are re-anchored at
    the current attachment point (rest length 0 afterwards)."""
    n, dt, k_on = 20000, 0.3, 1.0
    key, replica, step = (7, 0), 1, 4
    d = _loaded_bond_cell(n, [0.003, 0.0])
    d["spr_state"] = np.zeros(n)
    st, an, nt, _ = O.bell_update(O.make_cell(d), d["spr_state"], d["spr_anchor"], dt, k_on, 0.1, 4.0, key, replica,
                                  step)
    p = 1.0 - math.exp(-k_on * dt)
    u = np.array([O.uniform(key[0], key[1], replica, s, step) for s in range(n)])
    assert np.array_equal(st == 1.0, u < p)
    bound = st == 1.0
    an = an.reshape(-1, 2)
    assert np.all(an[bound] == np.array([0.25, 0.5]))
    assert np.array_equal(an[~bound], d["spr_anchor"].reshape(-1, 2)[~bound])
    assert abs(np.mean(bound) - p) <= 5.0 * math.sqrt(p * (1 - p) / n)


def _orc_exp(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(x)
    O.lib().orc_exp_array(C.c_int(len(x)), O.dp(x), O.dp(y))
    return y


def _ulp_distance(a, b):
    ia = a.view(np.int64).astype(np.int64)
    ib = b.view(np.int64).astype(np.int64)
    return np.abs(ia - ib)


def test_shared_exp_within_one_ulp_of_libc():
    """The kinetics' exp (Cody-Waite reduction, degree-13 Taylor polynomial)
    against the C library's exp: at most 1 ulp apart over the whole normal
    range, the kinetics' range, and the subnormal tail; exact at 0; the
    limits overflow and underflow as exp does."""
    rng = np.random.default_rng(11)
    x = np.concatenate([rng.uniform(-708.0, 709.7, 300000), rng.uniform(-40.0, 40.0, 300000),
                        rng.uniform(-1e-3, 1e-3, 50000), rng.uniform(-745.0, -708.5, 20000),
                        np.array([1.0, -1.0, 0.5, math.log(2.0), -math.log(2.0), 709.78, -708.39, 1e-300, -1e-300])])
    y = _orc_exp(x)
    ref = np.array([math.exp(v) for v in x])
    normal = ref >= 2.2250738585072014e-308
    assert np.max(_ulp_distance(y[normal], ref[normal])) <= 1
    # subnormal results: within 1 ulp of the subnormal spacing
    assert np.max(np.abs(y[~normal] - ref[~normal])) <= 5e-324
    sp = _orc_exp(np.array([0.0, -0.0, 710.5, 1000.0, -746.5, -1e4, np.inf, -np.inf, np.nan]))
    assert sp[0] == 1.0 and sp[1] == 1.0
    assert np.isinf(sp[2]) and np.isinf(sp[3]) and np.isinf(sp[6])
    assert sp[4] == 0.0 and sp[5] == 0.0 and sp[7] == 0.0
    assert np.isnan(sp[8])


def test_shared_exp_error_against_exact_values():
    """Error against exp computed exactly (60 significant digits): below one
    ulp at 2,000 points spread over the kinetics' arguments."""
    from decimal import Decimal, getcontext
    getcontext().prec = 60
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.uniform(-30.0, 30.0, 1500), rng.uniform(-0.4, 0.4, 500)])
    y = _orc_exp(x)
    worst = 0.0
    for xi, yi in zip(x, y):
        ex = Decimal(float(xi)).exp()
        ulp = math.ulp(float(ex))
        worst = max(worst, abs(float((Decimal(float(yi)) - ex) / Decimal(ulp))))
    assert worst < 1.0, worst
