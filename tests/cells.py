"""Shared synthetic cells for the parity tests: the same numpy data feeds the
GPU context (wholecellfe_amd.Context) and the oracle (oracle_py.make_cell)."""
import numpy as np

from wholecellfe_amd import mesh


def cst_cell(K=6, n_spr=40, seed=1, law_table=False, E=1000.0, k_lin=1000.0, jitter=0.05, time=1.0):
    rng = np.random.default_rng(seed)
    X, T, b = mesh.disk_rings(K, 10.0)
    X = mesh.jitter(X, T, b, jitter, rng)
    tri, bary, anchor = mesh.random_springs(X, T, n_spr, 10.0, rng)
    px, py, kk = mesh.stiffening_table()
    d = dict(X=X.reshape(-1), x=X.reshape(-1).copy(), tris=T.reshape(-1), law=1, E=E, nu=0.3, thick=0.2, time=time,
             spr_tri=tri, spr_bary=bary.reshape(-1), spr_anchor=anchor.reshape(-1),
             spr_state=np.ones(len(tri)), spring_law=1 if law_table else 0, k_lin=k_lin)
    if law_table:
        d.update(tab_px=px, tab_py=py, tab_k=kk)
    return d


def ul_cell(K=5, seed=2, table=True, time=1.3, deform=0.01):
    """Reference-model cell: UL membrane law and node-attached springs."""
    rng = np.random.default_rng(seed)
    X, T, b = mesh.disk_rings(K, 10.0)
    X = mesh.jitter(X, T, b, 0.05, rng)
    N = len(X)
    x = X + deform * rng.standard_normal(X.shape)
    ea = 0.02 * rng.standard_normal((N, 3))
    cC = rng.uniform(0.2, 1.0, N)
    fau = 0.004 * rng.standard_normal((N, 2))
    d = dict(X=X.reshape(-1), x=x.reshape(-1), tris=T.reshape(-1), law=0, mu=1000.0, thick=1.0, time=time,
             ea=ea.reshape(-1), node_springs=1, cC=cC, fau0=fau.reshape(-1), rho_max=100.0,
             spring_law=1 if table else 0, k_lin=1000.0)
    if table:
        # the kMD 6-knot table (UL_growth_explicit.cpp:1207-1214)
        d.update(tab_px=np.array([0., 2.51250583e-03, 4.40265201e-03, 5.91408228e-03, 11.76817302e-03, 17.001e-03]),
                 tab_py=np.array([-7.15514497, 71.80501308, 356.62420075, 356.96656556, 126.21811585,
                                  528.40481124]),
                 tab_k=np.array([31426.8556, 99485.86235, 67080.67086, -7908.39043, 21978.66923, 76858.39733]))
    return d


def add_to_context(ctx, d, table_id=0, kinetics=0, solver=0):
    """Register cell dict d on a wholecellfe_amd.Context (tables must be set
    before finalize)."""
    if "tab_px" in d:
        ctx.set_table(table_id, d["tab_px"], d["tab_py"], d["tab_k"])
    return ctx.add_cell(X=d["X"], x=d.get("x"), tris=d["tris"], law=d["law"], mu=d.get("mu", 1000.0),
                        E=d.get("E", 1000.0), nu=d.get("nu", 0.3), thick=d.get("thick", 0.2), ea=d.get("ea"),
                        node_springs=d.get("node_springs", 0), cC=d.get("cC"), fau=d.get("fau0"),
                        rho_max=d.get("rho_max", 100.0), spr_tri=d.get("spr_tri"), spr_bary=d.get("spr_bary"),
                        spr_anchor=d.get("spr_anchor"), spr_state=d.get("spr_state"),
                        spring_law=d.get("spring_law", 0), k_lin=d.get("k_lin", 1000.0), table_id=table_id,
                        kinetics=kinetics, solver=solver)
