"""ctypes binding of oracle/liboracle.so (the FP64 CPU oracle) for the tests.

Test infrastructure: the product never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(REPO, "oracle", "liboracle.so")

_lib = None

D = C.POINTER(C.c_double)
I32 = C.POINTER(C.c_int32)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "oracle"], check=True,
                           stdout=subprocess.DEVNULL)
        _lib = C.CDLL(LIB_PATH)
    return _lib


def dp(a):
    return a.ctypes.data_as(D) if a is not None else None


def ip(a):
    return a.ctypes.data_as(I32) if a is not None else None


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Cell(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int32), ("n_tris", C.c_int32),
        ("X", D), ("x", D), ("tris", I32),
        ("law", C.c_int32),
        ("mu", C.c_double), ("E", C.c_double), ("nu", C.c_double), ("thick", C.c_double), ("time", C.c_double),
        ("ea", D),
        ("node_springs", C.c_int32), ("cC", D), ("fau0", D), ("rho_max", C.c_double),
        ("n_spr", C.c_int32), ("spr_tri", I32), ("spr_bary", D), ("spr_anchor", D), ("spr_state", D),
        ("spring_law", C.c_int32), ("k_lin", C.c_double),
        ("tab_n", C.c_int32), ("tab_px", D), ("tab_py", D), ("tab_k", D),
        ("k_bend", C.c_double), ("k_area", C.c_double), ("area_0", C.c_double), ("drag", C.c_double),
        ("dt", C.c_double), ("ap_u", D), ("ap_rate", C.c_double), ("ap_quantum", C.c_double),
        ("ap_density", C.c_double), ("passive", C.c_int32), ("f_ext", D),
    ]


LAW_UL, LAW_CST = 0, 1
SPRING_LINEAR, SPRING_TABLE = 0, 1


def make_cell(model):
    """Build an oracle Cell from a dict; keeps the numpy arrays alive on the
    returned object (attribute `_keep`)."""
    keep = {}

    def arr(name, conv):
        v = model.get(name)
        if v is None:
            return None
        a = conv(v)
        keep[name] = a
        return a

    X = arr("X", f64)
    x = arr("x", f64)
    tris = arr("tris", i32)
    c = Cell()
    c.n_nodes = len(X) // 2
    c.n_tris = len(tris) // 3
    c.X, c.x, c.tris = dp(X), dp(x), ip(tris)
    c.law = model.get("law", LAW_CST)
    c.mu = model.get("mu", 1000.0)
    c.E = model.get("E", 1000.0)
    c.nu = model.get("nu", 0.3)
    c.thick = model.get("thick", 0.2)
    c.time = model.get("time", 1.0)
    c.ea = dp(arr("ea", f64))
    c.node_springs = model.get("node_springs", 0)
    c.cC = dp(arr("cC", f64))
    c.fau0 = dp(arr("fau0", f64))
    c.rho_max = model.get("rho_max", 100.0)
    spr_tri = arr("spr_tri", i32)
    c.n_spr = 0 if spr_tri is None else len(spr_tri)
    c.spr_tri = ip(spr_tri)
    c.spr_bary = dp(arr("spr_bary", f64))
    c.spr_anchor = dp(arr("spr_anchor", f64))
    c.spr_state = dp(arr("spr_state", f64))
    c.spring_law = model.get("spring_law", SPRING_LINEAR)
    c.k_lin = model.get("k_lin", 1000.0)
    px = arr("tab_px", f64)
    c.tab_n = 0 if px is None else len(px)
    c.tab_px, c.tab_py, c.tab_k = dp(px), dp(arr("tab_py", f64)), dp(arr("tab_k", f64))
    c.k_bend = model.get("k_bend", 0.0)
    c.k_area = model.get("k_area", 0.0)
    c.area_0 = model.get("area_0", 10 * 10 * 3.1416)
    c.drag = model.get("drag", 0.0)
    c.dt = model.get("dt", 0.01)
    c.ap_u = dp(arr("ap_u", f64))
    c.ap_rate = model.get("ap_rate", 10.0)
    c.ap_quantum = model.get("ap_quantum", 5.0)
    c.ap_density = model.get("ap_density", 100.0)
    c.passive = model.get("passive", 0)
    c.f_ext = dp(arr("f_ext", f64))
    c._keep = keep
    return c


# ---------------- Part 1 wrappers ----------------

def eval_element_re(X, x, ea, cA, cC, dt, time, centroid, mat=1):
    L = lib()
    out = [np.zeros(k) for k in (9, 18, 3, 3, 3)]
    L.orc_eval_element_re(dp(f64(X)), dp(f64(x)), dp(f64(ea)), dp(f64(cA)), dp(f64(cC)), C.c_double(dt),
                          C.c_double(time), dp(f64(centroid)), C.c_int(mat), *[dp(o) for o in out])
    return out


def area_tot(tris, x):
    out = np.zeros(4)
    lib().orc_area_tot(C.c_int(len(tris)), ip(i32(tris)), dp(f64(x)), dp(out))
    return out


def calculate_res(tris, X, x, ea, cA, cC, dt, time, centroid, mat=1):
    n = len(X)
    outs = [np.zeros(3 * n), np.zeros(6 * n), np.zeros(n), np.zeros(n), np.zeros(n)]
    lib().orc_calculate_res(C.c_int(n), C.c_int(len(tris)), ip(i32(tris)), dp(f64(X)), dp(f64(x)), dp(f64(ea)),
                            dp(f64(cA)), dp(f64(cC)), C.c_double(dt), C.c_double(time), dp(f64(centroid)),
                            C.c_int(mat), *[dp(o) for o in outs])
    return outs


def force_fa(tris, x, cC, fau, law, kint):
    n = len(x)
    F = np.zeros(3 * n)
    lib().orc_force_fa(C.c_int(n), C.c_int(len(tris)), ip(i32(tris)), dp(f64(x)), dp(f64(cC)), dp(f64(fau)),
                       C.c_int(law), C.c_double(kint), dp(F))
    return F


def update_fa(tris, X, x, cA, cC, fau, dt, stris, SX, Sx, law, kint):
    cA, cC, fau = f64(cA).copy(), f64(cC).copy(), f64(fau).copy()
    unloc = C.c_int(0)
    lib().orc_update_fa(C.c_int(len(X)), C.c_int(len(tris)), ip(i32(tris)), dp(f64(X)), dp(f64(x)), dp(cA), dp(cC),
                        dp(fau), C.c_double(dt), C.c_int(len(SX)), C.c_int(len(stris)), ip(i32(stris)),
                        dp(f64(SX)), dp(f64(Sx)), C.c_int(law), C.c_double(kint), C.byref(unloc))
    return cA, cC, fau, unloc.value


def force_fa_subs(tris, x_cell, cC, fau, stris, Sx, law, kint):
    n = len(x_cell)
    ns = len(Sx)
    F = np.zeros(3 * ns)
    lib().orc_force_fa_subs(C.c_int(n), C.c_int(len(tris)), ip(i32(tris)), dp(f64(x_cell)), dp(f64(cC)),
                            dp(f64(fau)), C.c_int(ns), C.c_int(len(stris)), ip(i32(stris)), dp(f64(Sx)),
                            C.c_int(law), C.c_double(kint), dp(F))
    return F


def boundary_candidates(n_nodes, tris):
    tris = i32(tris).reshape(-1)
    T = len(tris) // 3
    isb = np.zeros(n_nodes, dtype=np.int32)
    st = np.zeros(n_nodes + 1, dtype=np.int32)
    cand = np.zeros(12 * max(T, 1), dtype=np.int32)
    m = lib().orc_boundary_candidates(C.c_int(n_nodes), C.c_int(T), ip(tris), ip(isb), ip(st), ip(cand))
    return isb, st, cand[:2 * m].reshape(-1, 2)


def bending_force_2d(tris, x, k_bend=20.0):
    n = len(x)
    F = np.zeros(3 * n)
    nopen = lib().orc_bending_force_2d(C.c_int(n), C.c_int(len(tris)), ip(i32(tris)), dp(f64(x)),
                                       C.c_double(k_bend), dp(F))
    return F, nopen


def area_constraint_force(tris, x, tot_area, k_area=1.0, area_0=10 * 10 * 3.1416):
    n = len(x)
    F = np.zeros(3 * n)
    nopen = lib().orc_area_constraint_force(C.c_int(n), C.c_int(len(tris)), ip(i32(tris)), dp(f64(x)),
                                            C.c_double(tot_area), C.c_double(k_area), C.c_double(area_0), dp(F))
    return F, nopen


def force_ap(tris, x, dt, u, rate=10.0, quantum=5.0, density=100.0):
    n = len(x)
    F = np.zeros(3 * n)
    nopen = lib().orc_force_ap(C.c_int(n), C.c_int(len(tris)), ip(i32(tris)), dp(f64(x)), C.c_double(dt),
                               C.c_double(rate), C.c_double(quantum), C.c_double(density), dp(f64(u)), dp(F))
    return F, nopen


# ---------------- Part 2 wrappers ----------------

def color_elements(n_nodes, tris):
    tris = i32(tris).reshape(-1)
    col = np.zeros(len(tris) // 3, dtype=np.int32)
    nc = lib().orc_color_elements(C.c_int(n_nodes), C.c_int(len(col)), ip(tris), ip(col))
    return col, nc


def assemble(cell, color=None, n_colors=0, want_K=True):
    n = 2 * cell.n_nodes
    G = np.zeros(n)
    K = np.zeros((n, n)) if want_K else None
    lib().orc_assemble(C.byref(cell), ip(color) if color is not None else None, C.c_int(n_colors), dp(G),
                       dp(K) if K is not None else None)
    return G, K


def element_system(cell, e):
    Ge, Ke = np.zeros(6), np.zeros(36)
    lib().orc_element_system(C.byref(cell), C.c_int(e), None, dp(Ge), dp(Ke))
    return Ge, Ke.reshape(6, 6)


def lu_factor(A, pivoting=1, f32=False):
    A = np.ascontiguousarray(A, dtype=np.float32 if f32 else np.float64).copy()
    n = A.shape[0]
    ipiv = np.zeros(n, dtype=np.int32)
    if f32:
        r = lib().orc_lu_factor_f32(C.c_int(n), A.ctypes.data_as(C.POINTER(C.c_float)), ip(ipiv), C.c_int(pivoting))
    else:
        r = lib().orc_lu_factor(C.c_int(n), dp(A), ip(ipiv), C.c_int(pivoting))
    return A, ipiv, r


def lu_solve(LU, ipiv, b, pivoting=1):
    b = f64(b).copy()
    lib().orc_lu_solve(C.c_int(len(b)), dp(f64(LU)), ip(i32(ipiv)), C.c_int(pivoting), dp(b))
    return b


def pcg_csr(rowptr, col, val, b, x0=None, rtol=1e-10, maxit=10000):
    n = len(b)
    x = np.zeros(n) if x0 is None else f64(x0).copy()
    it = lib().orc_pcg_csr(C.c_int(n), ip(i32(rowptr)), ip(i32(col)), dp(f64(val)), dp(f64(b)), dp(x),
                           C.c_double(rtol), C.c_int(maxit))
    return x, it


def newton_solve(cell, x0, rtol=1e-10, atol=1e-12, maxit=20):
    x = f64(x0).copy()
    hist = np.zeros(maxit + 1)
    it = lib().orc_newton_solve(C.byref(cell), dp(x), C.c_double(rtol), C.c_double(atol), C.c_int(maxit), dp(hist))
    return x, it, hist


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def uniform(key_lo, key_hi, a, b, c):
    f = lib().orc_uniform
    f.restype = C.c_double
    return f(C.c_uint32(key_lo), C.c_uint32(key_hi), C.c_uint32(a), C.c_uint32(b), C.c_uint32(c))


def bell_update(cell, state, anchor, dt, k_on, k_off0, f_b, key, replica, step):
    state = f64(state).copy()
    anchor = f64(anchor).copy()
    mm = C.c_double(0)
    nt = lib().orc_bell_update(C.byref(cell), dp(state), dp(anchor), C.c_double(dt), C.c_double(k_on),
                               C.c_double(k_off0), C.c_double(f_b), C.c_uint32(key[0]), C.c_uint32(key[1]),
                               C.c_uint32(replica), C.c_uint32(step), C.byref(mm))
    return state, anchor, nt, mm.value


def spring_law(cell, Dv):
    f, df = C.c_double(0), C.c_double(0)
    lib().orc_spring_law(C.byref(cell), C.c_double(Dv), C.byref(f), C.byref(df))
    return f.value, df.value


def commit_ul_strain(cell):
    ea = np.zeros(3 * cell.n_nodes)
    lib().orc_commit_ul_strain(C.byref(cell), dp(ea))
    return ea


# ---------------- Part 2, sparse forms ----------------

def csr_pattern(n_nodes, tris):
    tris = i32(tris).reshape(-1)
    L = lib()
    nnz = L.orc_csr_pattern(C.c_int(n_nodes), C.c_int(len(tris) // 3), ip(tris), None, None, C.c_int(0))
    rowptr = np.zeros(2 * n_nodes + 1, dtype=np.int32)
    col = np.zeros(nnz, dtype=np.int32)
    L.orc_csr_pattern(C.c_int(n_nodes), C.c_int(len(tris) // 3), ip(tris), ip(rowptr), ip(col), C.c_int(nnz))
    return rowptr, col


def assemble_csr(cell, rowptr, col):
    val = np.zeros(len(col))
    G = np.zeros(2 * cell.n_nodes)
    rc = lib().orc_assemble_csr(C.byref(cell), ip(rowptr), ip(col), dp(val), dp(G))
    if rc != 0:
        raise ValueError("cell not supported by the CSR assembly")
    return G, val


def newton_solve_pcg(cell, x0, rtol=1e-12, atol=0.0, maxit=30, cg_rtol=1e-13, cg_maxit=200000):
    x = f64(x0).copy()
    hist = np.zeros(maxit + 1)
    it = lib().orc_newton_solve_pcg(C.byref(cell), dp(x), C.c_double(rtol), C.c_double(atol), C.c_int(maxit),
                                    C.c_double(cg_rtol), C.c_int(cg_maxit), dp(hist))
    return x, it, hist


def band_width(n_nodes, tris):
    tris = i32(tris).reshape(-1)
    order = np.zeros(n_nodes, dtype=np.int32)
    w = lib().orc_band_width(C.c_int(n_nodes), C.c_int(len(tris) // 3), ip(tris), ip(order))
    return w, order


def band_newton_step(cell, x0):
    x = f64(x0).copy()
    rc = lib().orc_band_newton_step(C.byref(cell), dp(x))
    return x, rc


def batch_band_newton_step(cells_, xs, n_threads=1):
    """One Newton iteration per cell (FP64 RCM band LU), OpenMP over cells;
    returns (seconds of the timed iterations, failures); xs updated in place."""
    L = lib()
    L.orc_batch_band_newton_step.restype = C.c_double
    arr = (Cell * len(cells_))(*cells_)
    ptrs = (D * len(xs))(*[dp(x) for x in xs])
    err = C.c_int(0)
    t = L.orc_batch_band_newton_step(C.c_int(len(cells_)), arr, ptrs, C.c_int(n_threads), C.byref(err))
    return t, err.value
