"""Pin the FP64 oracle's reference restatement against golden vectors from the
compiled reference (tests/golden/*.npz, written by tests/golden/make_golden.py
with oracle/_ref/refharness_*).

Floating point: the restatement follows Eigen 3.3.4's operation structure but
not every summation order, so agreement is asserted to 1e-12 relative to the
array's max magnitude (measured agreement is ~1e-15).
"""
import numpy as np
import pytest

import oracle_py as O
from conftest import golden

RTOL = 1e-12


def close(a, b, rtol=RTOL):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    scale = max(np.max(np.abs(b)), 1e-300)
    err = np.max(np.abs(a - b)) / scale
    assert err <= rtol, f"max rel err {err:.3e}"
    return err


@pytest.mark.parametrize("case", ["ul_disk296_t1p3", "ul_disk483_t3", "ul_disk296_ea3d"])
def test_eval_element_re_matches_reference(case):
    g = golden(case)
    X, x, ea, cA, cC, tris = g["X"], g["x"], g["ea"], g["cA"], g["cC"], g["tris"]
    cen = O.area_tot(tris, x)
    close(cen, g["areatot"])
    T = len(tris)
    outs = {k: np.zeros((T, n)) for k, n in [("Re", 9), ("Re_ea", 18), ("MMe", 3), ("Re_Adif", 3), ("Re_Cdif", 3)]}
    for e in range(T):
        v = tris[e]
        r = O.eval_element_re(X[v], x[v], ea[v], cA[v], cC[v], float(g["dt"]), float(g["time"]), cen[1:])
        for k, o in zip(outs, r):
            outs[k][e] = o
    for k in outs:
        close(outs[k], g["elem_" + k])


@pytest.mark.parametrize("case", ["ul_disk296_t1p3", "ul_disk483_t3", "ul_disk296_ea3d"])
def test_calculate_res_matches_reference(case):
    g = golden(case)
    X, x, ea, cA, cC, tris = g["X"], g["x"], g["ea"], g["cA"], g["cC"], g["tris"]
    cen = O.area_tot(tris, x)
    Res, Res_ea, MM, RA, RC = O.calculate_res(tris, X, x, ea, cA, cC, float(g["dt"]), float(g["time"]), cen[1:])
    close(Res, g["res_Res"])
    close(Res_ea, g["res_Res_ea"])
    close(MM, g["res_MM"])
    close(RA, g["res_Res_Adif"])
    close(RC, g["res_Res_Cdif"])
    close(cen[1:], g["res_centroid"])


def test_calculate_res_substrate_matches_reference():
    """calculateRes, 10-argument overload with mat = 2 (UL:931-987, the
    substrate's call at UL:451)."""
    g = golden("calcres_subs")
    n = len(g["X"])
    Res, Res_ea, MM, _, _ = O.calculate_res(g["tris"], g["X"], g["x"], g["ea"], np.zeros(n), np.zeros(n),
                                            float(g["dt"]), float(g["time"]), np.array([1e30, 0.0, 0.0]), mat=2)
    close(Res, g["res_Res"])
    close(Res_ea, g["res_Res_ea"])
    close(MM, g["res_MM"])


def test_substrate_law_is_the_passive_cell_law_scaled():
    """mat = 2 is mat = 1 with mu x 1000 and no active stress: with t_act(0) =
    0 the cell law at time 0 times 1000 is the substrate law at any time (the
    stress is linear in mu; UL:721-763)."""
    g = golden("calcres_subs")
    n = len(g["X"])
    z, far = np.zeros(n), np.array([1e30, 0.0, 0.0])
    R2, Rea2, MM2, _, _ = O.calculate_res(g["tris"], g["X"], g["x"], g["ea"], z, z, 0.005, 7.0, far, mat=2)
    R1, Rea1, MM1, _, _ = O.calculate_res(g["tris"], g["X"], g["x"], g["ea"], z, z, 0.005, 0.0, far, mat=1)
    np.testing.assert_allclose(R2, 1000.0 * R1, rtol=0, atol=1e-12 * np.abs(R2).max())
    np.testing.assert_array_equal(Rea2, Rea1)
    np.testing.assert_array_equal(MM2, MM1)
    R1t, _, _, _, _ = O.calculate_res(g["tris"], g["X"], g["x"], g["ea"], z, z, 0.005, 7.0, far, mat=1)
    assert np.abs(1000.0 * R1t - R2).max() > 1e-3 * np.abs(R2).max()  # the cell law at t = 7 is active


@pytest.mark.parametrize("case,law", [("forcefa_k1", 0), ("forcefa_k31", 0), ("forcefa_kMD", 1),
                                      ("forcefa_kMD_3d", 1)])
def test_force_fa_matches_reference(case, law):
    g = golden(case)
    F = O.force_fa(g["tris"], g["x"], g["cC"], g["fau"], law, float(g["kint"]))
    close(F, g["F_FA"])


@pytest.mark.parametrize("case,law", [("updatefa_k1_rigid", 0), ("updatefa_k31_moving", 0),
                                      ("updatefa_kMD_rigid", 1), ("updatefa_kMD_moving", 1)])
def test_update_fa_matches_reference(case, law):
    g = golden(case)
    cA, cC, fau, unloc = O.update_fa(g["tris"], g["X"], g["x"], g["cA"], g["cC"], g["fau"], float(g["dt"]),
                                     g["Stris"], g["SX"], g["Sx"], law, float(g["kint"]))
    assert unloc == 0
    close(cA, g["out_cA"])
    close(cC, g["out_cC"])
    close(fau, g["out_fau"])


@pytest.mark.parametrize("case,law", [("forcesubs_k31", 0), ("forcesubs_kMD", 1)])
def test_force_fa_subs_matches_reference(case, law):
    g = golden(case)
    F = O.force_fa_subs(g["tris"], g["x"], g["cC"], g["fau"], g["Stris"], g["Sx"], law, float(g["kint"]))
    close(F, g["F_subs"])


@pytest.mark.parametrize("case,law", [("forcesubs_k31", 0), ("forcesubs_kMD", 1)])
def test_force_fa_subs_balances_the_cell_force(case, law):
    """Action and reaction: with every anchor strictly inside one substrate
    triangle, the substrate's total force is minus the cell's
    (calculateForce_FA, pinned separately), and its first moment about the
    origin is minus sum_i F_FA[i] (x) p_i at the anchors p_i = x_i - fa_u_i,
    since barycentric weights reproduce the point they split."""
    g = golden(case)
    cC = g["cC"].copy()
    cC[0] = 0.0  # node 0 sits on a substrate vertex
    Fs = O.force_fa_subs(g["tris"], g["x"], cC, g["fau"], g["Stris"], g["Sx"], law, float(g["kint"])).reshape(-1, 3)
    Fc = O.force_fa(g["tris"], g["x"], cC, g["fau"], law, float(g["kint"])).reshape(-1, 3)
    p = g["x"] - g["fau"]
    scale = np.abs(Fc).sum()
    np.testing.assert_allclose(Fs.sum(axis=0), -Fc.sum(axis=0), rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(g["Sx"].T @ Fs, -(p.T @ Fc), rtol=0, atol=1e-12 * scale * np.abs(p).max())


def test_force_fa_subs_loads_every_triangle_at_a_vertex():
    """An anchor on a substrate vertex loads each triangle around it (no early
    exit, UL:1330-1395): the vertex receives valence x the single-hit force
    and the other vertices nothing."""
    g = golden("forcesubs_kMD")
    cC = np.zeros(len(g["x"]))
    cC[0] = 1.0
    Fs = O.force_fa_subs(g["tris"], g["x"], cC, g["fau"], g["Stris"], g["Sx"], 1, -1.0).reshape(-1, 3)
    Fc = O.force_fa(g["tris"], g["x"], cC, g["fau"], 1, -1.0).reshape(-1, 3)
    p = g["x"][0] - g["fau"][0]
    v = int(np.argmin(np.sum((g["Sx"] - p) ** 2, axis=1)))
    valence = int((g["Stris"] == v).any(axis=1).sum())
    assert valence >= 4
    np.testing.assert_allclose(Fs[v], -valence * Fc[0], rtol=1e-9)
    others = np.delete(Fs, v, axis=0)
    assert np.abs(others).max() < 1e-9 * np.abs(Fc[0]).max()


def test_golden_inputs_exercise_the_md_table_segments():
    """The force-law fixtures must cover every MD-table segment, the
    extrapolation beyond the last knot and D == 0 (UL:1247-1264)."""
    g = golden("forcefa_kMD")
    D = np.linalg.norm(g["fau"], axis=1)
    px = [0., 2.51250583e-03, 4.40265201e-03, 5.91408228e-03, 11.76817302e-03, 17.001e-03]
    seg = np.searchsorted(px, D, side="right")
    assert set(range(1, 7)) <= set(seg.tolist())
    assert (D == 0).any()
