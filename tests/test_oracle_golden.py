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


def test_golden_inputs_exercise_the_md_table_segments():
    """The force-law fixtures must cover every MD-table segment, the
    extrapolation beyond the last knot and D == 0 (UL:1247-1264)."""
    g = golden("forcefa_kMD")
    D = np.linalg.norm(g["fau"], axis=1)
    px = [0., 2.51250583e-03, 4.40265201e-03, 5.91408228e-03, 11.76817302e-03, 17.001e-03]
    seg = np.searchsorted(px, D, side="right")
    assert set(range(1, 7)) <= set(seg.tolist())
    assert (D == 0).any()
