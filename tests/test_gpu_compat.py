"""The reference-signature adapters (include/wcfe_compat.hpp), called from a
C++ program shaped like the reference's call sites
(wholecellfe_amd/_build/compat_check, built from tests/compat_check.cpp),
reproduce the compiled reference's golden vectors on the GPU."""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(REPO, "wholecellfe_amd", "_build", "compat_check")


def run(mode, arrays, meta):
    if not os.path.exists(BIN):
        pytest.fail("compat_check is not built (python -m wholecellfe_amd.build)")
    d = tempfile.mkdtemp(prefix="wcfe_compat_")
    try:
        for k, v in arrays.items():
            a = np.ascontiguousarray(v)
            a = a.astype(np.int32) if a.dtype.kind in "iu" else a.astype(np.float64)
            a.tofile(os.path.join(d, k + ".bin"))
        np.asarray(meta, dtype=np.float64).tofile(os.path.join(d, "meta.bin"))
        r = subprocess.run([BIN, d, mode], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stderr
        return {f[4:-4]: np.fromfile(os.path.join(d, f)) for f in os.listdir(d) if f.startswith("out_")}
    finally:
        shutil.rmtree(d)


def relerr(a, b):
    return np.max(np.abs(np.ravel(a) - np.ravel(b))) / np.max(np.abs(b))


def test_calculate_res_adapter():
    g = golden("ul_disk483_t3")
    o = run("calcres", dict(X=g["X"], x=g["x"], ea=g["ea"], tris=g["tris"]), [float(g["time"]), 0.005, 0, 0])
    assert relerr(o["Res"], g["res_Res"]) < 1e-10
    assert relerr(o["Res_ea"], g["res_Res_ea"]) < 1e-10
    assert relerr(o["MM"], g["res_MM"]) < 1e-12


def test_calculate_res_substrate_adapter():
    g = golden("calcres_subs")
    o = run("calcressubs", dict(X=g["X"], x=g["x"], ea=g["ea"], tris=g["tris"]), [float(g["time"]), 0.005, 0, 0])
    assert relerr(o["Res"], g["res_Res"]) < 1e-10
    assert relerr(o["Res_ea"], g["res_Res_ea"]) < 1e-10
    assert relerr(o["MM"], g["res_MM"]) < 1e-12


@pytest.mark.parametrize("case,law", [("forcefa_k31", 0), ("forcefa_kMD", 1)])
def test_force_fa_adapter(case, law):
    g = golden(case)
    o = run("forcefa", dict(X=g["X"], x=g["x"], cC=g["cC"], fau=g["fau"], tris=g["tris"]),
            [0.0, 0.0, float(g["kint"]), law])
    assert relerr(o["F_FA"], g["F_FA"]) < 1e-12


@pytest.mark.parametrize("case,law", [("updatefa_k1_rigid", 0), ("updatefa_kMD_moving", 1)])
def test_update_fa_adapter(case, law):
    g = golden(case)
    o = run("updatefa", dict(X=g["X"], x=g["x"], cA=g["cA"], cC=g["cC"], fau=g["fau"], SX=g["SX"], Sx=g["Sx"],
                             tris=g["tris"], Stris=g["Stris"]), [0.0, float(g["dt"]), float(g["kint"]), law])
    assert relerr(o["cC"], g["out_cC"]) < 1e-12
    assert relerr(o["fau"], g["out_fau"]) < 1e-11


@pytest.mark.parametrize("case", ["bound_disk296", "bound_disk296_notch"])
def test_boundary_force_adapters(case):
    import oracle_py as O
    g = golden(case)
    dt = 0.1
    o = run("boundary", dict(X=g["X"], x=g["x"], tris=g["tris"]), [0.0, dt, 0.0, 1])
    assert relerr(o["areatot"], g["areatot"]) < 1e-13
    assert relerr(o["F_bend"], g["F_bend"]) < 1e-13
    # tot_area comes from the GPU's tree reduction here (1e-16 apart from the
    # reference's sequential sum); the force scales with tot_area - area_0,
    # which cancels A ~ 310 against area_0 ~ 314 (about 80x amplification)
    assert relerr(o["F_area"], g["F_area"]) < 1e-11
    # actin polymerisation: each vertex either idle or the fired value
    n = len(g["x"])
    fired, _ = O.force_ap(g["tris"], g["x"], dt, np.zeros(n))
    a, f = o["F_ap"].reshape(-1, 3), fired.reshape(-1, 3)
    on = np.abs(a).max(axis=1) > 0
    assert on.sum() > 0
    assert relerr(a[on], f[on]) < 1e-13


@pytest.mark.parametrize("case,law", [("forcesubs_k31", 0), ("forcesubs_kMD", 1)])
def test_force_fa_subs_adapter(case, law):
    g = golden(case)
    o = run("forcesubs", dict(X=g["SX"], x=g["Sx"], tris=g["Stris"], x_cell=g["x"], cC=g["cC"], fau=g["fau"],
                              cell_tris=g["tris"]), [0.0, 0.0, float(g["kint"]), law])
    assert relerr(o["F_subs"], g["F_subs"]) < 1e-12
