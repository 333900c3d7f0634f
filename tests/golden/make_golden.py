"""Generate golden vectors from the compiled reference (this container only).

Runs oracle/_ref/refharness_{k1,k31,kMD} (built by `make -C oracle ref` from the
reference's own UL_growth_explicit.cpp + nondestructivetrimesh.cpp) on seeded
inputs over the reference's shipped meshes, and stores inputs and outputs as
small compressed .npz fixtures next to this script.  The fixtures are data
(mesh coordinates/connectivity read from the reference's scripts/*.txt input
files, seeded inputs, and the reference's outputs); nothing of the reference's
source travels.

Usage:  python tests/golden/make_golden.py [--only boundary|substrate|vtk]
"""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
SCRIPTS = os.path.join(REF, "wholeCellFE_kMD", "cellBP", "scripts")
HARNESS = os.path.join(REPO, "oracle", "_ref", "refharness_{}")


def read_ecomsol(path):
    """Parse the reference 'ecomsol' mesh format (iomesh.cpp:491-565); z := 0."""
    lines = open(path).read().split("\n")
    i = next(k for k, l in enumerate(lines) if "Nodes" in l)
    n = int(lines[i + 1].split()[0])
    pts = np.zeros((n, 3))
    for k in range(n):
        v = lines[i + 2 + k].split()
        pts[k, 0], pts[k, 1] = float(v[0]), float(v[1])
    j = next(k for k, l in enumerate(lines) if "Elements" in l)
    m = int(lines[j + 1].split()[0])
    tris = np.array([[int(v) for v in lines[j + 2 + k].split()[:3]] for k in range(m)], dtype=np.int64)
    return pts, tris


def run(variant, mode, inputs, meta):
    d = tempfile.mkdtemp(prefix="wcfe_golden_")
    try:
        for k, v in inputs.items():
            np.ascontiguousarray(v).tofile(os.path.join(d, k + ".bin"))
        with open(os.path.join(d, "meta.txt"), "w") as f:
            for k, v in meta.items():
                f.write("%s %.17g\n" % (k, v))
        subprocess.run([HARNESS.format(variant), mode, d], check=True)
        out = {}
        for fn in os.listdir(d):
            if fn.startswith("out_"):
                out[fn[4:-4]] = np.fromfile(os.path.join(d, fn), dtype=np.float64)
            elif fn == "out.vtk":
                out["vtk"] = np.frombuffer(open(os.path.join(d, fn), "rb").read(), dtype=np.uint8)
        return out
    finally:
        shutil.rmtree(d)


def deform(X, rng, amp):
    A = np.array([[1.04, 0.03], [-0.02, 0.96]])
    x = X.copy()
    x[:, :2] = X[:, :2] @ A.T + amp * rng.standard_normal((len(X), 2))
    return x


def boundary_cases(disk296, t296, disk483, t483):
    """calculateBendingForce2D, calculateAreaConstraintForce, calculateForce_AP.

    calculateForce_AP seeds rand() from the clock, so its fixture records one
    realisation: every boundary vertex's force is either 0 or the fired value.
    dt = 0.1 puts the firing probability at rate dt exp(-rate dt) = 1/e.
    """
    rng = np.random.default_rng(20250917)
    for name, X, T in [("bound_disk296", disk296, t296), ("bound_disk483", disk483, t483)]:
        x = deform(X, rng, 0.02)
        inp = dict(X=X, x=x, tris=T)
        ob = run("kMD", "bend", inp, {})
        oa = run("kMD", "areacon", inp, {})
        op = run("kMD", "ap", inp, dict(dt=0.1))
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **inp, ap_dt=0.1, F_bend=ob["F"],
                            F_area=oa["F"], areatot=oa["areatot"], F_ap=op["F"])
        print("wrote", name)
    # notched disk: two passes that drop every 5th triangle with a boundary
    # edge, so that some boundary vertices see up to 6 boundary neighbours
    # (the reference keeps the last one it meets)
    X, T = disk296, t296
    for r in range(2):
        edges = {}
        for t in T:
            for j in range(3):
                k = tuple(sorted((int(t[j]), int(t[(j + 1) % 3]))))
                edges[k] = edges.get(k, 0) + 1
        bt = [e for e, t in enumerate(T)
              if any(edges[tuple(sorted((int(t[j]), int(t[(j + 1) % 3]))))] == 1 for j in range(3))]
        keep = np.ones(len(T), bool)
        keep[bt[r::5]] = False
        T = T[keep]
    x = deform(X, rng, 0.02)
    inp = dict(X=X, x=x, tris=T)
    ob = run("kMD", "bend", inp, {})
    oa = run("kMD", "areacon", inp, {})
    np.savez_compressed(os.path.join(HERE, "bound_disk296_notch.npz"), **inp, F_bend=ob["F"], F_area=oa["F"],
                        areatot=oa["areatot"])
    print("wrote bound_disk296_notch")


def substrate_cases(disk296, t296, subs, tsubs):
    """calculateForce_FA_subs (k31 linear, kMD table) and the substrate
    overload of calculateRes (mat = 2)."""
    rng = np.random.default_rng(20251004)
    N = len(disk296)
    for name, variant, kint in [("forcesubs_k31", "k31", 31000.0), ("forcesubs_kMD", "kMD", -1.0)]:
        x = deform(disk296, rng, 0.01)
        D = rng.uniform(0.0, 0.02, N)
        th = rng.uniform(0, 2 * np.pi, N)
        fau = np.zeros((N, 3))
        fau[:, 0], fau[:, 1] = D * np.cos(th), D * np.sin(th)
        # node 0's substrate point on a substrate vertex: every triangle around
        # that vertex is loaded (the reference does not stop at the first hit)
        v = int(np.argmin(np.sum((subs[:, :2] - x[0, :2]) ** 2, axis=1)))
        fau[0, :2] = x[0, :2] - subs[v, :2]
        cC = rng.uniform(0, 1, N)
        SX = subs.copy()
        Sx = SX.copy()
        Sx[:, 0] += 0.01 * np.sin(SX[:, 1] / 7.0)
        Sx[:, 1] += 0.01 * np.cos(SX[:, 0] / 5.0)
        fau[0, :2] = x[0, :2] - Sx[v, :2]
        inp = dict(X=disk296, x=x, cC=cC, fau=fau, SX=SX, Sx=Sx, tris=t296, Stris=tsubs)
        o = run(variant, "forcefasubs", inp, dict(kint=kint))
        np.savez_compressed(os.path.join(HERE, name + ".npz"), kint=kint, **inp, F_subs=o["F_subs"])
        print("wrote", name)
    Ns = len(subs)
    x = subs.copy()
    x[:, :2] += 0.002 * rng.standard_normal((Ns, 2))
    ea = np.zeros((Ns, 6))
    ea[:, [0, 1, 3]] = 1e-4 * rng.standard_normal((Ns, 3))
    inp = dict(X=subs, x=x, ea=ea, tris=tsubs)
    o = run("kMD", "calcressubs", inp, dict(dt=0.005, time=1.0))
    np.savez_compressed(os.path.join(HERE, "calcres_subs.npz"), dt=0.005, time=1.0, **inp, res_Res=o["Res"],
                        res_Res_ea=o["Res_ea"], res_MM=o["MM"])
    print("wrote calcres_subs")


def vtk_cases(disk296, t296, subs, tsubs):
    """The per-frame VTK writers (iomesh.cpp:844 cell frame, per variant;
    iomesh.cpp:730 substrate frame).  The fixture holds the inputs and the
    written text."""
    rng = np.random.default_rng(20251102)
    N = len(disk296)
    x = deform(disk296, rng, 0.02)
    ea = np.zeros((N, 6))
    ea[:, [0, 1, 3]] = 0.05 * rng.standard_normal((N, 3))
    cA = rng.uniform(0, 100, N)
    cC = rng.uniform(0, 1, N)
    D = rng.uniform(0.0, 0.022, N)  # every MD-table segment and beyond
    th = rng.uniform(0, 2 * np.pi, N)
    fau = np.zeros((N, 3))
    fau[:, 0], fau[:, 1] = D * np.cos(th), D * np.sin(th)
    fau[0] = 0.0
    inp = dict(X=disk296, x=x, tris=t296, ea=ea, cA=cA, cC=cC, fau=fau)
    for variant in ("k1", "k31", "kMD"):
        o = run(variant, "vtkcell", inp, dict(time=1.2345678901, kint=-1.0))
        np.savez_compressed(os.path.join(HERE, "vtk_cell_%s.npz" % variant), time=1.2345678901, **inp, vtk=o["vtk"])
        print("wrote vtk_cell_%s" % variant)
    Ns = len(subs)
    xs = subs.copy()
    xs[:, :2] += 0.002 * rng.standard_normal((Ns, 2))
    eas = np.zeros((Ns, 6))
    eas[:, [0, 1, 3]] = 1e-4 * rng.standard_normal((Ns, 3))
    faus = np.zeros((Ns, 3))
    faus[:, :2] = 1e-3 * rng.standard_normal((Ns, 2))
    inp = dict(X=subs, x=xs, tris=tsubs, ea=eas, fau=faus)
    o = run("kMD", "vtksubs", inp, dict(time=0.04))
    np.savez_compressed(os.path.join(HERE, "vtk_subs.npz"), time=0.04, **inp, vtk=o["vtk"])
    print("wrote vtk_subs")


def main():
    if not os.path.exists(HARNESS.format("kMD")):
        sys.exit("build the reference harness first: make -C oracle ref")
    rng = np.random.default_rng(20240601)
    disk296, t296 = read_ecomsol(os.path.join(SCRIPTS, "disk_20um_diam_edited.txt"))
    disk483, t483 = read_ecomsol(os.path.join(SCRIPTS, "disk_20um_diam_N1025.txt"))
    subs, tsubs = read_ecomsol(os.path.join(SCRIPTS, "sq_subs_100um.txt"))
    if "--only" in sys.argv and sys.argv[sys.argv.index("--only") + 1] == "boundary":
        boundary_cases(disk296, t296, disk483, t483)
        return
    if "--only" in sys.argv and sys.argv[sys.argv.index("--only") + 1] == "substrate":
        substrate_cases(disk296, t296, subs, tsubs)
        return
    if "--only" in sys.argv and sys.argv[sys.argv.index("--only") + 1] == "vtk":
        vtk_cases(disk296, t296, subs, tsubs)
        return

    # --- element residual / calculateRes (law identical in all variants) ---
    for name, X, T, time, ea3d in [("ul_disk296_t1p3", disk296, t296, 1.3, False),
                                   ("ul_disk483_t3", disk483, t483, 3.0, False),
                                   ("ul_disk296_ea3d", disk296, t296, 0.7, True)]:
        N = len(X)
        x = deform(X, rng, 0.02)
        ea = np.zeros((N, 6))
        ea[:, [0, 1, 3]] = 0.05 * rng.standard_normal((N, 3))
        if ea3d:
            ea[:, [2, 4, 5]] = 0.02 * rng.standard_normal((N, 3))
        cA = rng.uniform(0, 100, N)
        cC = rng.uniform(0, 1, N)
        inp = dict(X=X, x=x, ea=ea, cA=cA, cC=cC, tris=T)
        meta = dict(dt=0.005, time=time)
        o1 = run("kMD", "elem", inp, meta)
        o2 = run("kMD", "calcres", inp, meta)
        o3 = run("kMD", "areatot", inp, meta)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), dt=0.005, time=time, **inp,
                            **{"elem_" + k: v for k, v in o1.items()},
                            **{"res_" + k: v for k, v in o2.items()},
                            areatot=o3["areatot"])
        print("wrote", name)

    # --- calculateForce_FA (linear k1/k31, MD table kMD) ---
    N = len(disk296)
    for name, variant, kint, planar in [("forcefa_k1", "k1", 1000.0, True), ("forcefa_k31", "k31", 31000.0, True),
                                        ("forcefa_kMD", "kMD", -1.0, True), ("forcefa_kMD_3d", "kMD", -1.0, False)]:
        x = deform(disk296, rng, 0.02)
        D = rng.uniform(0.0, 0.022, N)
        D[:4] = [0.0, 2.51250583e-03, 17.001e-03, 4.40265201e-03]  # exact knots and D == 0
        th = rng.uniform(0, 2 * np.pi, N)
        fau = np.zeros((N, 3))
        if planar:
            fau[:, 0], fau[:, 1] = D * np.cos(th), D * np.sin(th)
        else:
            ph = rng.uniform(0.2, np.pi - 0.2, N)
            fau[:, 0], fau[:, 1], fau[:, 2] = D * np.cos(th) * np.sin(ph), D * np.sin(th) * np.sin(ph), D * np.cos(ph)
        cC = rng.uniform(0, 1, N)
        inp = dict(X=disk296, x=x, cC=cC, fau=fau, tris=t296)
        o = run(variant, "forcefa", inp, dict(kint=kint))
        np.savez_compressed(os.path.join(HERE, name + ".npz"), kint=kint, **inp, F_FA=o["F_FA"])
        print("wrote", name)

    # --- updateFA (11-argument form) with the 100 um substrate ---
    for name, variant, kint, moving in [("updatefa_k1_rigid", "k1", 1000.0, False),
                                        ("updatefa_k31_moving", "k31", 31000.0, True),
                                        ("updatefa_kMD_rigid", "kMD", -1.0, False),
                                        ("updatefa_kMD_moving", "kMD", -1.0, True)]:
        N = len(disk296)
        X = disk296.copy()
        x = X.copy()
        x[:, :2] += 0.004 * rng.standard_normal((N, 2))
        D = rng.uniform(0.0, 0.02, N)
        th = rng.uniform(0, 2 * np.pi, N)
        fau = np.zeros((N, 3))
        fau[:, 0], fau[:, 1] = D * np.cos(th), D * np.sin(th)
        cA = np.full(N, 100.0)
        cC = rng.uniform(0, 1, N)
        cC[:3] = [0.0, 1.0, 1e-9]
        SX = subs.copy()
        Sx = SX.copy()
        if moving:
            Sx[:, 0] += 0.01 * np.sin(SX[:, 1] / 7.0)
            Sx[:, 1] += 0.01 * np.cos(SX[:, 0] / 5.0)
        inp = dict(X=X, x=x, cA=cA, cC=cC, fau=fau, SX=SX, Sx=Sx, tris=t296, Stris=tsubs)
        o = run(variant, "updatefa", inp, dict(dt=0.005, kint=kint))
        np.savez_compressed(os.path.join(HERE, name + ".npz"), dt=0.005, kint=kint, **inp,
                            out_cA=o["cA"], out_cC=o["cC"], out_fau=o["fau"])
        print("wrote", name)

    # config-1 template mesh (483 nodes / 905 triangles) for the bench and tests
    np.savez_compressed(os.path.join(HERE, "mesh_disk483.npz"), X=disk483[:, :2], tris=t483.astype(np.int32))
    print("wrote mesh_disk483")
    boundary_cases(disk296, t296, disk483, t483)
    substrate_cases(disk296, t296, subs, tsubs)
    vtk_cases(disk296, t296, subs, tsubs)


if __name__ == "__main__":
    main()
