#!/usr/bin/env python3
"""Benchmark: Newton steps/s of the batched mixed-precision GPU form vs the
plain FP64 GPU form, at <= 1e-8 relative linear residual (FP64 refinement).

Workload (default): 256 independent disk cells in one batch, triangle counts
log-uniform in [200, 5000], E uniform in [0.5, 5] kPa, nu = 0.3, thickness
0.2 um, spring law per cell drawn from {1 pN/nm, 31 pN/nm, 64-knot stiffening
table}, T/3 springs per cell; dense FP32 blocked LU (MFMA) + FP64 iterative
refinement; one hipGraph replay per batched Newton step.

A step is one Newton iteration on every cell of the batch (assemble, factor,
solve with refinement to 1e-8, update, FP64 residual and convergence test,
commit of converged load steps).  value = Newton steps/s summed over cells.

    python bench.py [--gpus 1] [--steps 10] [--warmup 3]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "Newton steps/s (batched mixed-prec vs plain FP64 GPU) at 1e-8 rel residual"
WORKLOAD_TEXT = {
    "batched256": "batched %(n)d disk cells, T log-uniform 200-5000 (total %(t)d triangles), E 0.5-5 kPa, k in {1, "
                  "31 pN/nm, 64-knot table}, T/3 springs/cell, FP32 LU (MFMA, %(piv)s pivoting; band LU in "
                  "reverse Cuthill-McKee numbering) + FP64 refinement to 1e-8, hipGraph",
    "single20k": "one disk cell, %(t)d triangles, 5000 springs on the 64-knot stiffening table, dense FP32 LU "
                 "(MFMA, %(piv)s pivoting) + FP64 refinement to 1e-8",
    "single500k": "one disk cell, %(t)d triangles, 100000 springs on the 64-knot stiffening table, block-CSR FP32 "
                  "block-Jacobi PCG + FP64 refinement to 1e-8",
    "replicas1024": "%(n)d replicas, T uniform 1000-10000 (total %(t)d triangles), Bell kinetics, solver per cell "
                    "by size (dense FP32 LU / block-CSR PCG) + FP64 refinement to 1e-8",
}
MFMA_FP32_PEAK = 157.3e12
HBM_PEAK = 8.0e12


def cpu_baseline(cells, budget_s=15.0):
    """The FP64 oracle (oracle/liboracle.so, one thread) timed on a sample of
    the batch's cells: one Newton step each (assemble K and G, partial-pivot
    LU, solve, update).  The batch rate is extrapolated with the measured
    per-cell cost scaled by (n / n_sample)^3 (dense LU dominates)."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_py as O
    from wholecellfe_amd import mesh

    order = sorted(range(len(cells)), key=lambda i: len(cells[i]["X"]))
    sample, spent, rec = [], 0.0, []
    px, py, kk = mesh.stiffening_table()
    # every cell with n <= 3200, smallest first, until the budget is spent
    for idx in order:
        c = cells[idx]
        n = 2 * len(c["X"])
        if n > 3200 or spent > budget_s:
            continue
        d = dict(X=c["X"].reshape(-1), x=c["X"].reshape(-1).copy(), tris=c["tris"].reshape(-1), law=1, E=c["E"],
                 nu=c["nu"], thick=c["thick"], time=0.5, spr_tri=c["spr_tri"], spr_bary=c["spr_bary"].reshape(-1),
                 spr_anchor=c["spr_anchor"].reshape(-1), spr_state=np.ones(len(c["spr_tri"])),
                 spring_law=c["spring_law"], k_lin=c["k_lin"], tab_px=px, tab_py=py, tab_k=kk)
        oc = O.make_cell(d)
        t0 = time.perf_counter()
        G, K = O.assemble(oc)
        LU, piv, _ = O.lu_factor(K, pivoting=1)
        dx = O.lu_solve(LU, piv, -G, pivoting=1)
        _ = d["x"] + dx
        dt = time.perf_counter() - t0
        spent += dt
        rec.append((n, dt))
        sample.append(idx)
    rec.sort()
    big = rec[len(rec) // 2:]
    coef = float(np.median([t / n ** 3 for n, t in big]))
    total = sum(coef * (2 * len(c["X"])) ** 3 for c in cells)
    return {"value": len(cells) / total, "unit": "Newton steps/s", "cores": 1, "kind": "port",
            "sample": "FP64 oracle Newton step (assemble, partial-pivot LU, solve) on %d of %d cells with n <= %d "
                      "(%.1f s measured); batch time extrapolated as sum of c n^3, c = %.3e s" %
                      (len(rec), len(cells), max(n for n, _ in rec), spent, coef)}


def measured_traffic(kernel, n_cells, pivoting):
    """HBM bytes per launch of the LU kernel, from the rocprofv3 FETCH_SIZE and
    WRITE_SIZE passes summarised in profiles/ (scaled by the same counters'
    reading for a plain copy of known size); None when no summary exists for
    this workload."""
    path = os.path.join(REPO, "profiles", "%s_batched%d_%s_traffic.txt" % (kernel, n_cells, pivoting))
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def gemm_roofline(prof, a, depth, diag_per_step=0, fused=False):
    """The LU's dominant kernel, latency bound under the band LU (each launch
    is one block step of the chain of dependent diagonal blocks).
    fused (one launch per block step, k_lu_step): achieved = useful 64^3 block
    products (trailing updates, L block solves, and the diagonal inverse, whose
    Gauss-Jordan elimination is 64^3 multiply-adds) x 2 64^3 flop / summed
    launch time (HIP events); the inverses that L items recompute are not
    counted.  Otherwise (k_lu_gemm): trailing-update products x 2 64^3 +
    diagonal blocks x (4/3) 64^3 flop."""
    g = prof["lu_gemm"]
    tile_flops = 2.0 * 64 ** 3
    diag_flops = 0.0 if fused else (4.0 / 3.0) * 64 ** 3
    flops = g["work"] * tile_flops + diag_per_step * diag_flops
    achieved = flops / (g["ms"] * 1e-3) if g["ms"] > 0 else 0.0
    launches = max(g["launches"], 1)
    kernel = "k_lu_step" if fused else "k_lu_gemm"
    traffic = measured_traffic(kernel, a.cells, a.pivoting) if a.workload == "batched256" else None
    if fused:
        name = ("k_lu_step<float> (fused band-LU block step: trailing update, Gauss-Jordan inverse of the "
                "diagonal block, L block solves; v_mfma_f32_16x16x4f32)")
    else:
        name = ("k_lu_gemm<float,false> (LU trailing update + fused diagonal factorisation, "
                "v_mfma_f32_16x16x4f32)")
    out = {"kernel": name, "bound": "mfma",
           "achieved": achieved / 1e12, "peak": MFMA_FP32_PEAK / 1e12, "unit": "TFLOP/s",
           "frac": achieved / MFMA_FP32_PEAK, "traffic": traffic,
           "launches_per_step": g["launches"], "block_products_per_step": g["work"],
           "flops_per_block_product": tile_flops, "avg_launch_us": g["ms"] * 1e3 / launches}
    if not fused:
        out.update({"c_tile_bytes_per_launch": g["work"] / launches / depth * 2 * 64 * 64 * 4,
                    "panels_per_update": depth, "diag_blocks_per_step": diag_per_step,
                    "flops_per_diag_block": diag_flops})
    return out


def pcg_roofline(cells, prof, st):
    """Block-CSR PCG (FP32): HBM-bound; algorithmic bytes per iteration
    20 nnzb (2 x 2 FP32 values + column index per block) + 116 N (row
    pointers, the five PCG vectors, the 2 x 2 diagonal inverses) over the PCG
    iterations of the solve phase."""
    nnzb, nodes = 0, 0
    for c in cells:
        if c.get("solver", 0) != 2:
            continue
        T = np.asarray(c["tris"]).reshape(-1, 3).astype(np.int64)
        n = len(c["X"])
        pairs = np.concatenate([T[:, i] * n + T[:, j] for i in range(3) for j in range(3)])
        nnzb += len(np.unique(pairs))
        nodes += n
    iters = max(st.get("max_cg_iters", 0), 1) * max(st.get("max_refine_used", 1), 1)
    bytes_iter = 20.0 * nnzb + 116.0 * nodes
    ms = prof["solve"]["ms"]
    achieved = bytes_iter * iters / (ms * 1e-3) if ms > 0 else 0.0
    return {"kernel": "block-CSR PCG (k_pcg_spmv_dot + k_pcg_update + k_pcg_paxpy)", "bound": "hbm",
            "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK,
            "traffic": None, "nnz_blocks": nnzb, "nodes": nodes, "pcg_iterations": iters,
            "bytes_per_iteration": bytes_iter}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cells", type=int, default=256, help="cells of the batched256 workload")
    ap.add_argument("--workload", default="batched256",
                    help="batched256 (default: the headline batch), single20k, single500k or replicas1024")
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--pivoting", choices=["static", "partial"], default="static")
    ap.add_argument("--autotune", choices=["none", "solver", "lu", "all"], default=None,
                    help="run wcfe_autotune on the batch before timing (default: solver for replicas1024, "
                         "none otherwise)")
    ap.add_argument("--fused-step", type=int, choices=[0, 1, 2, 3], default=None,
                    help="band-LU block-step mode (library default when omitted)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only time a plain device copy of MIB MiB (HBM probe / counter reference) and exit")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported (the cells of a batch are "
              "independent systems: several GPUs would each take a share of them, with no exchange)" % a.gpus,
              file=sys.stderr)
        sys.exit(2)

    import wholecellfe_amd as W
    from wholecellfe_amd import configs

    if a.copy_probe:
        with W.Context() as c:
            nbytes = a.copy_probe << 20
            ms = c.copy_probe(nbytes, reps=20)
        print(json.dumps({"copy_bytes": nbytes, "ms_per_copy": ms, "hbm_GBps": 2 * nbytes / (ms * 1e-3) / 1e9}))
        return

    log = lambda *m: print("[bench]", *m, file=sys.stderr, flush=True)
    t_setup = time.perf_counter()
    if a.workload == "batched256":
        cells, tables = configs.batched_cells(n_cells=a.cells)
    else:
        cells, tables = configs.workload(a.workload)
        a.no_cpu = True  # the CPU baseline samples the dense batch only
    n_tri = sum(len(c["tris"]) for c in cells)
    ndofs = [2 * len(c["X"]) for c in cells]
    log("built %d cells: %d triangles, dof %d..%d in %.1fs" % (len(cells), n_tri, min(ndofs), max(ndofs),
                                                               time.perf_counter() - t_setup))
    piv = W.PIVOT_STATIC if a.pivoting == "static" else W.PIVOT_PARTIAL
    common = dict(lin_rtol=1e-8, newton_rtol=1e-7, newton_atol=1e-12, t0=0.5, dt=0.005, pivoting=piv)
    if a.fused_step is not None:
        common["fused_step"] = a.fused_step

    # ---- batched mixed precision, one hipGraph replay per Newton step ----
    ctx = W.Context(precision=W.PREC_MIXED, batched=1, use_graph=1, max_refine=8, **common)
    configs.load(ctx, cells, tables)
    tune = a.autotune or ("solver" if a.workload == "replicas1024" else "none")
    tuned = None
    if tune != "none":
        flags = {"solver": W.TUNE_SOLVER, "lu": W.TUNE_LU, "all": W.TUNE_SOLVER | W.TUNE_LU}[tune]
        t_tune = time.perf_counter()
        tuned = ctx.autotune(flags=flags, steps=2)
        log("autotune (%s): %s in %.1fs" % (tune, tuned, time.perf_counter() - t_tune))
    log("mixed context ready")
    for _ in range(a.warmup):
        ctx.newton_step(stats=False)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        ctx.newton_step(stats=False)
    ctx.synchronize()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / a.steps
    value = len(cells) * a.steps / (t1 - t0)
    log("mixed: %.2f ms/step, %.0f Newton steps/s" % (ms_step, value))
    st = ctx.newton_step(stats=True)
    prof = ctx.profile_step()
    sched = ctx.schedule()
    ctx.close()
    log("check step:", st)
    log("phases:", {k: round(v["ms"], 3) for k, v in prof.items()})
    log("LU schedule:", sched)
    dense = [c for c in cells if c.get("solver", 0) != 2]
    # diagonal blocks factored per step inside the trailing-update launches
    # (block 0 of each cell has its own kernel)
    diag_per_step = sum(max((2 * len(c["X"]) + 63) // 64 - 1, 0) for c in dense)
    roofline = (gemm_roofline(prof, a, sched["lu_depth"], 0 if sched["fused_step"] else diag_per_step,
                              fused=bool(sched["fused_step"]))
                if prof["lu_gemm"]["launches"] > 0 else pcg_roofline(cells, prof, st))
    lu_alg = configs.lu_flops(dense, symmetric=a.pivoting == "static")
    lu_ms = prof["lu_gemm"]["ms"] + prof["lu_strip"]["ms"] + prof["lu_panel"]["ms"]

    # ---- plain FP64 GPU form: same kernels in FP64, one cell at a time, no graph ----
    fp64 = None
    if not a.no_fp64:
        c64 = W.Context(precision=W.PREC_FP64, batched=0, use_graph=0, max_refine=3, **common)
        configs.load(c64, cells, tables)
        c64.newton_step(stats=False)
        c64.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.fp64_steps):
            c64.newton_step(stats=False)
        c64.synchronize()
        t1 = time.perf_counter()
        st64 = c64.newton_step(stats=True)
        c64.close()
        v64 = len(cells) * a.fp64_steps / (t1 - t0)
        fp64 = {"value": v64, "unit": "Newton steps/s", "ms_per_step": (t1 - t0) * 1e3 / a.fp64_steps,
                "steps": a.fp64_steps, "lin_relres_max": st64["max_lin_relres"]}
        log("fp64 plain: %.1f ms/step, %.0f Newton steps/s" % (fp64["ms_per_step"], v64))

    cpu = None if a.no_cpu else cpu_baseline(cells)
    if cpu:
        log("cpu baseline:", cpu)

    out = {
        "metric": METRIC, "value": value, "unit": "Newton steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32+fp64", "data": "synthetic",
        "config": {"workload": WORKLOAD_TEXT.get(a.workload, a.workload) % dict(n=len(cells), t=n_tri, piv=a.pivoting),
                   "cells": len(cells), "triangles": n_tri, "dof_min": min(ndofs), "dof_max": max(ndofs),
                   "parallelism": "single GPU", "lu_schedule": sched},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "fp64_plain": fp64,
        "speedup_vs_fp64_plain": (value / fp64["value"]) if fp64 else None,
        "element_assemblies_per_s": n_tri * a.steps / ((ms_step * a.steps) * 1e-3),
        "lin_relres_max": st["max_lin_relres"],
        "refine_sweeps_max": st["max_refine_used"],
        "static_pivots": st["n_static_pivots"],
        # the flop rate a dense (full-matrix) LU would need to factor the batch in
        # the time the band LU takes, against the FP32 MFMA peak
        "lu_dense_equivalent_tflops": (lu_alg / (lu_ms * 1e-3)) / 1e12 if lu_ms > 0 else None,
        "lu_dense_equivalent_over_mfma_peak": (lu_alg / (lu_ms * 1e-3)) / MFMA_FP32_PEAK if lu_ms > 0 else None,
        "phases_ms": {k: round(v["ms"], 4) for k, v in prof.items()},
        "autotuned": tuned,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
