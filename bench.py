#!/usr/bin/env python3
"""Benchmark: Newton steps/s of the batched mixed-precision GPU form vs the
plain FP64 GPU form, at <= 1e-8 relative linear residual (FP64 refinement)
and a 1e-8 relative Newton tolerance.

Headline workload (default): 256 independent disk cells in one batch,
triangle counts log-uniform in [200, 5000], E uniform in [0.5, 5] kPa,
nu = 0.3, thickness 0.2 um, spring law per cell drawn from {1 pN/nm,
31 pN/nm, 64-knot stiffening table}, T/3 springs per cell; FP32 band LU on
MFMA + FP64 iterative refinement; one hipGraph replay per batched Newton
step.  The other workload families (one 864-triangle cell in FP64, one
20k-triangle cell, one 500k-triangle cell on block-CSR PCG, 1,024 Bell
replicas) are timed after it, each under "workloads" with its own roofline.

A step is one Newton iteration on every cell of the batch (assemble, factor,
solve with refinement to 1e-8, update, FP64 residual and convergence test,
commit of converged load steps).  value = Newton steps/s summed over cells.

    python bench.py [--gpus 1] [--steps 10] [--warmup 3] [--no-extra]
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
                  "31 pN/nm, 64-knot table}, T/3 springs/cell, FP32 LU (MFMA, %(piv)s pivoting; band LU in reverse "
                  "Cuthill-McKee numbering) + FP64 refinement to 1e-8, hipGraph",
    "single905": "one disk cell, %(t)d triangles, 300 linear integrins (1 pN/nm) anchored on a rigid 50 x 50 grid, "
                 "plain FP64 (band LU on FP64 MFMA, FP64 residual), hipGraph",
    "single20k": "one disk cell, %(t)d triangles, 5000 springs on the 64-knot stiffening table, dense FP32 LU "
                 "(MFMA, %(piv)s pivoting; band LU) + FP64 refinement to 1e-8",
    "single500k": "one disk cell, %(t)d triangles, 100000 springs on the 64-knot stiffening table, block-CSR FP32 "
                  "block-Jacobi PCG + FP64 refinement to 1e-8",
    "replicas1024": "%(n)d replicas, T uniform 1000-10000 (total %(t)d triangles), Bell kinetics, FP32 band LU "
                    "for every cell + FP64 refinement to 1e-8, hipGraph",
}
MFMA_FP32_PEAK = 157.3e12
MFMA_FP64_PEAK = 78.6e12
HBM_PEAK = 8.0e12
T0, DT = 0.5, 0.005


def log(*m):
    print("[bench]", *m, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# CPU leg: the FP64 oracle as the baseline and as the checker
# ---------------------------------------------------------------------------
def _oracle():
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_py as O
    return O


def oracle_dict(c, tables, x=None, time=T0):
    px, py, kk = tables[c.get("table_id", 0)]
    X = np.asarray(c["X"], dtype=np.float64).reshape(-1)
    return dict(X=X, x=X.copy() if x is None else np.asarray(x, dtype=np.float64).reshape(-1),
                tris=np.asarray(c["tris"]).reshape(-1), law=c["law"], E=c["E"], nu=c["nu"], thick=c["thick"],
                time=time, spr_tri=c["spr_tri"], spr_bary=np.asarray(c["spr_bary"]).reshape(-1),
                spr_anchor=np.asarray(c["spr_anchor"]).reshape(-1), spr_state=np.ones(len(c["spr_tri"])),
                spring_law=c["spring_law"], k_lin=c["k_lin"], tab_px=px, tab_py=py, tab_k=kk)


def cpu_baseline(cells, tables, budget_s=20.0):
    """The FP64 oracle (oracle/liboracle.so) running one Newton iteration on
    every cell of the batch -- assemble K and G, the same reverse Cuthill-McKee
    band LU the GPU uses (FP64, no pivoting), solve, update -- with OpenMP over
    the cells, on 1 thread and on the host's cores (at most 16); numbering and
    band storage are prepared before the timed part.  No extrapolation: every
    cell is stepped."""
    O = _oracle()
    order = sorted(range(len(cells)), key=lambda i: -len(cells[i]["tris"]))  # largest first
    ds = [oracle_dict(cells[i], tables) for i in order]
    ocs = [O.make_cell(d) for d in ds]

    def run(threads):
        xs = [d["x"].copy() for d in ds]
        t, err = O.batch_band_newton_step(ocs, xs, n_threads=threads)
        if err:
            raise RuntimeError("oracle band LU failed on %d cells" % err)
        return t

    threads = max(1, min(16, os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16") or 16)))
    t1 = [run(1)]
    while sum(t1) < budget_s / 2 and len(t1) < 5:
        t1.append(run(1))
    tn = [run(threads)]
    while sum(tn) < budget_s / 4 and len(tn) < 9:
        tn.append(run(threads))
    s1, sn = float(np.median(t1)), float(np.median(tn))
    return {"value": len(cells) / sn, "unit": "Newton steps/s", "cores": threads, "kind": "port",
            "sample": "FP64 oracle, one Newton iteration on all %d cells (assemble, RCM band LU, solve, update), "
                      "median of %d runs on %d threads; 1 thread: %.3f s per batch step (median of %d)"
                      % (len(cells), len(tn), threads, s1, len(t1)),
            "single_thread": {"value": len(cells) / s1, "unit": "Newton steps/s", "cores": 1},
            "seconds_per_batch_step": sn}


def oracle_check(cells, tables, common, n_sample=16):
    """Error of the GPU's converged positions against the FP64 oracle on a
    size-spread sample of the workload's cells (one load step at t0 in a
    batch of those cells, the headline's options): x_err = max |x - x_o| /
    max |x_o - X| with x_o the oracle's FP64 Newton solution (Jacobi-PCG
    linear solves), force_err = ||G(x_gpu)|| / ||G(X)|| (the oracle's FP64
    residual at the GPU's positions)."""
    import wholecellfe_amd as W
    from wholecellfe_amd import configs
    O = _oracle()
    order = sorted(range(len(cells)), key=lambda i: len(cells[i]["tris"]))
    pick = sorted({order[int(round(k * (len(order) - 1) / (n_sample - 1)))] for k in range(n_sample)})
    sub = [cells[i] for i in pick]
    opts = dict(common, newton_rtol=1e-10, newton_atol=0.0)
    ctx = W.Context(precision=W.PREC_MIXED, batched=1, use_graph=1, max_refine=4, **opts)
    configs.load(ctx, sub, tables)
    # each cell's positions right after the Newton iteration that committed
    # its first load step (cells that converge first go on to the next one)
    xs = [None] * len(sub)
    for _ in range(30):
        ctx.newton_step()
        for i in range(len(sub)):
            if xs[i] is None and ctx.cell_info(i)["step"] >= 1:
                xs[i] = ctx.positions(i)
        if all(x is not None for x in xs):
            break
    ctx.close()
    if any(x is None for x in xs):
        raise RuntimeError("oracle check: a cell did not converge in 30 Newton iterations")
    ts = [T0] * len(sub)
    xe, fe = 0.0, 0.0
    for c, x, t in zip(sub, xs, ts):
        X = np.asarray(c["X"]).reshape(-1)
        G, _ = O.assemble(O.make_cell(oracle_dict(c, tables, x=x, time=t)), want_K=False)
        G0, _ = O.assemble(O.make_cell(oracle_dict(c, tables, time=t)), want_K=False)
        fe = max(fe, float(np.linalg.norm(G) / np.linalg.norm(G0)))
        x_o, it, _ = O.newton_solve_pcg(O.make_cell(oracle_dict(c, tables, time=t)), X, rtol=1e-11, atol=0.0, maxit=30)
        if it < 0:
            raise RuntimeError("oracle Newton did not converge")
        xe = max(xe, float(np.max(np.abs(x - x_o)) / np.max(np.abs(x_o - X))))
    sizes = [len(c["tris"]) for c in sub]
    return {"x_err_vs_oracle": xe, "force_err_vs_oracle": fe,
            "oracle_sample": "%d cells, %d..%d triangles, one load step each" % (len(sub), min(sizes), max(sizes))}


# ---------------------------------------------------------------------------
# rooflines
# ---------------------------------------------------------------------------
def measured_traffic(kernel, workload, pivoting):
    """HBM bytes per launch of a kernel from the rocprofv3 FETCH_SIZE and
    WRITE_SIZE passes summarised in profiles/ (scaled by the same counters'
    reading for a plain copy of known size); None without a summary."""
    path = os.path.join(REPO, "profiles", "%s_%s_%s_traffic.txt" % (kernel, workload, pivoting))
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def lu_roofline(prof, fused, fp64=False, traffic=None):
    """The band LU's dominant kernel.  fused (one launch per block step,
    k_lu_step): achieved = useful 64^3 block products (trailing updates, L
    block solves and the diagonal inverse, whose Gauss-Jordan elimination is
    64^3 multiply-adds) x 2 64^3 flop / summed launch time (HIP events at the
    kernel's own start and end); the inverses that L items form again are not
    counted.  Otherwise (k_lu_gemm): trailing-update products x 2 64^3 flop."""
    g = prof["lu_gemm"]
    tile_flops = 2.0 * 64 ** 3
    flops = g["work"] * tile_flops
    achieved = flops / (g["ms"] * 1e-3) if g["ms"] > 0 else 0.0
    peak = MFMA_FP64_PEAK if fp64 else MFMA_FP32_PEAK
    kernel = ("k_lu_step<%s> (fused band-LU block step: trailing update, Gauss-Jordan inverse of the diagonal block, "
              "L block solves)" if fused else "k_lu_gemm<%s> (LU trailing update)") % ("double" if fp64 else "float")
    return {"kernel": kernel, "bound": "mfma", "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
            "frac": achieved / peak, "traffic": traffic, "launches_per_step": g["launches"],
            "block_products_per_step": g["work"], "flops_per_block_product": tile_flops,
            "avg_launch_us": g["ms"] * 1e3 / max(g["launches"], 1)}


def _nnz_blocks(c):
    T = np.asarray(c["tris"]).reshape(-1, 3).astype(np.int64)
    n = len(c["X"])
    pairs = np.concatenate([T[:, i] * n + T[:, j] for i in range(3) for j in range(3)])
    return len(np.unique(pairs))


def pcg_roofline(cells, prof, st):
    """Block-CSR PCG (FP32, single-reduction form): HBM-bound; algorithmic
    bytes per iteration 20 nnzb (2 x 2 FP32 values + column index per block)
    + 124 N (row pointers 4; update: p, s, w, x, r read and p, s, x, r, u
    written, 8 each, the 2 x 2 diagonal inverse 16; SpMV: u, r read and w
    written, 8 each) over the PCG iterations of the solve phase."""
    nnzb = sum(_nnz_blocks(c) for c in cells if c.get("solver", 0) == 2)
    nodes = sum(len(c["X"]) for c in cells if c.get("solver", 0) == 2)
    iters = max(st.get("max_cg_iters", 0), 1) * max(st.get("max_refine_used", 1), 1)
    bytes_iter = 20.0 * nnzb + 124.0 * nodes
    ms = prof["solve"]["ms"]
    achieved = bytes_iter * iters / (ms * 1e-3) if ms > 0 else 0.0
    return {"kernel": "block-CSR PCG (k_pcg_update + k_pcg_spmv_dot with the per-cell scalars)", "bound": "hbm",
            "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK,
            "traffic": None, "nnz_blocks": nnzb, "nodes": nodes, "pcg_iterations": iters,
            "bytes_per_iteration": bytes_iter}


def phase_rooflines(cells, layouts, prof, st, s_sys=4):
    """HBM fractions of the other phases from their algorithmic bytes:
    assembly  12 T (triangles) + 8 d N (node state, d = 4 doubles: X, x) +
              n_spr (4 + 8 (3 + 2 + 1)) (spring triangle, weights, anchor,
              state) + s (4 nnzb + 2 N) (system entries and right-hand side
              written; s = 4 for FP32 K) + 8 (2 N) (FP64 G);
    refinement (per sweep) 8 21 T + 12 T + 8 3 (2 N) (symmetric FP64 K_e,
              triangles, d/b/r vectors), times the sweeps of the step;
    solves    (2 bw + 1) nblk 64^2 s per sweep of a dense cell (L and U band
              blocks and the diagonal inverse), times the sweeps, less the
              first forward sweep (it runs inside the fused block steps)."""
    T = sum(len(c["tris"]) for c in cells)
    N = sum(len(c["X"]) for c in cells)
    S = sum(len(c["spr_tri"]) for c in cells)
    nnzb = sum(_nnz_blocks(c) for c in cells)
    sweeps = max(st.get("max_refine_used", 1), 1)
    b_asm = 12.0 * T + 8 * 4.0 * N + S * (4.0 + 8 * 6) + s_sys * (4.0 * nnzb + 2.0 * N) + 8 * 2.0 * N
    b_ir = (8 * 21.0 * T + 12.0 * T + 8 * 3 * 2.0 * N) * sweeps
    b_sol = 0.0
    for L in layouts:
        if L["solver"] == 2:
            continue
        blk = 64 * 64 * s_sys * L["nblk"]
        b_sol += (2 * L["bw"] + 1) * blk * sweeps - L["bw"] * blk
    out = {}
    for name, b, ph in (("assemble", b_asm, "assemble"), ("refine", b_ir, "refine"), ("solve", b_sol, "solve")):
        ms = prof[ph]["ms"]
        gbs = b / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        out[name] = {"ms": ms, "alg_bytes": b, "GBps": gbs, "hbm_frac": gbs * 1e9 / HBM_PEAK,
                     "launches": prof[ph]["launches"]}
    return out


# ---------------------------------------------------------------------------
# GPU runs
# ---------------------------------------------------------------------------
def timed_steps(ctx, steps, warmup):
    for _ in range(warmup):
        ctx.newton_step(stats=False)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.newton_step(stats=False)
    ctx.synchronize()
    return time.perf_counter() - t0


REPLICA_ATOL = 1e-8


def replica_run(W, configs, cells, tables, common, load_steps, time_limit=120.0, **opt):
    """The Bell-replica family as a run: load_steps calls of wcfe_load_step
    (Newton iterations until every replica has committed one more load step;
    replicas that converge first go on), one graph replay per iteration.
    Reports replica load steps per second and the bound fraction after.

    The absolute Newton tolerance is REPLICA_ATOL: once the contraction has
    saturated (t >= 2 s) a replica whose bonds did not change in a step starts
    that load step in equilibrium, its residual is FP64 rounding noise
    (~1e-11 here) and no relative criterion can be met; 1e-8 sits above that
    floor and far below the forces (springs 1000 x displacement).  A load step
    that still misses max_newton is counted, not raised, and ends the run."""
    o = dict(common, **opt)
    o["newton_atol"] = max(o.get("newton_atol", 0.0), REPLICA_ATOL)
    ctx = W.Context(precision=W.PREC_MIXED, batched=1, use_graph=1, max_refine=4, **o)
    configs.load(ctx, cells, tables)
    n = len(cells)
    bound0 = np.mean(np.concatenate([ctx.springs(i)[0] for i in range(n)]))
    ctx.newton_step(stats=False)  # graph capture
    ctx.synchronize()
    step0 = np.array([ctx.cell_info(i)["step"] for i in range(n)])
    iters, calls, failed = 0, 0, None
    t0 = time.perf_counter()
    for k in range(load_steps):
        try:
            st = ctx.load_step()
        except W.WcfeError as e:
            failed = "load step %d: %s" % (k + 1, e)
            log("replica run: " + failed)
            break
        iters += st["iterations"]
        calls += 1
        if k % 100 == 99:
            log("replica run: %d load steps, %d Newton iterations, %.1f s" % (k + 1, iters, time.perf_counter() - t0))
        if time.perf_counter() - t0 > time_limit:
            failed = "stopped at the %.0f s limit after %d load steps" % (time_limit, k + 1)
            break
    ctx.synchronize()
    dt = time.perf_counter() - t0
    steps = np.array([ctx.cell_info(i)["step"] for i in range(n)]) - step0
    bound = np.mean(np.concatenate([ctx.springs(i)[0] for i in range(n)]))
    ctx.close()
    return {"load_step_calls": calls, "load_steps_per_replica_min": int(steps.min()),
            "load_steps_per_replica_max": int(steps.max()), "newton_iterations": iters, "seconds": dt,
            "replica_load_steps_per_s": float(steps.sum()) / dt, "ms_per_newton_iteration": dt * 1e3 / max(iters, 1),
            "bound_fraction_start": float(bound0), "bound_fraction_end": float(bound),
            "newton_atol": o["newton_atol"], "incomplete": failed}


def run_workload(W, configs, name, cells, tables, steps, warmup, common, precision, **opt):
    """Time one workload family; returns its entry of "workloads"."""
    t_setup = time.perf_counter()
    ctx = W.Context(precision=precision, batched=1, use_graph=1, max_refine=4, **dict(common, **opt))
    configs.load(ctx, cells, tables)
    setup = time.perf_counter() - t_setup
    dt = timed_steps(ctx, steps, warmup)
    st = ctx.newton_step(stats=True, check=False)
    prof = ctx.profile_step()
    sched = ctx.schedule()
    sched["two_chain_cells"] = sum(ctx.cell_layout(i)["chain_top"] > 0 for i in range(len(cells)))
    ctx.close()
    n_tri = sum(len(c["tris"]) for c in cells)
    fp64 = precision == W.PREC_FP64
    dense = any(c.get("solver", 0) != 2 for c in cells)
    roof = lu_roofline(prof, bool(sched["fused_step"]), fp64) if dense and prof["lu_gemm"]["launches"] else \
        pcg_roofline(cells, prof, st)
    value = len(cells) * steps / dt
    log("%s: %.3f ms/step, %.0f Newton steps/s (setup %.1f s)" % (name, dt * 1e3 / steps, value, setup))
    return {"value": value, "unit": "Newton steps/s", "ms_per_step": dt * 1e3 / steps, "steps": steps,
            "warmup": warmup, "dtype": "fp64" if fp64 else "fp32+fp64",
            "config": WORKLOAD_TEXT[name] % dict(n=len(cells), t=n_tri, piv="static"), "cells": len(cells),
            "triangles": n_tri, "element_assemblies_per_s": n_tri * steps / dt, "lin_relres_max": st["max_lin_relres"],
            "step_error": st["error"], "refine_sweeps_max": st["max_refine_used"], "pcg_iters_max": st["max_cg_iters"],
            "lu_schedule": sched, "roofline": roof, "phases_ms": {k: round(v["ms"], 4) for k, v in prof.items()}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cells", type=int, default=256, help="cells of the batched256 workload")
    ap.add_argument("--workload", default="batched256",
                    help="headline workload: batched256 (default), single905, single20k, single500k or replicas1024")
    ap.add_argument("--fp64-steps", type=int, default=5)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the oracle check")
    ap.add_argument("--no-extra", action="store_true", help="time the headline workload only")
    ap.add_argument("--pivoting", choices=["static", "partial"], default="static")
    ap.add_argument("--autotune", choices=["none", "solver", "lu", "pcg", "all"], default="none",
                    help="run wcfe_autotune on the headline batch before timing")
    ap.add_argument("--fused-step", type=int, choices=[0, 1, 2, 3, 4, 5], default=None,
                    help="band-LU block-step mode (library default when omitted)")
    ap.add_argument("--twist", type=int, choices=[0, 1, 2], default=None,
                    help="two-chain band LU: 0 off, 1 automatic (default), 2 always")
    ap.add_argument("--pcg-lanes", type=int, choices=[4, 8, 16], default=None,
                    help="block-CSR PCG: lanes per block row in the SpMV (default 8)")
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

    t_setup = time.perf_counter()
    if a.workload == "batched256":
        cells, tables = configs.batched_cells(n_cells=a.cells)
    else:
        cells, tables = configs.workload(a.workload)
    n_tri = sum(len(c["tris"]) for c in cells)
    ndofs = [2 * len(c["X"]) for c in cells]
    log("built %d cells: %d triangles, dof %d..%d in %.1fs" % (len(cells), n_tri, min(ndofs), max(ndofs),
                                                               time.perf_counter() - t_setup))
    piv = W.PIVOT_STATIC if a.pivoting == "static" else W.PIVOT_PARTIAL
    common = dict(lin_rtol=1e-8, newton_rtol=1e-8, newton_atol=1e-12, t0=T0, dt=DT, pivoting=piv)
    if a.fused_step is not None:
        common["fused_step"] = a.fused_step
    if a.twist is not None:
        common["twist"] = a.twist
    if a.pcg_lanes is not None:
        common["pcg_lanes"] = a.pcg_lanes
    all_dense = dict(dense_max_dof=1 << 30) if a.workload == "replicas1024" else {}

    # ---- headline: batched mixed precision, one hipGraph replay per Newton step ----
    ctx = W.Context(precision=W.PREC_MIXED, batched=1, use_graph=1, max_refine=4, **dict(common, **all_dense))
    configs.load(ctx, cells, tables)
    tuned = None
    if a.autotune != "none":
        flags = {"solver": W.TUNE_SOLVER, "lu": W.TUNE_LU, "pcg": W.TUNE_PCG,
                 "all": W.TUNE_SOLVER | W.TUNE_LU | W.TUNE_PCG}[a.autotune]
        tuned = ctx.autotune(flags=flags, steps=2)
        log("autotune (%s): %s" % (a.autotune, tuned))
    dt_run = timed_steps(ctx, a.steps, a.warmup)
    ms_step = dt_run * 1e3 / a.steps
    value = len(cells) * a.steps / dt_run
    log("mixed: %.3f ms/step, %.0f Newton steps/s" % (ms_step, value))
    st = ctx.newton_step(stats=True, check=False)
    prof = ctx.profile_step()
    sched = ctx.schedule()
    layouts = [ctx.cell_layout(i) for i in range(len(cells))]
    sched["two_chain_cells"] = sum(L["chain_top"] > 0 for L in layouts)
    ctx.close()
    log("check step:", st)
    log("phases:", {k: round(v["ms"], 3) for k, v in prof.items()})
    log("LU schedule:", sched)
    dense = [c for c, L in zip(cells, layouts) if L["solver"] != 2]
    if prof["lu_gemm"]["launches"] > 0:
        traffic = measured_traffic("k_lu_step" if sched["fused_step"] else "k_lu_gemm", a.workload, a.pivoting) \
            if a.workload == "batched256" and a.cells == 256 else None
        roofline = lu_roofline(prof, bool(sched["fused_step"]), traffic=traffic)
    else:
        roofline = pcg_roofline(cells, prof, st)
    lu_alg = configs.lu_flops(dense, symmetric=a.pivoting == "static")
    lu_ms = prof["lu_gemm"]["ms"] + prof["lu_strip"]["ms"] + prof["lu_panel"]["ms"]

    # ---- plain FP64 GPU form: same kernels in FP64, one cell at a time, no graph ----
    fp64 = None
    if not a.no_fp64:
        c64 = W.Context(precision=W.PREC_FP64, batched=0, use_graph=0, max_refine=3, **dict(common, **all_dense))
        configs.load(c64, cells, tables)
        dt64 = timed_steps(c64, a.fp64_steps, 1)
        st64 = c64.newton_step(stats=True, check=False)
        c64.close()
        v64 = len(cells) * a.fp64_steps / dt64
        fp64 = {"value": v64, "unit": "Newton steps/s", "ms_per_step": dt64 * 1e3 / a.fp64_steps,
                "steps": a.fp64_steps, "lin_relres_max": st64["max_lin_relres"]}
        log("fp64 plain: %.1f ms/step, %.0f Newton steps/s" % (fp64["ms_per_step"], v64))

    # ---- CPU leg: baseline and checker ----
    cpu, check = None, {}
    if not a.no_cpu and a.workload == "batched256":
        cpu = cpu_baseline(cells, tables)
        log("cpu baseline:", cpu)
        check = oracle_check(cells, tables, common)
        log("oracle check:", check)

    # ---- the other workload families ----
    extra = {}
    if not a.no_extra and a.workload == "batched256":
        for name, prec, steps, opt in (("single905", W.PREC_FP64, 20, {}), ("single20k", W.PREC_MIXED, 10, {}),
                                       ("single500k", W.PREC_MIXED, 5, {}),
                                       ("replicas1024", W.PREC_MIXED, 10, dict(dense_max_dof=1 << 30))):
            cs, tb = configs.workload(name)
            extra[name] = run_workload(W, configs, name, cs, tb, steps, 2, common, prec, **opt)
            if name == "replicas1024":  # the family's run: 500 load steps of every replica
                extra[name]["run_500_load_steps"] = replica_run(W, configs, cs, tb, common, 500, **opt)

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
        "element_assemblies_per_s": n_tri * a.steps / dt_run,
        "newton_rtol": common["newton_rtol"], "lin_rtol": common["lin_rtol"],
        "lin_relres_max": st["max_lin_relres"],
        "step_error": st["error"],
        "refine_sweeps_max": st["max_refine_used"],
        "static_pivots": st["n_static_pivots"],
        "phases": phase_rooflines(cells, layouts, prof, st),
        # the flop rate a dense (full-matrix) LU would need to factor the batch in
        # the time the band LU takes, against the FP32 MFMA peak
        "lu_dense_equivalent_tflops": (lu_alg / (lu_ms * 1e-3)) / 1e12 if lu_ms > 0 else None,
        "phases_ms": {k: round(v["ms"], 4) for k, v in prof.items()},
        "autotuned": tuned,
        "workloads": extra or None,
    }
    out.update(check)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
