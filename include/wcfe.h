/* wcfe.h -- C ABI of the MI355X-native per-step cell mechanics hot path.
 *
 * The reference (dredremontes/wholeCellFE) advances each cell with explicit
 * central differences; its hot path is a set of free C++ functions called from
 * advance_sim (wholeCellFE_kMD/cellBP/main.cpp:356-363):
 *   updateFA               UL_growth_explicit.cpp:2026 (kMD) / :2005 (k1, k31)
 *   updateForces_Explicit  UL_growth_explicit.cpp:508  (cell overload)
 *     -> calculateRes      UL_growth_explicit.cpp:632
 *     -> evalElementRe     UL_growth_explicit.cpp:710
 *     -> calculateForce_FA UL_growth_explicit.cpp:1198 (kMD) / :1196 (k1, k31)
 * This library replaces them with HIP kernels for gfx950 behind plain C entry
 * points (pointers and sizes only).  Two families:
 *
 *  (1) Reference-shaped, stateless calls with the reference's argument meaning
 *      (3 components per node, host arrays in and out): wcfe_calculate_res,
 *      wcfe_force_fa, wcfe_force_fa_subs, wcfe_update_fa and the boundary
 *      forces.  They are what include/wcfe_compat.hpp
 *      forwards the reference call sites to.
 *  (2) The implicit (Newton) step the reference lacks: a context holding a
 *      batch of cells resident in HBM, assembled with an analytic tangent,
 *      colour-ordered scatter, dense blocked LU (FP32 MFMA, FP64 iterative
 *      refinement) and an FP64 residual/convergence test.
 *
 * Units follow the reference: um, pN, s, Pa = pN/um^2.
 * Every function returns WCFE_OK (0) or a WCFE_ERR_* code; the message of the
 * last error of a context is available from wcfe_last_error().
 * A context is not thread safe: one host thread per context.
 */
#ifndef WCFE_H
#define WCFE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ---- */
#define WCFE_OK 0
#define WCFE_ERR_ARG 1            /* invalid argument or size */
#define WCFE_ERR_HIP 2            /* HIP runtime error (message has details) */
#define WCFE_ERR_OOM 3            /* device allocation failed */
#define WCFE_ERR_DEGENERATE 4     /* zero-area or inverted element */
#define WCFE_ERR_ZERO_PIVOT 5     /* exact zero pivot in partial-pivot LU */
#define WCFE_ERR_NOT_CONVERGED 6  /* Newton or refinement did not converge */
#define WCFE_ERR_STATE 7          /* call out of order (e.g. before finalize) */
#define WCFE_ERR_NO_DEVICE 8      /* no usable gfx950 device */

/* ---- enumerations ---- */
enum wcfe_law { WCFE_LAW_UL = 0, WCFE_LAW_CST = 1 };
enum wcfe_spring_law { WCFE_SPRING_LINEAR = 0, WCFE_SPRING_TABLE = 1 };
enum wcfe_kinetics { WCFE_KIN_NONE = 0, WCFE_KIN_MEANFIELD = 1, WCFE_KIN_BELL = 2 };
enum wcfe_precision { WCFE_PREC_MIXED = 0, WCFE_PREC_FP64 = 1 };
enum wcfe_pivoting { WCFE_PIVOT_STATIC = 0, WCFE_PIVOT_PARTIAL = 1 };
/* linear solver of a cell: dense blocked LU, or block-CSR + block-Jacobi PCG
 * (both inside the FP64 iterative refinement); AUTO picks CSR above
 * wcfe_options.dense_max_dof unknowns */
enum wcfe_solver { WCFE_SOLVER_AUTO = 0, WCFE_SOLVER_DENSE = 1, WCFE_SOLVER_CSR = 2 };
/* frame file layouts: the reference's cell and substrate write_vtk */
enum wcfe_vtk_style { WCFE_VTK_CELL = 0, WCFE_VTK_SUBSTRATE = 1 };

typedef struct wcfe_ctx wcfe_ctx;

typedef struct wcfe_options {
  int32_t device;        /* HIP device ordinal */
  int32_t precision;     /* wcfe_precision: MIXED = FP32 assembly + LU on MFMA, FP64 residual/IR */
  int32_t pivoting;      /* wcfe_pivoting for the dense LU */
  int32_t batched;       /* 1: one launch sequence over all cells; 0: one cell at a time */
  int32_t use_graph;     /* 1: capture a Newton step in a hipGraph and replay it */
  int32_t max_refine;    /* max iterative-refinement sweeps per linear solve */
  double lin_rtol;       /* ||b - K d|| <= lin_rtol ||b|| (FP64) ends refinement */
  double newton_rtol;    /* ||G(x)|| <= newton_rtol ||G(x0)|| ends a load step */
  double newton_atol;    /* ... or ||G(x)|| <= newton_atol; keep it above the FP64
                            rounding floor of G when a load step can start in
                            equilibrium (a load that has stopped changing) */
  int32_t max_newton;    /* a load step that needs more iterations is reported */
  double dt;             /* load/time step [s] */
  double t0;             /* initial time [s] (active stress schedule UL:743-757) */
  uint32_t seed_lo, seed_hi; /* Philox key for Bell kinetics */
  double k_on, k_off0, f_bell; /* Bell kinetics: binding rate, unloaded off-rate, force scale */
  double kappa;          /* active-stress isotropy (reference: 0.5, UL:872) */
  int32_t symmetric;     /* 1: cells whose tangent is symmetric by construction (no node-attached
                            springs) are factored LDL^T-style with static pivoting: only the lower
                            triangle is updated and U_kj = D_k L_jk^T (half the flops);
                            0: always the full LU */
  int32_t dense_max_dof; /* WCFE_SOLVER_AUTO: cells with more unknowns use the CSR solver */
  double cg_rtol;        /* PCG (working precision) stops at ||r|| <= tol ||r0|| per refinement sweep:
                            tol = cg_rtol in the first sweep; in a later one, when the cell's
                            FP64 relative residual so far is e, tol = max(cg_rtol,
                            0.1 lin_rtol / e) capped at 0.1 (the rest of the way to lin_rtol) */
  int32_t cg_max_iter;   /* PCG iteration cap per refinement sweep */
  int32_t lu_groups;     /* batched dense LU: cells split in this many groups factored on concurrent
                            streams (1..16; 0 = automatic: 1 when every dense band is narrow -- 2 when
                            the batch has two-chain cells -- else 4) */
  int32_t lu_depth;      /* static pivoting: block steps per delayed trailing update (1..8; 0 = automatic:
                            1 when every dense band is narrow, else 4) */
  int32_t band;          /* 1: dense cells under static pivoting are numbered to narrow their band
                            (reverse Cuthill-McKee) and factored inside it (fill stays in the band);
                            0: natural numbering, full matrix */
  int32_t fused_step;    /* with one panel per round (lu_depth 1) in FP32, one launch per block step:
                            trailing update, next diagonal inverse and its block solves together.
                            1: per launch, L items (each inverts the diagonal block itself: shortest
                            chain) while they are at most 2 per compute unit; else split panel
                            items while those are (one item inverts and solves the first L block,
                            the other L blocks are L items, one item takes the cell's trailing
                            blocks); else one panel item per cell (one inverse, then every L
                            block: least work) and one item for the trailing blocks;
                            2: L items only; 3: panel items after the first step; 4: one
                            launch for the whole factorization, each cell's block steps in
                            order in one workgroup (no launch between two steps of a cell's
                            chain); 5: as 1, split panel items wherever 1 takes panel items;
                            0: separate launches */
  int32_t twist;         /* with fused block steps and one-workgroup sweeps in FP32 (block bandwidths
                            up to 4): cells whose chain of block steps is longer than the batch needs
                            are factored as two chains at once (top down and bottom up, meeting in a
                            middle window of bw blocks), about half the dependent block steps.
                            1: when the batch's block steps are latency bound (at most 8 block
                            products per compute unit and step); 2: always; 0: one chain per cell */
  int32_t pcg_lanes;     /* block-CSR PCG: lanes per block row in the SpMV, 4, 8 or 16 (0: 4); a
                            workgroup of 1024 threads per chunk of 256 block rows */
} wcfe_options;

typedef struct wcfe_cell_desc {
  int32_t n_nodes, n_tris;
  const double *X;        /* 2N reference positions: start of step (UL law) / rest (CST law) */
  const double *x;        /* 2N current positions, NULL = X */
  const int32_t *tris;    /* 3T node indices, 0-based */
  int32_t law;            /* wcfe_law */
  double mu;              /* UL law shear modulus [Pa] (reference 1000, UL:724) */
  double E, nu;           /* CST law */
  double thick;           /* [um] (reference 1, UL:768) */
  const double *ea;       /* 3N planar Euler-Almansi strain [xx,yy,xy] (UL law), NULL = 0 */
  /* reference node-attached integrin springs (UL:1198-1303) */
  int32_t node_springs;   /* 0/1 */
  const double *cC;       /* N bound fraction */
  const double *fau;      /* 2N spring extension fa_u */
  double rho_max;         /* [um^-2] (reference 100, UL:1218) */
  /* discrete springs at barycentric points of cell triangles */
  int32_t n_spr;
  const int32_t *spr_tri;   /* host triangle of each spring */
  const double *spr_bary;   /* 3 barycentric coordinates per spring */
  const double *spr_anchor; /* 2 per spring, substrate point */
  const double *spr_state;  /* 1 per spring: 1 bound / 0 unbound, NULL = all bound */
  /* force law shared by both spring kinds */
  int32_t spring_law;     /* wcfe_spring_law */
  double k_lin;           /* [pN/um] linear law */
  int32_t table_id;       /* table law: id given to wcfe_set_table */
  int32_t kinetics;       /* wcfe_kinetics */
  int32_t solver;         /* wcfe_solver */
  /* boundary and body forces of the reference cell model (UL law), in the
   * implicit residual G(x) = ... + drag (x - X)/dt - F_bend(x) - F_area(x) - F_AP(X):
   * bending and area terms at the current positions with their exact tangent,
   * actin polymerisation at the start-of-step positions X with one Philox draw
   * per boundary vertex and load step (key wcfe_options.seed_lo/hi, counter
   * (cell, vertex, step, 1)).  Zero = off. */
  double k_bend;          /* calculateBendingForce2D (UL:1571-1660), reference 20 pN/um */
  double k_area;          /* calculateAreaConstraintForce (UL:1691-1784), reference 1 pN/um^3 */
  double area_0;          /* target area [um^2], reference 10*10*3.1416 (UL:1777) */
  double drag;            /* viscous drag (UL:568-574), reference 0.001 pN s/um; velocity (x - X)/dt */
  int32_t ap;             /* 1: calculateForce_AP (UL:1477-1568) */
  double ap_rate;         /* [1/s], reference 10 (UL:1552) */
  double ap_quantum;      /* [pN] per firing trial, reference 5 (UL:1561) */
  double ap_density;      /* trials per um of boundary (x L/2), reference 100 (UL:1555) */
  /* the reference's substrate (updateForces_Explicit substrate overload,
   * UL:429-504) as a cell of the batch */
  int32_t passive;        /* 1: no active stress (mat = 2: t_act = 0, UL:763); with mu = 1e6 (UL:727) */
  const double *f_ext;    /* 2N constant nodal load, G -= f_ext (calculateForce_FA_subs, UL:467-470);
                             NULL = none; update between load steps with wcfe_set_external_forces */
} wcfe_cell_desc;

typedef struct wcfe_step_stats {
  int32_t n_cells;
  int32_t n_converged;    /* cells whose load step converged at this Newton step */
  int32_t max_refine_used;
  int32_t n_static_pivots;/* tiny pivots replaced (static pivoting) */
  int32_t n_failed;       /* cells whose refinement missed lin_rtol */
  double max_lin_relres;  /* max over cells of the final ||b - K d|| / ||b|| */
  double max_newton_rel;  /* max over cells of ||G(x_new)|| / ||G(x0)|| */
  int32_t max_cg_iters;   /* CSR cells: most PCG iterations of one refinement sweep */
  int32_t n_cg_breakdown; /* CSR cells whose PCG met a non-positive curvature p.Kp <= 0 */
} wcfe_step_stats;

/* ---- lifecycle ---- */
void wcfe_default_options(wcfe_options *opt);
int wcfe_create(const wcfe_options *opt, wcfe_ctx **out);
int wcfe_destroy(wcfe_ctx *ctx);
const char *wcfe_last_error(const wcfe_ctx *ctx);
/* gfx arch name of the context's device, e.g. "gfx950" */
int wcfe_device_arch(const wcfe_ctx *ctx, char *buf, int len);

/* ---- batch definition (before wcfe_finalize) ---- */
/* Force table in the reference's (px, py, k) form (UL:1204-1214): for
 * D in [px_j, px_{j+1}) the force is py_j + k_j (D - px_j); beyond the last
 * knot the last segment extrapolates.  n <= 64. */
int wcfe_set_table(wcfe_ctx *ctx, int32_t table_id, int32_t n, const double *px, const double *py,
                   const double *k);
int wcfe_add_cell(wcfe_ctx *ctx, const wcfe_cell_desc *desc, int32_t *cell_id);
/* Colour elements, lay the batch out in HBM, allocate the dense systems. */
int wcfe_finalize(wcfe_ctx *ctx);

/* ---- topology changes and checkpoints (after wcfe_finalize) ---- */
/* Replace cell `cell` by a new description -- e.g. after remeshing changed its
 * connectivity (edge flips, splits, defragmentation); the caller maps its
 * state onto the new nodes.  The cell keeps its load time and step; every
 * other cell keeps its current state; the batch is coloured and laid out
 * again. */
int wcfe_replace_cell(wcfe_ctx *ctx, int32_t cell, const wcfe_cell_desc *desc);
/* After the host's remeshing of one cell (El Topo's edge flips, splits and
 * collapses, then defragmentation: main.cpp:190-191): the cell's new
 * triangles, with its state mapped on the device from its old nodes.  New
 * node i takes w[i] old[src0[i]] + (1 - w[i]) old[src1[i]] of the positions
 * X and x, the strain ea, the bound fraction cC and the spring extension fau:
 * a kept (or renumbered) node has src0 = src1 = its old index and w = 1, a
 * vertex split from edge (a, b) has a, b and w = 1/2 (El Topo's midpoint
 * interpolation, edgesplitter.cpp:563-577).  src0, src1, w NULL: the nodes
 * are kept as they are.  x_new (2N, optional) replaces the mapped current
 * positions (El Topo's smoothed position of a new vertex).  spr_tri,
 * spr_bary (optional, one per existing spring): the springs' new host
 * triangles and weights; bond states and anchors stay with their springs.
 * The topology is compared by a hash of the triangle list (then element by
 * element): when neither the triangles nor the maps nor the springs change,
 * only x_new is written and nothing is laid out again (relaid = 0);
 * otherwise the cell starts its load step again on the new mesh, the batch is
 * laid out again and every other cell keeps its state bit for bit
 * (relaid = 1).  A constant nodal load (f_ext) is reset to zero: set it again
 * with wcfe_set_external_forces. */
int wcfe_remesh_cell(wcfe_ctx *ctx, int32_t cell, int32_t n_nodes, int32_t n_tris, const int32_t *tris,
                     const int32_t *src0, const int32_t *src1, const double *w, const double *x_new,
                     const int32_t *spr_tri, const double *spr_bary, int32_t *relaid);
/* Binary checkpoint of every cell (mesh, material, positions, strains, node
 * springs, bond states and anchors, load time and step -- the Bell stream
 * position) and the force tables.  Options are not stored: load into a
 * context created with the options wanted, then call wcfe_finalize. */
int wcfe_checkpoint_save(wcfe_ctx *ctx, const char *path);
int wcfe_checkpoint_load(wcfe_ctx *ctx, const char *path);

/* ---- Newton path (after wcfe_finalize) ---- */
/* One batched Newton iteration on every cell: assemble K, G at x; factor;
 * solve K d = -G with refinement to lin_rtol; x += d; FP64 residual at the new
 * x; cells meeting the convergence test commit their load step (kinetics,
 * strain update, time += dt).  stats may be NULL (then nothing is copied back
 * and the call does not synchronise, unless the batch has CSR cells, whose
 * PCG checks convergence on the host every few iterations; such batches are
 * not captured in a graph).  With stats, the step's errors are returned:
 * WCFE_ERR_DEGENERATE, WCFE_ERR_ZERO_PIVOT, and WCFE_ERR_NOT_CONVERGED when a
 * cell's linear solve missed lin_rtol within max_refine sweeps (stats are
 * filled in every case; n_failed counts those cells). */
int wcfe_newton_step(wcfe_ctx *ctx, wcfe_step_stats *stats);
/* Newton iterations until every cell has committed one load step. */
int wcfe_load_step(wcfe_ctx *ctx, int32_t *iterations, wcfe_step_stats *stats);
int wcfe_synchronize(wcfe_ctx *ctx);

/* Autotuner: times `steps` Newton steps of the finalized batch for each
 * candidate setting and keeps the fastest.  WCFE_TUNE_SOLVER: dense / CSR
 * threshold (dense_max_dof, cells with solver AUTO; candidates that would not
 * fit in free device memory are skipped); WCFE_TUNE_LU: delayed-update depth
 * (1, 2, 4, 6) x concurrent cell groups (1, 2, 4), at depth 1 with and without
 * the fused block step, with split panel items (fused_step 5) and (mixed
 * precision, one group) the whole factorization in one launch (fused_step
 * 4); then, with fused block steps,
 * one chain per cell against two-chain cells (twist 0 / 2); WCFE_TUNE_PCG
 * (batches with block-CSR cells): lanes per block row of the PCG's SpMV
 * (pcg_lanes 4, 8, 16).  Every trial starts from the state the cells were
 * added with, and so does the batch afterwards: call it after wcfe_finalize
 * and before any Newton step.  chosen (optional, 6 entries): lu_depth,
 * lu_groups, dense_max_dof, fused_step, the number of two-chain cells and
 * pcg_lanes in effect; ms_best (optional): ms per step. */
#define WCFE_TUNE_SOLVER 1
#define WCFE_TUNE_LU 2
#define WCFE_TUNE_PCG 4
int wcfe_autotune(wcfe_ctx *ctx, int32_t flags, int32_t steps, int32_t *chosen, double *ms_best);
/* The dense-LU schedule in effect after wcfe_finalize (the options, or their
 * automatic choice): delayed-update depth, concurrent cell groups, the widest
 * block band of the dense cells (64-row blocks) and whether the triangular
 * sweeps run one workgroup per cell along the band (1) or spread over
 * workgroups (0), and whether each block step is one fused launch (1) or one
 * launch runs every block step, one workgroup per cell (2). */
int wcfe_get_schedule(wcfe_ctx *ctx, int32_t *lu_depth, int32_t *lu_groups, int32_t *max_band, int32_t *band_solve,
                      int32_t *fused_step);

/* Diagnostics: time `reps` plain device copies of `bytes` bytes (one read and
 * one write of every byte, 16 bytes per thread) on the context's stream; the
 * HBM bandwidth probe and the known-size reference for byte counters. */
int wcfe_copy_probe(wcfe_ctx *ctx, int64_t bytes, int32_t reps, double *ms_per_copy);

/* Diagnostics: the library's FP64 exp (the one the bond kinetics use:
 * Cody-Waite reduction, degree-13 Taylor polynomial in fused multiply-adds,
 * within 1 ulp of exp) evaluated on the device over n host values. */
int wcfe_eval_exp(wcfe_ctx *ctx, int64_t n, const double *x, double *y);

/* Phases of a Newton step, for wcfe_profile_step. */
#define WCFE_PH_FILL 0     /* zero the dense systems and residuals */
#define WCFE_PH_ASSEMBLE 1 /* colour-ordered element assembly of K and G (work = elements) */
#define WCFE_PH_LU_GEMM 2  /* LU trailing updates on MFMA (work = 64x64x64 block products) */
#define WCFE_PH_LU_PANEL 3 /* LU diagonal/panel factorisation, interchanges, block solves */
#define WCFE_PH_SOLVE 4    /* triangular solves */
#define WCFE_PH_REFINE 5   /* FP64 refinement residuals, norms, control */
#define WCFE_PH_RESIDUAL 6 /* update x, FP64 residual at the new x, convergence test */
#define WCFE_PH_COMMIT 7   /* strain update, bond kinetics, time advance */
#define WCFE_PH_LU_STRIP 8 /* LU next-panel strip updates on MFMA (work = block products) */
#define WCFE_NPHASES 9
/* Run one Newton step eagerly with HIP events around every phase; per phase:
 * summed kernel milliseconds, launches, and work units (see above).  The
 * state advances exactly as in wcfe_newton_step.  Diagnostics: the extra
 * events make this slower than a plain step. */
int wcfe_profile_step(wcfe_ctx *ctx, double *phase_ms /* [WCFE_NPHASES] */,
                      int32_t *phase_launches /* [WCFE_NPHASES], optional */,
                      double *phase_work /* [WCFE_NPHASES], optional */);

/* ---- state access ---- */
int wcfe_get_positions(wcfe_ctx *ctx, int32_t cell, double *x /* 2N */);
int wcfe_set_positions(wcfe_ctx *ctx, int32_t cell, const double *x /* 2N */);
int wcfe_get_ul_state(wcfe_ctx *ctx, int32_t cell, double *X /* 2N */, double *ea /* 3N */, double *cC /* N */,
                      double *fau /* 2N */);
/* Restore a cell's start-of-step state (any pointer may be NULL; time NaN
 * keeps the current load time), e.g. to redo a load step at a smaller dt after
 * the host's collision handling shortened it; the next Newton iteration
 * starts a new load step. */
int wcfe_set_ul_state(wcfe_ctx *ctx, int32_t cell, const double *X /* 2N */, const double *ea /* 3N */,
                      const double *cC /* N */, const double *fau /* 2N */, double time);
/* Replace the constant nodal load of a cell added with f_ext != NULL (2N,
 * NULL = zero), e.g. the cell's integrin forces on the substrate from
 * wcfe_force_fa_subs in a staggered cell/substrate step.  It enters the next
 * residual evaluation; call it between load steps. */
int wcfe_set_external_forces(wcfe_ctx *ctx, int32_t cell, const double *f_ext);
/* ---- frame output (the reference's per-frame write_vtk, main.cpp:447-454) ----
 * wcfe_frame_capture snapshots every cell's committed state (start-of-step
 * positions, strain, bound fraction, spring extension, load time) on the
 * device in stream order after the work already enqueued, and copies it to
 * pinned host memory on a separate stream while the following Newton steps
 * run.  Two frames can be in flight: a capture reuses the slot of the frame
 * captured two captures earlier.  wcfe_frame_get and wcfe_frame_write_vtk
 * wait for the frame's copy; a batch laid out again (wcfe_replace_cell)
 * invalidates earlier frames. */
int wcfe_frame_capture(wcfe_ctx *ctx, int32_t *frame);
int wcfe_frame_get(wcfe_ctx *ctx, int32_t frame, int32_t cell, double *X /* 2N */, double *ea /* 3N */,
                   double *cC /* N */, double *fau /* 2N */, double *time);
/* One cell of a captured frame as a VTK file (wcfe_write_vtk's layout; z = 0;
 * cA NULL = the reference's initial 100, main.cpp:674-677). */
int wcfe_frame_write_vtk(wcfe_ctx *ctx, int32_t frame, int32_t cell, const char *path, int32_t style,
                         const double *cA /* N */, int32_t law, double kint);
int wcfe_get_springs(wcfe_ctx *ctx, int32_t cell, double *state /* n_spr */, double *anchor /* 2 n_spr */);
/* Number of cells of the context and one cell's sizes (after wcfe_add_cell or
 * wcfe_checkpoint_load; no wcfe_finalize needed): callers size the host
 * buffers of the getters below from these. */
int wcfe_num_cells(const wcfe_ctx *ctx, int32_t *n_cells);
int wcfe_get_cell_sizes(wcfe_ctx *ctx, int32_t cell, int32_t *n_nodes, int32_t *n_tris, int32_t *n_spr);
/* How a cell is solved after wcfe_finalize: solver (WCFE_SOLVER_DENSE or
 * WCFE_SOLVER_CSR), 64-row blocks of its dense system, their block bandwidth
 * in the band numbering (nblk - 1: full matrix), symmetric factorization, and
 * for a two-chain factorization (wcfe_options.twist) the top blocks of chain
 * A and the bottom blocks of chain B (0, 0: one chain). */
int wcfe_get_cell_layout(wcfe_ctx *ctx, int32_t cell, int32_t *solver, int32_t *nblk, int32_t *bw, int32_t *sym,
                         int32_t *chain_top, int32_t *chain_bottom);
int wcfe_get_cell_info(wcfe_ctx *ctx, int32_t cell, double *time, int32_t *step, int32_t *newton_iter,
                       int32_t *n_colors, int32_t *ndof);
/* FP64 residual G(x) = R_int(x) - F_spring(x) at the current positions (2N). */
int wcfe_residual_fp64(wcfe_ctx *ctx, int32_t cell, double *G);
/* Element systems in the cell's element order: Ge (6 per element) and Ke
 * (36 per element, row-major) with springs folded in, FP64. */
int wcfe_element_systems(wcfe_ctx *ctx, int32_t cell, double *Ge, double *Ke);
/* Assembled dense tangent (FP64 copy of the working-precision matrix, ndof x
 * ndof, row-major) and residual, as the solver sees them. */
int wcfe_assembled_system(wcfe_ctx *ctx, int32_t cell, double *K, double *G);
/* Element colours (int32 per element) in the cell's element order. */
int wcfe_get_colors(wcfe_ctx *ctx, int32_t cell, int32_t *color);

/* ---- dense LU on its own (parity tests, autotuning) ----
 * Factor and solve a batch of row-major n_i x n_i systems in the context's
 * working precision (FP32 for MIXED, FP64 otherwise).  A holds the matrices
 * back to back; b the right-hand sides; x receives the solutions after
 * refinement in FP64 against the FP64 A.  relres (per system) receives
 * ||b - A x|| / ||b||, iters the refinement sweeps used. */
int wcfe_dense_solve(wcfe_ctx *ctx, int32_t n_sys, const int32_t *n, const double *A, const double *b, double *x,
                     double *relres, int32_t *iters, int32_t *ipiv_out /* optional, sum n */);

/* ---- reference-shaped stateless calls (3 components per node) ----
 * wcfe_calculate_res: calculateRes (UL:632-707 cell, mat = 1: mu 1000,
 * active stress; UL:931-987 substrate, mat = 2: mu 1e6, no active stress;
 * constants UL:721-768) on planar meshes (z = 0, planar ea).  X, x: 3N; ea:
 * 6N Voigt [xx,yy,zz,xy,xz,yz]; outputs Res 3N, Res_ea 6N, MM N. */
int wcfe_calculate_res(wcfe_ctx *ctx, int32_t n_nodes, int32_t n_tris, const int32_t *tris, const double *X,
                       const double *x, const double *ea, int32_t mat, double time, double *Res, double *Res_ea,
                       double *MM);
/* calculateForce_FA (UL:1198 kMD table, law 1; UL1:1196 linear kint, law 0).
 * x: 3N, cC: N, fau: 3N, F: 3N. */
int wcfe_force_fa(wcfe_ctx *ctx, int32_t n_nodes, int32_t n_tris, const int32_t *tris, const double *x,
                  const double *cC, const double *fau, int32_t law, double kint, double *F);
/* calculateForce_FA on device-resident arrays, allocating nothing: no
 * counterpart in the reference; for callers that cannot allocate device
 * memory (stream capture, serving loops, harnesses that forbid hipMalloc).
 * The arguments mean what wcfe_force_fa's do, but every array is a device
 * pointer (d_tris 3T, d_x 3N, d_cC N, d_fau 3N, d_F 3N); d_n_bad (optional,
 * device) receives the number of triangle corners outside [0, n_nodes)
 * (skipped).  The one-ring lists live in a static per-device scratch array
 * of the library: n_nodes <= 262144 and n_tris <= 524288 (WCFE_ERR_ARG
 * beyond).  The call is enqueued on the context's stream and returns without
 * waiting (wcfe_synchronize waits); the result equals wcfe_force_fa's bit for
 * bit.  Calls that overlap on one device share the scratch array and race. */
int wcfe_force_fa_noalloc(wcfe_ctx *ctx, int32_t n_nodes, int32_t n_tris, const int32_t *d_tris, const double *d_x,
                          const double *d_cC, const double *d_fau, int32_t law, double kint, double *d_F,
                          int32_t *d_n_bad);
/* updateFA 11-argument form (UL:2026 kMD, law 1; UL1:2005 k1/k31, law 0).
 * cA, cC, fau are updated in place.  n_unlocated (optional) receives the
 * number of anchors outside every substrate triangle (subs_u := 0 there; the
 * reference leaves it uninitialised). */
int wcfe_update_fa(wcfe_ctx *ctx, int32_t n_nodes, const double *X, const double *x, double *cA, double *cC,
                   double *fau, double dt, int32_t n_snodes, int32_t n_stris, const int32_t *stris, const double *SX,
                   const double *Sx, int32_t law, double kint, int32_t *n_unlocated);
/* calculateForce_FA_subs (UL:1305-1475 kMD table, law 1; UL1:1297 linear
 * kint, law 0): the cell's integrin forces applied to the substrate.  Each
 * cell node's force acts at x_cell - fau on every substrate triangle that
 * contains that point in the substrate's current positions Sx, split by
 * barycentric weight and scaled by c_C * (one-ring area / 3) * rho_max (100).
 * Cell: x_cell 3N, cC N, fau 3N; substrate: Sx 3M; F: 3M (overwritten).
 * A point may lie in at most 16 substrate triangles (WCFE_ERR_ARG beyond). */
int wcfe_force_fa_subs(wcfe_ctx *ctx, int32_t n_nodes, int32_t n_tris, const int32_t *tris, const double *x_cell,
                       const double *cC, const double *fau, int32_t n_snodes, int32_t n_stris, const int32_t *stris,
                       const double *Sx, int32_t law, double kint, double *F);
/* write_vtk (iomesh.cpp:844-1047, style WCFE_VTK_CELL: c_A, theta_e_t, c_C and
 * the spring force per node -- law 1: the MD table at |fa_u| (kMD), law 0:
 * |kint fa_u_xy| (k1, k31) -- then fa_u; iomesh.cpp:730-842, style
 * WCFE_VTK_SUBSTRATE: theta_e_t, then fa_u, fixed notation).  Host only, no
 * context.  x, fau: 3N; ea: 6N Voigt (NULL = 0); cA, cC: N (cell style; NULL
 * = 100 and 0). */
int wcfe_write_vtk(const char *path, int32_t style, int32_t n_nodes, int32_t n_tris, const int32_t *tris,
                   const double *x, const double *ea, const double *fau, const double *cA, const double *cC,
                   int32_t law, double kint, double time);
/* calculateAreaTot (UL:1662-1688): out[4] = {area, centroid x, y, z}; x: 3N. */
int wcfe_area_tot(wcfe_ctx *ctx, int32_t n_nodes, int32_t n_tris, const int32_t *tris, const double *x,
                  double *out);
/* Boundary forces, one GPU thread per boundary vertex (El Topo's boundary
 * rule, nondestructivetrimesh.cpp:527-543; the reference's neighbour search
 * and "last candidate wins" rule, UL:1497-1533).  x, F: 3N; F is overwritten
 * (0 at interior vertices).  n_open (optional) receives the number of boundary
 * vertices without both curve neighbours: their force is 0 (the reference
 * reads an uninitialised vector there).
 * calculateBendingForce2D (UL:1571-1660); reference k_bend = 20 pN/um. */
int wcfe_bending_force_2d(wcfe_ctx *ctx, int32_t n_nodes, int32_t n_tris, const int32_t *tris, const double *x,
                          double k_bend, double *F, int32_t *n_open);
/* calculateAreaConstraintForce (UL:1691-1784) with the caller's tot_area
 * (calculateAreaTot at x, UL:527); reference k_area = 1, area_0 = 10*10*3.1416. */
int wcfe_area_constraint_force(wcfe_ctx *ctx, int32_t n_nodes, int32_t n_tris, const int32_t *tris, const double *x,
                               double tot_area, double k_area, double area_0, double *F, int32_t *n_open);
/* calculateForce_AP (UL:1477-1568): per boundary vertex one draw u in [0,1)
 * (u: N values, the reference's rand()/(RAND_MAX+1.)); the force is
 * quantum * ceil(density L / 2) along the vertex normal when
 * u < rate dt exp(-rate dt), else 0.  u == NULL draws from the context's
 * Philox key with counter (draw, vertex, 0, 2).  Reference: rate 10 1/s,
 * quantum 5 pN, density 100 per um. */
int wcfe_force_ap(wcfe_ctx *ctx, int32_t n_nodes, int32_t n_tris, const int32_t *tris, const double *x, double dt,
                  double rate, double quantum, double density, const double *u, uint32_t draw, double *F,
                  int32_t *n_open);

#ifdef __cplusplus
}
#endif
#endif
