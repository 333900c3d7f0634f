/* FP64 CPU oracle for the wholeCellFE per-step mechanics hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (wholecellfe_amd/) links,
 * includes or calls this code; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, as the checker.
 *
 * Part 1 restates the reference's functions (wholeCellFE_kMD/cellBP/
 * UL_growth_explicit.cpp = "UL", wholeCellFE_k1/... = "UL1") in plain FP64 C++,
 * 3 components per node exactly as the reference stores them.  It is pinned by
 * golden vectors produced by the compiled reference (oracle/refharness.cpp).
 *
 * Part 2 holds the pieces the reference does not have (planar Newton model,
 * tangent, colouring, dense LU, Jacobi-PCG, Bell kinetics, Philox).  The
 * tangent is the forward-mode derivative of the Part 1 residual; LU/CG/Philox
 * are textbook and pinned by closed forms and published test vectors.
 */
#ifndef WCFE_ORACLE_H
#define WCFE_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- Part 1: reference restatement (3 DOF per node) -------- */

/* UL:710-928 evalElementRe, cell overload.  X,x: 3 nodes x 3; ea: 3 x 6 Voigt
 * [xx,yy,zz,xy,xz,yz]; outputs Re[9], Re_ea[18], MMe[3], ReA[3], ReC[3]. */
void orc_eval_element_re(const double *X, const double *x, const double *ea, const double *cA, const double *cC,
                         double dt, double time, const double *centroid, int mat, double *Re, double *Re_ea,
                         double *MMe, double *ReA, double *ReC);

/* UL:1662-1688 calculateAreaTot -> out[4] = {area, cx, cy, cz}. */
void orc_area_tot(int n_tris, const int32_t *tris, const double *x, double *out);

/* UL:632-707 calculateRes (cell overload, element-order scatter). */
void orc_calculate_res(int n_nodes, int n_tris, const int32_t *tris, const double *X, const double *x,
                       const double *ea, const double *cA, const double *cC, double dt, double time,
                       const double *centroid, int mat, double *Res, double *Res_ea, double *MM, double *ResA,
                       double *ResC);

/* UL:1198-1303 (law 1: 6-knot MD table) / UL1:1196-1295 (law 0: linear kint).
 * fau is 3N; F is 3N. */
void orc_force_fa(int n_nodes, int n_tris, const int32_t *tris, const double *x, const double *cC, const double *fau,
                  int law, double kint, double *F);

/* UL:2026-2174 / UL1:2005-2115 updateFA (11-argument form).  In/out cA, cC, fau. */
void orc_update_fa(int n_nodes, int n_tris, const int32_t *tris, const double *X, const double *x, double *cA,
                   double *cC, double *fau, double dt, int n_snodes, int n_stris, const int32_t *stris,
                   const double *SX, const double *Sx, int law, double kint, int *n_unlocated);

/* UL:1305-1475 calculateForce_FA_subs (law 1: MD table) / UL1:1297 (law 0:
 * linear kint): reaction of the integrin springs on the substrate, F (3 Ns).
 * x_cell, fau: 3N; Sx: current substrate positions, 3 Ns. */
void orc_force_fa_subs(int n_nodes, int n_tris, const int32_t *tris, const double *x_cell, const double *cC,
                       const double *fau, int n_snodes, int n_stris, const int32_t *stris, const double *Sx, int law,
                       double kint, double *F);

/* Boundary vertices and the neighbour candidates the reference's boundary
 * loops visit (UL:1497-1528 and its copies at :1594-1607, :1714-1746), by El
 * Topo's rule (odd edge valence, nondestructivetrimesh.cpp:527-543).
 * is_bnd: N; cand_start: N+1; cand: 2 ints (nj, nint) per candidate,
 * capacity 12 T.  Returns the number of candidates. */
int orc_boundary_candidates(int n_nodes, int n_tris, const int32_t *tris, int32_t *is_bnd, int32_t *cand_start,
                            int32_t *cand);
/* UL:1571-1660 calculateBendingForce2D (k_bend: reference 20 pN/um).  x, F:
 * 3N; only boundary entries of F are written.  Returns the number of boundary
 * vertices missing a curve neighbour (their force is set to 0; the reference
 * reads an uninitialised vector there). */
int orc_bending_force_2d(int n_nodes, int n_tris, const int32_t *tris, const double *x, double k_bend, double *F);
/* UL:1691-1784 calculateAreaConstraintForce (k_area 1, area_0 10*10*3.1416). */
int orc_area_constraint_force(int n_nodes, int n_tris, const int32_t *tris, const double *x, double tot_area,
                              double k_area, double area_0, double *F);
/* UL:1477-1568 calculateForce_AP with the per-vertex draw u[N] in [0,1)
 * supplied (the reference: rand()/(RAND_MAX+1.)); rate 10 1/s, quantum 5 pN,
 * density 100 per um in the reference. */
int orc_force_ap(int n_nodes, int n_tris, const int32_t *tris, const double *x, double dt, double rate,
                 double quantum, double density, const double *u, double *F);

/* ---------------- Part 2: planar Newton model (2 DOF per node) ----------- */

enum { ORC_LAW_UL = 0, ORC_LAW_CST = 1 };
enum { ORC_SPRING_LINEAR = 0, ORC_SPRING_TABLE = 1 };

typedef struct orc_cell {
  int32_t n_nodes, n_tris;
  const double *X;       /* 2N: start-of-step (UL) or rest (CST) positions */
  const double *x;       /* 2N: current positions */
  const int32_t *tris;   /* 3T */
  int32_t law;           /* ORC_LAW_* */
  double mu, E, nu, thick, time;
  const double *ea;      /* 3N planar Euler-Almansi strain [xx,yy,xy] (UL law) */
  /* node-attached springs (reference calculateForce_FA model) */
  int32_t node_springs;
  const double *cC;      /* N */
  const double *fau0;    /* 2N spring extension at the start of the step */
  double rho_max;
  /* barycentric springs (bench configurations) */
  int32_t n_spr;
  const int32_t *spr_tri;
  const double *spr_bary;   /* 3 per spring */
  const double *spr_anchor; /* 2 per spring */
  const double *spr_state;  /* 1 per spring: bound (1) / unbound (0) or bound fraction */
  /* spring force law, shared by both kinds */
  int32_t spring_law;
  double k_lin;
  int32_t tab_n;
  const double *tab_px, *tab_py, *tab_k;
  /* boundary and body forces of the reference cell model, implicit form
   * (UL law): G += drag (x - X)/dt (UL:568-574 with v = (x - X)/dt)
   *             - F_bend(x) (UL:1571-1660) - F_area(x) (UL:1691-1784, area at x)
   *             - F_AP(X)  (UL:1477-1568 at the start-of-step positions, draws ap_u) */
  double k_bend, k_area, area_0, drag, dt;
  const double *ap_u;    /* N draws in (0,1), NULL = no actin polymerisation force */
  double ap_rate, ap_quantum, ap_density;
  /* substrate form (UL:429-504): passive = 1 drops the active stress (mat = 2,
   * t_act = 0, UL:763); f_ext (2N, NULL = none) is a constant nodal load,
   * G -= f_ext (the cell's integrin forces on the substrate, UL:467-470) */
  int32_t passive;
  const double *f_ext;
} orc_cell;

/* Spring force magnitude f(D) and slope f'(D) (UL:1258-1264 table form). */
void orc_spring_law(const orc_cell *c, double D, double *f, double *df);

/* Greedy first-fit element colouring in element-index order.  Returns the
 * number of colours (or -1 if more than 64 are needed). */
int orc_color_elements(int n_nodes, int n_tris, const int32_t *tris, int32_t *color);

/* Equilibrium residual G(x) = R_int(x) - F_spring(x) - boundary/body terms
 * (2N) and, if K != NULL, the dense tangent dG/dx (2N x 2N row-major),
 * assembled colour by colour in element-index order within a colour (color
 * may be NULL: element order), then the node terms.  The area term's
 * rank-one part uses dA/dx_m = R (x_j2 - x_j1) / 2 at boundary vertices
 * (R v = (v_y, -v_x)), the exact area gradient when every boundary vertex has
 * two boundary neighbours. */
void orc_assemble(const orc_cell *c, const int32_t *color, int n_colors, double *G, double *K);

/* Element-level pieces for parity tests: G_e[6], K_e[36] (row-major) of one
 * element including the springs and node-spring terms folded into it. */
void orc_element_system(const orc_cell *c, int e, const int32_t *spr_start, double *Ge, double *Ke);

/* Dense LU, row-major n x n, in place.  pivoting: 1 partial (LAPACK getrf
 * semantics, ipiv[k] = row swapped with k), 0 none.  Returns 0 or (k+1) of the
 * first zero pivot. */
int orc_lu_factor(int n, double *A, int32_t *ipiv, int pivoting);
void orc_lu_solve(int n, const double *LU, const int32_t *ipiv, int pivoting, double *b);
int orc_lu_factor_f32(int n, float *A, int32_t *ipiv, int pivoting);

/* Jacobi-preconditioned CG on CSR (FP64).  Returns iterations or -1. */
int orc_pcg_csr(int n, const int32_t *rowptr, const int32_t *col, const double *val, const double *b, double *x,
                double rtol, int maxit);

/* One full Newton solve of G(x)=0 for one cell (x updated in place).
 * Returns the number of Newton iterations, or -1 on non-convergence.
 * res_hist (optional, maxit+1 entries) receives ||G|| per iterate. */
int orc_newton_solve(orc_cell *c, double *x, double rtol, double atol, int maxit, double *res_hist);

/* ---------------- Part 2, sparse forms (2 DOF per node) ------------------- */
/* Scalar CSR pattern of the tangent: rows 2i and 2i+1 hold columns 2j, 2j+1
 * of every node j sharing a triangle with i (i included), increasing.
 * rowptr: 2N+1, col: capacity cap.  Returns nnz; writes only when cap >= nnz. */
int orc_csr_pattern(int n_nodes, int n_tris, const int32_t *tris, int32_t *rowptr, int32_t *col, int cap);
/* orc_assemble into that pattern (element order).  Cells with bending or
 * area terms (whose tangent leaves the pattern) are refused: returns -1. */
int orc_assemble_csr(const orc_cell *c, const int32_t *rowptr, const int32_t *col, double *val, double *G);
/* Newton solve with FP64 Jacobi-PCG linear solves (cg_rtol per solve);
 * returns iterations or -1 (no convergence), -2 (unsupported cell). */
int orc_newton_solve_pcg(orc_cell *c, double *x, double rtol, double atol, int maxit, double cg_rtol, int cg_maxit,
                         double *res_hist);
/* Reverse Cuthill-McKee node order (order: N, optional); returns the node
 * bandwidth max |pos(i) - pos(j)| over neighbours. */
int orc_band_width(int n_nodes, int n_tris, const int32_t *tris, int32_t *order);
/* One Newton iteration (assemble, FP64 band LU in RCM numbering, solve,
 * x += dx); returns 0, -1 (unsupported cell) or k+1 at a zero pivot. */
int orc_band_newton_step(orc_cell *c, double *x);
/* orc_band_newton_step on every cell (x[i]: 2N each) with OpenMP over the
 * cells, n_threads threads; numbering and storage are prepared before the
 * timed part.  Returns the seconds of the Newton iterations; err receives the
 * number of cells that failed. */
double orc_batch_band_newton_step(int n_cells, orc_cell *cells, double *const *x, int n_threads, int *err);

/* The shared FP64 exp of the kinetics (wholecellfe_amd/csrc/wcfe_exp.hpp). */
void orc_exp_array(int n, const double *x, double *y);

/* Philox4x32-10 (Salmon et al., SC'11). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Uniform in (0,1) from the first word of Philox(ctr=(a,b,c,0), key). */
double orc_uniform(uint32_t key_lo, uint32_t key_hi, uint32_t a, uint32_t b, uint32_t c);

/* Bell-model bond update for barycentric springs after a converged step.
 * state in/out (0/1), anchor in/out (re-anchored at the current point on
 * binding).  Returns number of transitions; min_margin receives min |u - p|. */
int orc_bell_update(const orc_cell *c, double *state, double *anchor, double dt, double k_on, double k_off0,
                    double f_b, uint32_t key_lo, uint32_t key_hi, uint32_t replica, uint32_t step,
                    double *min_margin);

/* Planar commit of the reference model after convergence: ea += Res_ea/MM
 * (UL:607-612) evaluated at x with X, using the planar strain storage. */
void orc_commit_ul_strain(const orc_cell *c, double *ea_out);

#ifdef __cplusplus
}
#endif
#endif
