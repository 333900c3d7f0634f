// Host-side launcher declarations shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "wcfe_device.hpp"

namespace wcfe {

// Work list of one launch: map[item] = (cell << 20) | (item index within the
// cell), for item in [0, total).
struct WorkList {
  const int *map;
  int total;
};

// ---- assembly and vector kernels (k_assemble.hip) ----
enum AsmMode : int { ASM_FULL = 0, ASM_RESIDUAL = 1, ASM_STRAIN = 2 };

template <class T>
void launch_fill_systems(hipStream_t s, const CellParams *cp, const int *cells, int ncell, int max_npad, T *K,
                         int lower_only = 0);
void launch_fill(hipStream_t s, double *p, long long n, double value);
void launch_copy(hipStream_t s, const void *src, void *dst, long long bytes);
// K: dense systems; Kb: block-CSR values of sparse cells (View::el_bpos)
template <class T>
void launch_assemble(hipStream_t s, const View &v, int e_begin, int e_end, int mode, T *K, T *Kb, double *G,
                     double *Ke64, double *Rea, double *MM);
// whole-batch assembly: element pass over [e_begin, e_end) (all colours),
// then a node gather over [n0, n1) (ASM_FULL or ASM_RESIDUAL)
struct GatherPlan {
  const int *start;   // per node: groups [start[n], start[n+1])
  const int *col;     // per group: column node
  const int *bpos;    // per group: block-CSR position (sparse cells), else -1
  const int *cstart;  // per group: contributions [cstart[g], cstart[g+1])
  const int *ke;      // per contribution: offset of the 2 x 2 block in Ke64
  const int *node;    // per group: its row node
  int g0, g1;         // groups of the node range [n0, n1)
};
template <class T>
void launch_assemble_gather(hipStream_t s, const View &v, int e_begin, int e_end, int mode, const int *node_cell,
                            const int *adj_start, const int *adj, const GatherPlan &gp, int n0, int n1, double *Ge,
                            double *Ke64, double *G, T *K, T *Kb, int lower_only = 0);
void launch_element_dump(hipStream_t s, const View &v, int e_begin, int e_end, double *Ge, double *Ke);
// per-cell norms of cells above NORM_CHUNK entries: partial sums per chunk
// (stride per cell = the batch's largest chunk count) and an arrival counter
// per cell (zero between launches)
constexpr int NORM_CHUNK = 32768;
struct NormWork {
  double *part = nullptr;
  unsigned *cnt = nullptr;
  int stride = 1;
};
// which 1 with lin_rtol >= 0 also applies the refinement control to each
// refining cell (as launch_ir_control with reset 0)
void launch_cell_norm(hipStream_t s, const CellParams *cp, CellState *cs, const int *cells, int ncell,
                      const double *vec, int which, int only_active, double lin_rtol = -1.0,
                      const NormWork &nw = NormWork());
void launch_neg_copy(hipStream_t s, const double *src, double *dst, long long n);
void launch_matvec_ir(hipStream_t s, const View &v, const int *node_cell, const int *adj_start, const int *adj,
                      const double *Ke64, const double *G, const double *delta, double *r, int node_begin,
                      int node_end, const CellState *cs, const BndView &bv);
void launch_ir_control(hipStream_t s, CellState *cs, const int *cells, int ncell, double lin_rtol, int reset);
void launch_perm_gather(hipStream_t s, const int *node_cell, const CellParams *cp, const int *kpos, const double *r,
                        double *rp, int n0, int n1);
void launch_perm_scatter(hipStream_t s, const int *node_cell, const CellParams *cp, const CellState *cs,
                         const int *kpos, const double *dp, double *delta, int n0, int n1, int accumulate);
void launch_update_x(hipStream_t s, double2 *x, const double *delta, int node_begin, int node_end);
void launch_newton_control(hipStream_t s, CellState *cs, const int *cells, int ncell, double rtol, double atol);
void launch_commit_nodes(hipStream_t s, const View &v, const int *node_cell, double2 *X, double2 *x, double *ea,
                         double *cC, double2 *fau, const double *Rea, const double *MM, int node_begin, int node_end,
                         double dt);
void launch_commit_springs(hipStream_t s, const View &v, const int *sp_cell, const int *sp_elem, const int *sp_local,
                           double *state, double *anchor, int s_begin, int s_end, double dt, double k_on,
                           double k_off0, double f_b, unsigned key_lo, unsigned key_hi, const int *replica_of_cell,
                           int *transitions);
void launch_commit_cells(hipStream_t s, CellState *cs, const int *cells, int ncell, double dt);
void launch_remap_nodes(hipStream_t s, int n_new, const int *src0, const int *src1, const double *w, int old_off,
                        const double2 *X, const double2 *x, const double *ea, const double *cC, const double2 *fau,
                        double2 *Xn, double2 *xn, double *ean, double *cCn, double2 *faun);
template <class T>
void launch_diag_tau(hipStream_t s, const CellParams *cp, CellState *cs, const int *cells, int ncell, const T *K);

// ---- boundary and body forces (k_boundary.hip) ----
enum : int { REF_BEND = 0, REF_AREA = 1, REF_AP = 2 };
struct RefBoundaryArgs {
  double k_bend, tot_area, k_area, area_0, dt, rate, quantum, density;
  unsigned key_lo, key_hi, draw;
};
// cst: N+1 candidate offsets (empty range = interior vertex), cand: (nj, nint)
void launch_ref_boundary(hipStream_t s, int which, int n_nodes, const int *cst, const int2 *cand, const double *x,
                         const RefBoundaryArgs &a, const double *u, double *F, int *n_open);
void launch_ref_area_tot(hipStream_t s, int n_tris, const int *tris, const double *x, double *out);
void launch_bnd_frames(hipStream_t s, const BndView &bv, const double2 *x, int b0, int b1);
void launch_cell_area(hipStream_t s, const CellParams *cp, const int *cells, int ncell, const int *crange, int ncolor,
                      const int *el_tri, const double2 *x, double *area);
void launch_cell_gdot(hipStream_t s, const CellParams *cp, const int *cells, int ncell, const BndView &bv,
                      const double *delta);
template <class T>
void launch_node_terms(hipStream_t s, bool full, const View &v, const BndView &bv, const int *node_cell, int n0, int n1,
                       T *K, T *Kb, const int *row_ptr, const int *bcol, const int *diag_pos, double *G);

// ---- dense LU and triangular solves (k_lu.hip) ----
// Dinv holds, per cell and diagonal block, the inverses of the block's unit
// lower and upper triangles (2 x 64 x 64 per block, npad * 128 per cell).
template <class T>
void launch_lu_diag_inv(hipStream_t s, const CellParams *cp, CellState *cs, const WorkList &wl, int k, T *K, T *Dinv);
template <class T>
void launch_lu_panel_pp(hipStream_t s, const CellParams *cp, CellState *cs, const WorkList &wl, int k, T *K,
                        int *ipiv, T *Dinv);
template <class T>
void launch_lu_laswp(hipStream_t s, const CellParams *cp, const WorkList &wl, int k, T *K, const int *ipiv);
template <class T>
void launch_lu_trsm(hipStream_t s, const CellParams *cp, CellState *cs, const WorkList &wl, int k, T *K,
                    const T *Dinv);
template <class T>
void launch_lu_gemm(hipStream_t s, const CellParams *cp, CellState *cs, const WorkList &wl, int k, T *K, T *Dinv,
                    int next_diag, int npanel, int strip, hipEvent_t ev_start = nullptr, hipEvent_t ev_stop = nullptr);
// fused block step k (narrow bands, static pivoting): trailing update of
// panel k, factor of diagonal block k + 1 and the block solves of panel k + 1
// the forward substitution of the first solve inside the fused block steps:
// y (band numbering, per cell at sys_off) = L^-1 P (-G); kinv: per band row
// its batch dof (-1: padding); y null: no forward items work
struct StepFwd {
  float *y = nullptr;
  const int *kinv = nullptr;
  const double *G = nullptr;
};
template <class T>
void launch_lu_step(hipStream_t s, const CellParams *cp, CellState *cs, const WorkList &wl, int k, T *K, T *Dinv,
                    int panel, StepFwd fwd, hipEvent_t ev_start = nullptr, hipEvent_t ev_stop = nullptr);
// every block step of every listed cell in one launch, one workgroup per cell
template <class T>
void launch_lu_cell(hipStream_t s, const CellParams *cp, CellState *cs, const WorkList &wl, T *K, T *Dinv,
                    StepFwd fwd, hipEvent_t ev_start = nullptr, hipEvent_t ev_stop = nullptr);
template <class T>
void launch_lu_solve(hipStream_t s, const CellParams *cp, const CellState *cs, const int *cells, int ncell,
                     const T *K, const T *Dinv, const int *ipiv, int pivoting, const double *r, double *delta,
                     int max_npad, int accumulate);
// static pivoting: forward and backward sweeps with one workgroup per block
// row (fwd: items (cell << 20 | block) by cell, block ascending; bwd: block
// descending) over the cells `cells` (for the any-refining test); yw: the
// exchange buffer, lu_solve_exchange_bytes(system rows) bytes, zeroed once;
// sync: SOLVE_SYNC_INTS ints, zeroed once
constexpr int SOLVE_SYNC_INTS = 4;
template <class T>
void launch_lu_solve_mw(hipStream_t s, const CellParams *cp, CellState *cs, const int *cells, int ncell,
                        const WorkList &fwd, const WorkList &bwd, const T *K, const T *Dinv, const double *r,
                        double *delta, void *yw, int *sync, int accumulate);
size_t lu_solve_exchange_bytes(long long rows, int elem_size);
size_t lu_solve_lds_bytes(int npad, int elem_size);
size_t lu_solve_max_dynamic(size_t max_lds);
// one workgroup per cell, band-limited left-looking sweeps with the next
// block's data prefetched (static pivoting, max_bw <= SOLVE_WG_MAX_BW); ysw
// (batches with two-chain cells): a forward and a backward launch with one
// workgroup per chain, the forward result kept there (sys elements)
template <class T>
void launch_lu_solve_band(hipStream_t s, const CellParams *cp, const CellState *cs, const int *cells, int ncell,
                          const T *K, const T *Dinv, const int *kpos, const double *r, double *delta, int accumulate,
                          int max_npad, int max_bw, const float *yfwd = nullptr, T *ysw = nullptr);
// block bandwidth up to which one workgroup per cell runs the triangular sweeps
constexpr int SOLVE_WG_MAX_BW = 8;
// concurrent cell groups of the dense LU chosen automatically for wide bands
constexpr int AUTO_GROUPS_MAX = 4;
hipError_t lu_prepare_attributes(size_t max_lds);

// ---- sparse path: block-CSR systems and batched PCG (k_csr.hip) ----
constexpr int CG_ROWS = 1024;  // nodes per chunk (one thread per node)
enum CgMode : int { CG_INIT = 0, CG_STEP = 1 };
// per-cell conjugate-gradient state (one entry per cell; used by sparse cells)
struct CgState {
  double rz, rr, rr0, pq, alpha, beta;
  double tol;             // this solve's relative tolerance (set at its start)
  int iters, done, breakdown, total_iters;
  int chunk_off, nchunk;  // the cell's chunks in the chunk list / partials
  int pad_[2];
};
template <class T> struct CgVectors {
  T *x, *r, *z, *p, *q, *s, *dinv;  // 2 N each (dinv: 4 N); z = M^-1 r, q = K z, s = K p
  long long pstride;                // partial sums: three arrays of pstride (one per chunk of the batch)
};
template <class T> void launch_zero(hipStream_t s, T *p, long long n);
template <class T>
void launch_bsr_diag_inv(hipStream_t s, const int4 *chunks, int nchunk, const int *diag_pos, const T *val, T *dinv);
// chunks [cbase, cbase + nchunk) of the batch's chunk list
// lanes: 4, 8 or 16 lanes per block row in the SpMV
// rtol: cg_rtol; lin_rtol: the refinement's target (a later sweep's solve
// stops at max(cg_rtol, 0.1 lin_rtol / the cell's relative residual so far))
template <class T>
void launch_pcg_init(hipStream_t s, const int4 *chunks, int cbase, int nchunk, const int *cells, int ncell,
                     const CellState *cs, CgState *cg, const int *row_ptr, const int *col, const T *val,
                     const double *rhs, const CgVectors<T> &v, double *partials, double rtol, double lin_rtol,
                     int max_iter, int lanes);
// one iteration: the update (step: with the scalars formed from the last
// SpMV's partials, gin -> gout; else with gin's), then the SpMV
template <class T>
void launch_pcg_iteration(hipStream_t s, const int4 *chunks, int cbase, int nchunk, const CgState *gin,
                          CgState *gout, int step, const int *row_ptr, const int *col, const T *val,
                          const CgVectors<T> &v, double *partials, double rtol, int max_iter, int lanes);
// the scalars of the last SpMV by one workgroup per cell, gin -> gout
void launch_pcg_check(hipStream_t s, const int *cells, int ncell, const CellState *cs, const CgState *gin,
                      CgState *gout, const double *partials, long long pstride, double rtol, int max_iter);

template <class T>
void launch_pcg_finish(hipStream_t s, const int4 *chunks, int nchunk, const CellState *cs, const T *x, double *delta,
                       int accumulate);

// ---- calculateForce_FA on device pointers with no allocation (wcfe_noalloc.hip) ----
hipError_t noalloc_init(int device);
int noalloc_force_fa(int device, hipStream_t s, int n_nodes, int n_tris, const int *d_tris, const double *d_x,
                     const double *d_cC, const double *d_fau, int law, double kint, double *d_F, int *d_n_bad,
                     std::string &err);

// ---- reference-shaped stateless kernels (k_reference.hip) ----
void launch_eval_exp(hipStream_t s, long long n, const double *x, double *y);
void launch_ref_force_fa(hipStream_t s, int n_nodes, const int *ring_start, const int *ring, const int *tris,
                         const double *x, const double *cC, const double *fau, int law, double kint, double *F);
// uniform grid over substrate triangles (point location); list[start[c] ..
// start[c+1]) = triangles whose widened bounding box meets cell c, ascending
struct PointGrid {
  double ox, oy, inv_h;
  int gx, gy;
  const int *start, *list;
};
constexpr int SUBS_MAX_HITS = 16;  // containing substrate triangles per cell node (shared vertices)
void launch_ref_update_fa(hipStream_t s, int n_nodes, const double *X, const double *x, double *cA, double *cC,
                          double *fau, double dt, const PointGrid &grid, const int *stris, const double *SX,
                          const double *Sx, int law, double kint, int *n_unlocated);
void launch_ref_force_fa_subs(hipStream_t s, int n_nodes, const int *ring_start, const int *ring, const int *tris,
                              const double *x_cell, const double *cC, const double *fau, const PointGrid &grid,
                              const int *stris, const double *Sx, int law, double kint, int *hit_tri,
                              double *hit_val, int *n_hits, int *overflow);
void launch_dense_residual(hipStream_t s, const CellParams *cp, const CellState *cs, const int *cells, int ncell,
                           const double *A64, const long long *a_off, const double *b, const double *x, double *r);

// Host-side VTK frame writer (wcfe_vtk.hip): one cell or substrate, the
// reference's file layout.  Positions and fa_u with 2 (z = 0) or 3 components
// per node; strain 3 (planar [xx,yy,xy]) or 6 (Voigt) per node.
struct VtkInput {
  const char *path = nullptr;
  int style = 0, n_nodes = 0, n_tris = 0;
  const int32_t *tris = nullptr;
  const double *x = nullptr, *fau = nullptr, *ea = nullptr, *cA = nullptr, *cC = nullptr;
  int x_stride = 3, fau_stride = 3, ea_kind = 6;
  double cA_default = 100.0;  // the reference's initial c_A (main.cpp:674-677)
  int law = 1;
  double kint = 0.0, time = 0.0;
};
int write_vtk_file(const VtkInput &in, std::string &err);

}  // namespace wcfe
