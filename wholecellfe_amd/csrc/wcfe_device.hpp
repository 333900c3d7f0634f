// Device-side data views and element mechanics for the gfx950 hot path.
//
// Planar (2 DOF per node) formulation of the reference cell membrane:
//   reference law  UL_growth_explicit.cpp:710-928 (evalElementRe, mat = 1)
//   node springs   UL_growth_explicit.cpp:1198-1303 (calculateForce_FA)
// plus the analytic tangent the reference never formed (UL:889 "lets leave
// the tangent alone").  All element arithmetic is FP64; the dense system may
// be stored in FP32 (mixed precision) or FP64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wcfe_exp.hpp"

namespace wcfe {

constexpr int NB = 64;          // dense LU block size
constexpr int MAX_TABLE = 64;   // knots per force table

enum : int { LAW_UL = 0, LAW_CST = 1 };
enum : int { SPRING_LINEAR = 0, SPRING_TABLE = 1 };
enum : int { KIN_NONE = 0, KIN_MEANFIELD = 1, KIN_BELL = 2 };

// Per-cell parameters, one array entry per cell (device resident).
struct CellParams {
  int law, node_springs, spring_law, table_off, table_n, kinetics;
  int node_off, n_nodes, elem_off, n_elems, spr_off, n_spr;
  int ndof, npad, nblk, sys_off;  // sys_off: offset of the per-cell pivot vector
  int vec_off;                    // offset of the cell's dofs in the batch vectors (2 node_off)
  int sym;                        // 1: symmetric tangent, only the lower triangle is updated
                                  //    and U_kj = D_k L_jk^T is written by transposition
  int csr;                        // 1: block-CSR system + PCG (no dense storage: npad = nblk = 0)
  int bterm;                      // BT_* flags: boundary and body forces of the reference cell model
  int bnd_off, n_bnd;             // the cell's boundary vertices in the batch's boundary tables
  int bw;                         // block bandwidth of the dense system: A_ij = 0 for |i - j| > bw
                                  // (nblk - 1 = full); the LU keeps its fill inside the band
  int env;                        // blocks |i - j| <= env are zeroed before assembly (bw + delayed
                                  // panels - 1: the trailing update reads that far off the band)
  int passive;                    // 1: no active stress (the reference's substrate, mat = 2)
  int tw_m, tw_nb;                // two-chain band LU (tw_m > 0): chain A eliminates blocks
                                  // 0 .. tw_m-1, chain B blocks nblk-1 down to nblk-tw_nb, at
                                  // the same time; then A eliminates the middle window of bw blocks
  long long k_off;                // offset of the dense matrix (elements)
  double mu, E, nu, thick, k_lin, rho_max;
  double k_bend, k_area, area_0, drag, ap_rate, ap_quantum, ap_density;
};

enum : int { BT_BEND = 1, BT_AREA = 2, BT_AP = 4, BT_DRAG = 8, BT_EXT = 16 };

// One chain of fused band-LU block steps over a cell's dense system, in
// virtual block numbering: real block rb(v) = flip ? rmax - v : v (64 x 64
// blocks keep their orientation: the chain factors the block-reversed matrix
// P A P, P the block reversal, which has the same band).  Block steps
// k = -1 .. eliminate blocks 0 .. elim-1; when elim < nblk the step
// k = elim - 1 only applies the last panel's update to the blocks past it
// (the middle window that the other chain eliminates).
//   plain cell      one chain: all nblk blocks
//   two-chain cell  A = blocks 0 .. tw_m + bw - 1 (the top, then the middle
//                   window), B = blocks nblk-1 down to tw_m + bw (the bottom,
//                   reversed) with the middle window as its virtual blocks
//                   tw_nb .. tw_nb + bw - 1, updated and not eliminated.
// A and B touch disjoint blocks except the middle window, which B updates
// up to its step tw_nb - 1 and A from its step tw_m - bw on (tw_nb <=
// tw_m - bw): never in the same launch.
struct Chain {
  int nblk, bw, sym, elim, flip, rmax;
  __host__ __device__ int rb(int v) const { return flip ? rmax - v : v; }
  __host__ __device__ long long off(int bi, int bj, long long ld) const {
    return (long long)(rb(bi) * NB) * ld + (long long)rb(bj) * NB;
  }
};
__host__ __device__ inline Chain chain_of(const CellParams &p, int which) {
  Chain v;
  v.bw = p.bw;
  v.sym = p.sym;
  if (p.tw_m <= 0) {
    v.nblk = v.elim = p.nblk;
    v.flip = v.rmax = 0;
  } else if (which == 0) {
    v.nblk = v.elim = p.tw_m + p.bw;
    v.flip = v.rmax = 0;
  } else {
    v.nblk = p.tw_nb + p.bw;
    v.elim = p.tw_nb;
    v.flip = 1;
    v.rmax = p.nblk - 1;
  }
  return v;
}
// block steps of the cell's longest chain (steps k = -1 .. chain_steps - 2)
__host__ __device__ inline int chain_steps(const CellParams &p) { return p.tw_m > 0 ? p.tw_m + p.bw : p.nblk; }
// Work items of one chain in fused block step k (s0 = k + 1): panel form
// (k >= 0): the panel item, the U items (non-symmetric cells) and one item for
// the trailing blocks and the forward block; L-item form: the diagonal item,
// the L items, the U items, one item per trailing block and the forward
// block; the update-only step: one item per updated block and one for the
// chain's part of the forward sweep of the window rows.
__host__ __device__ inline int chain_items(const Chain &v, int k, int panel) {
  const int s0 = k + 1;
  if (s0 >= v.nblk || s0 > v.elim) return 0;
  if (s0 == v.elim) {  // the window's blocks, then one item for the chain's part of the window rows' forward sweep
    if (k < 0) return 0;
    const int nu = (k + v.bw < v.nblk - 1 ? k + v.bw : v.nblk - 1) - s0 + 1;
    return nu <= 0 ? 0 : (v.sym ? nu * (nu + 1) / 2 : nu * nu) + 1;
  }
  const int w = (v.nblk - s0 - 1 < v.bw) ? v.nblk - s0 - 1 : v.bw;
  int npl = 0;
  if (k >= 0) {
    const int e = (v.nblk - 1 < k + v.bw) ? v.nblk - 1 : k + v.bw;
    npl = e - s0 > 0 ? e - s0 : 0;
  }
  const int tiles = v.sym ? npl * (npl + 1) / 2 : npl * npl;
  // panel 1: one item inverts and solves every L block, one takes the
  // trailing tiles; 2: the same, but the L blocks after the first are L items
  if (panel == 1 && k >= 0) return 2 + (v.sym ? 0 : w);
  if (panel == 2 && k >= 0) return 2 + (w > 1 ? w - 1 : 0) + (v.sym ? 0 : w);
  return 2 + w + (v.sym ? 0 : w) + tiles;
}
__host__ __device__ inline int cell_step_items(const CellParams &p, int k, int panel) {
  int t = chain_items(chain_of(p, 0), k, panel);
  if (p.tw_m > 0) t += chain_items(chain_of(p, 1), k, panel);
  return t;
}

// Boundary tables and per-evaluation scratch of the node terms (batch wide;
// bidx == nullptr when no cell has boundary or body forces).
struct BndView {
  const int *bidx;    // per node: index into the boundary tables, -1 for interior nodes
  const int *node;    // per boundary vertex: global node id
  const int *cst;     // per boundary vertex: candidate range [cst[b], cst[b+1])
  const int2 *cand;   // (nj, nint) candidates, global node ids, in the reference's visiting order
  int2 *frame;        // per boundary vertex: (j1, j2) at the current positions, (-1, -1) if incomplete
  double2 *g;         // per boundary vertex: dA/dx = R (x_j2 - x_j1) / 2
  double *J;          // per boundary vertex: -dF/d(x_i, x_j1, x_j2), 2 x 6 row-major
  double *area;       // per cell: triangle area sum at the current positions
  double *gd;         // per cell: sum_b g_b . d_b (refinement matvec)
  const int *replica; // per cell: Philox stream id
  const double *fext; // per node: constant external load (2 per node; cells with BT_EXT)
  double dt;
  unsigned key_lo, key_hi;
};

// Per-cell mutable solver state.
struct CellState {
  double time;       // load time of the current load step
  double g0;         // ||G|| at the first iterate of the load step
  double gnorm;      // ||G|| of the latest residual
  double bnorm;      // ||rhs|| of the current linear solve
  double rnorm;      // ||b - K d|| of the current linear solve
  double lin_relres; // final relative residual of the last linear solve
  double tau;        // static pivot threshold
  int step;          // committed load steps
  int newton_it;     // Newton iterations in the current load step
  int ir_done;       // refinement converged for the current solve
  int ir_iters;      // refinement sweeps used
  int converged;     // set when the current iterate passed the Newton test
  int n_static;      // tiny pivots replaced in the last factorisation
  int zero_pivot;    // partial pivoting met an exactly zero column
  int degenerate;    // an element had zero area or was inverted
  int lin_failed;    // refinement missed lin_rtol
  int solve_stall;   // a triangular-solve wait hit its bound (results invalid)
};

// Everything an element kernel needs, passed by value.
struct View {
  // nodes (global numbering over the batch)
  const double2 *X;   // reference positions (UL: start of step; CST: rest)
  const double2 *x;   // current positions
  const double *ea;   // 3 per node (UL law)
  const double *cC;   // node springs: bound fraction
  const double2 *fau; // node springs: extension at the start of the step
  // elements (global numbering, colour-major order)
  const int *el_tri;  // 3 per element, global node ids
  const int *el_cell; // owning cell
  const int *el_spr;  // 2 per element: [begin, end) of its hosted barycentric springs
  const int *el_bpos; // 9 per element of a sparse cell: block-CSR index of block (a, b)
  const int *kpos;    // per node: its position in the cell's dense numbering (row block 2 kpos)
  // barycentric springs (grouped by host element)
  const double *sp_bary;   // 3 per spring
  const double *sp_anchor; // 2 per spring
  const double *sp_state;  // 1 per spring
  // per cell
  const CellParams *cp;
  const CellState *cs;
  // force tables
  const double *tab_px, *tab_py, *tab_k;
  double kappa;
};

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
__device__ __forceinline__ void philox4x32_10(unsigned c[4], unsigned k0, unsigned k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const unsigned long long p0 = (unsigned long long)0xD2511F53u * c[0];
    const unsigned long long p1 = (unsigned long long)0xCD9E8D57u * c[2];
    const unsigned n0 = (unsigned)(p1 >> 32) ^ c[1] ^ k0;
    const unsigned n2 = (unsigned)(p0 >> 32) ^ c[3] ^ k1;
    c[1] = (unsigned)p1;
    c[3] = (unsigned)p0;
    c[0] = n0;
    c[2] = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// Uniform in (0, 1): (w + 0.5) 2^-32 with w the first word of Philox(ctr).
__device__ __forceinline__ double philox_uniform(unsigned a, unsigned b, unsigned c, unsigned d, unsigned k0,
                                                 unsigned k1) {
  unsigned ctr[4] = {a, b, c, d};
  philox4x32_10(ctr, k0, k1);
  return ((double)ctr[0] + 0.5) * (1.0 / 4294967296.0);
}

__device__ __forceinline__ double t_act_of(double time) {  // UL:743-757
  return (time < 2.0) ? time * 200.0 : 400.0;
}

// Spring force vector f(u) and its Jacobian J = df/du (row-major 2x2).
// Linear: f = k u.  Table (UL:1258-1264 form): f = F(|u|) u/|u| with F
// piecewise affine, segment j for px_j <= D < px_{j+1}, the last segment
// extrapolating; at |u| = 0 the limit k_0 u is used.
__device__ __forceinline__ void spring_force(const View &v, const CellParams &p, double ux, double uy, double f[2],
                                             double J[4]) {
  if (p.spring_law == SPRING_LINEAR) {
    f[0] = p.k_lin * ux;
    f[1] = p.k_lin * uy;
    J[0] = p.k_lin; J[1] = 0.0; J[2] = 0.0; J[3] = p.k_lin;
    return;
  }
  const double *px = v.tab_px + p.table_off, *py = v.tab_py + p.table_off, *kk = v.tab_k + p.table_off;
  double D2 = ux * ux + uy * uy;
  if (D2 == 0.0) {
    f[0] = kk[0] * ux;
    f[1] = kk[0] * uy;
    J[0] = kk[0]; J[1] = 0.0; J[2] = 0.0; J[3] = kk[0];
    return;
  }
  double D = sqrt(D2);
  // number of knots <= D, by bisection over the (increasing) knots
  int lo = 0, hi = p.table_n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (px[mid] <= D) lo = mid + 1; else hi = mid;
  }
  int j = lo - 1;
  if (j < 0) j = 0;
  double F = py[j] + kk[j] * (D - px[j]);
  double dF = kk[j];
  double nx = ux / D, ny = uy / D;
  double s = F / D;
  f[0] = F * nx;
  f[1] = F * ny;
  J[0] = dF * nx * nx + s * (1.0 - nx * nx);
  J[1] = (dF - s) * nx * ny;
  J[2] = J[1];
  J[3] = dF * ny * ny + s * (1.0 - ny * ny);
}

// Force magnitude |F(D)| of a spring at extension u, D = |u|, with every
// operation rounded as written (no contraction) and the table segment rule of
// UL:1258-1264 (segment j for px_j <= D < px_{j+1}, the last one
// extrapolating): the force the Bell off-rate sees, rounded as a plain FP64
// CPU evaluation of the same expressions rounds it.
__device__ inline double spring_magnitude(const View &v, const CellParams &p, double ux, double uy) {
#pragma clang fp contract(off)
  const double D = sqrt(ux * ux + uy * uy);
  if (p.spring_law == SPRING_LINEAR) return fabs(p.k_lin * D);
  const double *px = v.tab_px + p.table_off, *py = v.tab_py + p.table_off, *kk = v.tab_k + p.table_off;
  int lo = 0, hi = p.table_n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (px[mid] <= D) lo = mid + 1; else hi = mid;
  }
  const int j = lo > 0 ? lo - 1 : 0;
  return fabs(py[j] + kk[j] * (D - px[j]));
}

// Element residual contribution Ge[6] and (if K) tangent Ke[36] (row-major,
// dof order x0,y0,x1,y1,x2,y2), FP64, springs folded in.  Returns false for a
// zero-area or inverted element.
template <bool WANT_K>
__device__ __forceinline__ bool element_system(const View &v, int e, double Ge[6], double Ke[36]) {
  const int c = v.el_cell[e];
  const CellParams &p = v.cp[c];
  const int n0 = v.el_tri[3 * e], n1 = v.el_tri[3 * e + 1], n2 = v.el_tri[3 * e + 2];
  const double2 X0 = v.X[n0], X1 = v.X[n1], X2 = v.X[n2];
  const double2 x0 = v.x[n0], x1 = v.x[n1], x2 = v.x[n2];
  const double time = v.cs[c].time;
  bool ok = true;

  // reference shape-function gradients dN_a (UL:771-774, 810-816 in 2D)
  const double g1x = X1.x - X0.x, g1y = X1.y - X0.y, g2x = X2.x - X0.x, g2y = X2.y - X0.y;
  const double detG = g1x * g2y - g2x * g1y;
  if (!(detG != 0.0)) ok = false;
  const double idG = 1.0 / detG;
  double dN[3][2];
  dN[1][0] = g2y * idG;  dN[1][1] = -g2x * idG;
  dN[2][0] = -g1y * idG; dN[2][1] = g1x * idG;
  dN[0][0] = -(dN[1][0] + dN[2][0]);
  dN[0][1] = -(dN[1][1] + dN[2][1]);
  const double At = 0.5 * fabs(detG) * p.thick;
  const double xs[3][2] = {{x0.x, x0.y}, {x1.x, x1.y}, {x2.x, x2.y}};

  if (WANT_K)
    for (int k = 0; k < 36; ++k) Ke[k] = 0.0;

  if (p.law == LAW_UL) {
    // F = sum_a x_a (x) dN_a  (UL:819-827, planar block)
    double F[2][2];
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) F[i][j] = xs[0][i] * dN[0][j] + xs[1][i] * dN[1][j] + xs[2][i] * dN[2][j];
    const double th = F[0][0] * F[1][1] - F[0][1] * F[1][0];
    if (!(th > 0.0)) ok = false;  // inverted element
    // be_t = (I - 2 ea_avg)^-1 from the nodal strains (UL:782-792)
    const double *e0 = v.ea + 3 * n0, *e1 = v.ea + 3 * n1, *e2 = v.ea + 3 * n2;
    const double exx = (e0[0] + e1[0] + e2[0]) / 3.0, eyy = (e0[1] + e1[1] + e2[1]) / 3.0,
                 exy = (e0[2] + e1[2] + e2[2]) / 3.0;
    const double m00 = 1.0 - 2.0 * exx, m11 = 1.0 - 2.0 * eyy, m01 = -2.0 * exy;
    const double detM = m00 * m11 - m01 * m01;
    const double b00 = m11 / detM, b11 = m00 / detM, b01 = -m01 / detM;
    // S = mu be_t + sigma_a I + p C^-1 with p = -mu / (theta^2 theta_e_t^2)
    // (UL:848-880); first PK-type stress P = F S = F S0 - (beta/theta^2) F^-T
    const double sa = p.passive ? 0.0 : v.kappa * t_act_of(time);
    const double S0[2][2] = {{p.mu * b00 + sa, p.mu * b01}, {p.mu * b01, p.mu * b11 + sa}};
    const double beta = p.mu * detM;  // mu / theta_e_t^2
    const double ith = 1.0 / th;
    const double H[2][2] = {{F[1][1] * ith, -F[1][0] * ith}, {-F[0][1] * ith, F[0][0] * ith}};  // F^-T
    const double c2 = beta * ith * ith;
    double P[2][2];
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) P[i][j] = F[i][0] * S0[0][j] + F[i][1] * S0[1][j] - c2 * H[i][j];
    double h[3][2], Sd[3][2];
    for (int a = 0; a < 3; ++a) {
      // residual Re(a,p) = A t (P dN_a)_p  (UL:893-909)
      Ge[2 * a] = At * (P[0][0] * dN[a][0] + P[0][1] * dN[a][1]);
      Ge[2 * a + 1] = At * (P[1][0] * dN[a][0] + P[1][1] * dN[a][1]);
      h[a][0] = H[0][0] * dN[a][0] + H[0][1] * dN[a][1];
      h[a][1] = H[1][0] * dN[a][0] + H[1][1] * dN[a][1];
      Sd[a][0] = S0[0][0] * dN[a][0] + S0[0][1] * dN[a][1];
      Sd[a][1] = S0[1][0] * dN[a][0] + S0[1][1] * dN[a][1];
    }
    if (WANT_K) {
      // K(a,p;b,q) = A t [ d_pq dN_b.S0 dN_a + beta/theta^2 (2 h_b[q] h_a[p] + h_b[p] h_a[q]) ]
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          const double geo = dN[b][0] * Sd[a][0] + dN[b][1] * Sd[a][1];
          for (int pp = 0; pp < 2; ++pp)
            for (int q = 0; q < 2; ++q)
              Ke[(2 * a + pp) * 6 + 2 * b + q] =
                  At * ((pp == q ? geo : 0.0) + c2 * (2.0 * h[b][q] * h[a][pp] + h[b][pp] * h[a][q]));
        }
    }
  } else {
    // linear plane-stress CST, small strain about the rest positions, with
    // isotropic active stress kappa * t_act(t)
    const double Dc = p.E / (1.0 - p.nu * p.nu);
    const double Xs[3][2] = {{X0.x, X0.y}, {X1.x, X1.y}, {X2.x, X2.y}};
    double exx = 0, eyy = 0, gxy = 0;
    for (int a = 0; a < 3; ++a) {
      const double ux = xs[a][0] - Xs[a][0], uy = xs[a][1] - Xs[a][1];
      exx += dN[a][0] * ux;
      eyy += dN[a][1] * uy;
      gxy += dN[a][1] * ux + dN[a][0] * uy;
    }
    const double sa = p.passive ? 0.0 : v.kappa * t_act_of(time);
    const double sxx = Dc * (exx + p.nu * eyy) + sa;
    const double syy = Dc * (p.nu * exx + eyy) + sa;
    const double sxy = Dc * 0.5 * (1.0 - p.nu) * gxy;
    for (int a = 0; a < 3; ++a) {
      Ge[2 * a] = At * (dN[a][0] * sxx + dN[a][1] * sxy);
      Ge[2 * a + 1] = At * (dN[a][1] * syy + dN[a][0] * sxy);
    }
    if (WANT_K) {
      const double G2 = Dc * 0.5 * (1.0 - p.nu);
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          const double ba = dN[a][0], da = dN[a][1], bb = dN[b][0], db = dN[b][1];
          double *K = Ke + (2 * a) * 6 + 2 * b;
          K[0] = At * (ba * Dc * bb + da * G2 * db);
          K[1] = At * (ba * Dc * p.nu * db + da * G2 * bb);
          K[6] = At * (da * Dc * p.nu * bb + ba * G2 * db);
          K[7] = At * (da * Dc * db + ba * G2 * bb);
        }
    }
  }

  // barycentric springs hosted by this element: G_a += b_a s f(u), u = p(x) - anchor
  for (int s = v.el_spr[2 * e]; s < v.el_spr[2 * e + 1]; ++s) {
    const double st = v.sp_state[s];
    if (st == 0.0) continue;
    const double *bb = v.sp_bary + 3 * s;
    const double ux = bb[0] * xs[0][0] + bb[1] * xs[1][0] + bb[2] * xs[2][0] - v.sp_anchor[2 * s];
    const double uy = bb[0] * xs[0][1] + bb[1] * xs[1][1] + bb[2] * xs[2][1] - v.sp_anchor[2 * s + 1];
    double f[2], J[4];
    spring_force(v, p, ux, uy, f, J);
    for (int a = 0; a < 3; ++a) {
      Ge[2 * a] += bb[a] * st * f[0];
      Ge[2 * a + 1] += bb[a] * st * f[1];
    }
    if (WANT_K)
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          const double w = bb[a] * bb[b] * st;
          double *K = Ke + (2 * a) * 6 + 2 * b;
          K[0] += w * J[0]; K[1] += w * J[1]; K[6] += w * J[2]; K[7] += w * J[3];
        }
  }

  // node-attached springs (UL:1226-1269): G_a += (area_e / 3) cC rho f(u_a),
  // u_a = fau_a + (x_a - X_a), area_e at the current positions
  if (p.node_springs) {
    const double cz = (xs[1][0] - xs[0][0]) * (xs[2][1] - xs[0][1]) - (xs[1][1] - xs[0][1]) * (xs[2][0] - xs[0][0]);
    const double area = 0.5 * fabs(cz);
    const double sg = cz >= 0.0 ? 0.5 : -0.5;
    const double dA[6] = {sg * (xs[1][1] - xs[2][1]), sg * (xs[2][0] - xs[1][0]), sg * (xs[2][1] - xs[0][1]),
                          sg * (xs[0][0] - xs[2][0]), sg * (xs[0][1] - xs[1][1]), sg * (xs[1][0] - xs[0][0])};
    const int nodes[3] = {n0, n1, n2};
    const double2 Xa[3] = {X0, X1, X2};
    for (int a = 0; a < 3; ++a) {
      const int nd = nodes[a];
      const double w = v.cC[nd] * p.rho_max;
      const double2 u0 = v.fau[nd];
      double f[2], J[4];
      spring_force(v, p, u0.x + (xs[a][0] - Xa[a].x), u0.y + (xs[a][1] - Xa[a].y), f, J);
      Ge[2 * a] += (area / 3.0) * w * f[0];
      Ge[2 * a + 1] += (area / 3.0) * w * f[1];
      if (WANT_K) {
        for (int pp = 0; pp < 2; ++pp)
          for (int k = 0; k < 6; ++k) Ke[(2 * a + pp) * 6 + k] += (w * f[pp] / 3.0) * dA[k];
        double *K = Ke + (2 * a) * 6 + 2 * a;
        const double s3 = area / 3.0 * w;
        K[0] += s3 * J[0]; K[1] += s3 * J[1]; K[6] += s3 * J[2]; K[7] += s3 * J[3];
      }
    }
  }
  return ok;
}

// Strain-update pieces of the reference law at the current x (UL:884-887,
// 910-914): Re_ea (3 per corner, [xx,yy,xy]) and lumped mass MMe (1 per corner).
__device__ __forceinline__ void element_strain(const View &v, int e, double Rea[9], double MMe[3]) {
  const int c = v.el_cell[e];
  const CellParams &p = v.cp[c];
  const int nn[3] = {v.el_tri[3 * e], v.el_tri[3 * e + 1], v.el_tri[3 * e + 2]};
  double Xs[3][2], xs[3][2];
  for (int a = 0; a < 3; ++a) {
    const double2 Xa = v.X[nn[a]], xa = v.x[nn[a]];
    Xs[a][0] = Xa.x; Xs[a][1] = Xa.y; xs[a][0] = xa.x; xs[a][1] = xa.y;
  }
  const double g1x = Xs[1][0] - Xs[0][0], g1y = Xs[1][1] - Xs[0][1], g2x = Xs[2][0] - Xs[0][0],
               g2y = Xs[2][1] - Xs[0][1];
  const double detG = g1x * g2y - g2x * g1y, idG = 1.0 / detG;
  double dN[3][2];
  dN[1][0] = g2y * idG;  dN[1][1] = -g2x * idG;
  dN[2][0] = -g1y * idG; dN[2][1] = g1x * idG;
  dN[0][0] = -(dN[1][0] + dN[2][0]);
  dN[0][1] = -(dN[1][1] + dN[2][1]);
  const double At = 0.5 * fabs(detG) * p.thick;
  double F[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) F[i][j] = xs[0][i] * dN[0][j] + xs[1][i] * dN[1][j] + xs[2][i] * dN[2][j];
  const double *e0 = v.ea + 3 * nn[0], *e1 = v.ea + 3 * nn[1], *e2 = v.ea + 3 * nn[2];
  const double exx = (e0[0] + e1[0] + e2[0]) / 3.0, eyy = (e0[1] + e1[1] + e2[1]) / 3.0,
               exy = (e0[2] + e1[2] + e2[2]) / 3.0;
  const double m00 = 1.0 - 2.0 * exx, m11 = 1.0 - 2.0 * eyy, m01 = -2.0 * exy;
  const double detM = m00 * m11 - m01 * m01;
  const double b00 = m11 / detM, b11 = m00 / detM, b01 = -m01 / detM;
  // be = F be_t F^T, ea_s = (I - be^-1) / 2
  const double t00 = F[0][0] * b00 + F[0][1] * b01, t01 = F[0][0] * b01 + F[0][1] * b11;
  const double t10 = F[1][0] * b00 + F[1][1] * b01, t11 = F[1][0] * b01 + F[1][1] * b11;
  const double be00 = t00 * F[0][0] + t01 * F[0][1], be01 = t00 * F[1][0] + t01 * F[1][1];
  const double be11 = t10 * F[1][0] + t11 * F[1][1];
  const double dbe = be00 * be11 - be01 * be01;
  const double es[3] = {0.5 * (1.0 - be11 / dbe), 0.5 * (1.0 - be00 / dbe), 0.5 * (be01 / dbe)};
  const double *ea[3] = {e0, e1, e2};
  for (int a = 0; a < 3; ++a) {
    for (int q = 0; q < 3; ++q) Rea[3 * a + q] = (1.0 / 3.0) * (es[q] - ea[a][q]) * At;
    MMe[a] = (1.0 / 3.0) * At;
  }
}

}  // namespace wcfe
