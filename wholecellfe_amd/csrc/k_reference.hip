// Kernels behind the reference-shaped stateless entry points (3 components
// per node, as the reference stores them) and the dense FP64 residual used by
// wcfe_dense_solve.
#include "wcfe_internal.hpp"

namespace wcfe {

// 6-knot force table of the kMD variant, UL_growth_explicit.cpp:1207-1214
// (extension in um, force in pN, stiffness in pN/um).
__constant__ double c_md_px[6] = {0., 2.51250583e-03, 4.40265201e-03, 5.91408228e-03, 11.76817302e-03, 17.001e-03};
__constant__ double c_md_py[6] = {-7.15514497, 71.80501308, 356.62420075, 356.96656556, 126.21811585, 528.40481124};
__constant__ double c_md_k[6] = {31426.8556, 99485.86235, 67080.67086, -7908.39043, 21978.66923, 76858.39733};

// UL:1258-1264: segment j for px_j <= D < px_{j+1}; D >= px_5 extrapolates the
// last segment.  D < 0 cannot occur (D is a norm).
__device__ __forceinline__ double md_force(double D) {
  int j = 0;
  if (D >= c_md_px[5]) {
    j = 5;
  } else {
    for (int q = 0; q < 5; ++q)
      if (D >= c_md_px[q] && D < c_md_px[q + 1]) j = q;
  }
  return c_md_py[j] + c_md_k[j] * (D - c_md_px[j]);
}

// calculateForce_FA (UL:1198-1303, law 1; UL1:1196-1295, law 0), one thread
// per node; the one-ring lists are in increasing triangle order.
__global__ __launch_bounds__(256) void k_ref_force_fa(int n_nodes, const int *ring_start, const int *ring,
                                                      const int *tris, const double *x, const double *cC,
                                                      const double *fau, int law, double kint, double *F) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_nodes) return;
  double Asum = 0.0;
  for (int k = ring_start[i]; k < ring_start[i + 1]; ++k) {
    const int *t = tris + 3 * ring[k];
    const double *p0 = x + 3 * t[0], *p1 = x + 3 * t[1], *p2 = x + 3 * t[2];
    const double e1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    const double e2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    const double nx = e1[1] * e2[2] - e1[2] * e2[1], ny = e1[2] * e2[0] - e1[0] * e2[2],
                 nz = e1[0] * e2[1] - e1[1] * e2[0];
    Asum += 0.5 * sqrt(nx * nx + ny * ny + nz * nz) / 3.0;
  }
  const double *u = fau + 3 * i;
  double fs[3];
  if (law == 0) {
    for (int c = 0; c < 3; ++c) fs[c] = kint * u[c];
  } else {
    const double D = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    const double f = md_force(D);
    for (int c = 0; c < 3; ++c) fs[c] = D == 0.0 ? 0.0 : f * (u[c] / D);
  }
  const double rho_i_max = 100.0;  // UL:1218
  for (int c = 0; c < 3; ++c) F[3 * i + c] = -cC[i] * fs[c] * Asum * rho_i_max;
}

void launch_ref_force_fa(hipStream_t s, int n_nodes, const int *ring_start, const int *ring, const int *tris,
                         const double *x, const double *cC, const double *fau, int law, double kint, double *F) {
  if (n_nodes <= 0) return;
  hipLaunchKernelGGL(k_ref_force_fa, dim3((n_nodes + 255) / 256), dim3(256), 0, s, n_nodes, ring_start, ring, tris, x,
                     cC, fau, law, kint, F);
}

// Point location on the substrate: the reference's containment test (signed
// z of two sub-triangle normals, A1 = A_tri - A2 - A3 >= 0; UL:2070-2095,
// UL:1335-1347) on the triangles of the point's cell of a uniform grid.  A
// cell lists every triangle whose (slightly widened) bounding box meets it,
// in increasing index, so scanning it visits every containing triangle in
// the reference's order.
__device__ __forceinline__ int grid_cell(const PointGrid &g, double px, double py) {
#pragma clang fp contract(off)
  const double fx = floor((px - g.ox) * g.inv_h), fy = floor((py - g.oy) * g.inv_h);
  const int ix = fx >= 0.0 ? (fx < g.gx ? (int)fx : g.gx - 1) : 0;  // NaN falls to cell 0 (no triangle contains it)
  const int iy = fy >= 0.0 ? (fy < g.gy ? (int)fy : g.gy - 1) : 0;
  return iy * g.gx + ix;
}
__device__ __forceinline__ bool tri_contains(const double *q0, const double *q1, const double *q2, const double *p,
                                             double &A_tri, double &A1, double &A2, double &A3) {
#pragma clang fp contract(off)
  const double e1[3] = {q1[0] - q0[0], q1[1] - q0[1], q1[2] - q0[2]};
  const double e2[3] = {q2[0] - q0[0], q2[1] - q0[1], q2[2] - q0[2]};
  const double ep[3] = {p[0] - q0[0], p[1] - q0[1], p[2] - q0[2]};
  const double nn[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
  A_tri = 0.5 * sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
  const double a3[3] = {e1[1] * ep[2] - e1[2] * ep[1], e1[2] * ep[0] - e1[0] * ep[2], e1[0] * ep[1] - e1[1] * ep[0]};
  A3 = -1.0;
  if (a3[2] >= 0.0) A3 = 0.5 * sqrt(a3[0] * a3[0] + a3[1] * a3[1] + a3[2] * a3[2]);
  const double a2[3] = {ep[1] * e2[2] - ep[2] * e2[1], ep[2] * e2[0] - ep[0] * e2[2], ep[0] * e2[1] - ep[1] * e2[0]};
  A2 = -1.0;
  if (a2[2] >= 0.0) A2 = 0.5 * sqrt(a2[0] * a2[0] + a2[1] * a2[1] + a2[2] * a2[2]);
  A1 = A_tri - A2 - A3;
  return A1 >= 0.0 && A2 >= 0.0 && A3 >= 0.0;
}

// updateFA 11-argument form (UL:2026-2174, law 1; UL1:2005-2115, law 0), one
// thread per cell node.  The substrate point X - fa_u is located among the
// substrate triangles (reference configuration) of its grid cell with the
// reference's z-sign tests; the last containing triangle wins (UL:2098).
__global__ __launch_bounds__(256) void k_ref_update_fa(int n_nodes, const double *X, const double *x, double *cA,
                                                       double *cC, double *fau, double dt, PointGrid grid,
                                                       const int *stris, const double *SX, const double *Sx, int law,
                                                       double kint, int *n_unlocated) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_nodes) return;
  double cell_u[3], sX[3], subs_u[3] = {0.0, 0.0, 0.0};
  for (int c = 0; c < 3; ++c) {
    cell_u[c] = x[3 * i + c] - X[3 * i + c];
    sX[c] = X[3 * i + c] - fau[3 * i + c];
  }
  bool located = false;
  const int gc = grid_cell(grid, sX[0], sX[1]);
  for (int k = grid.start[gc]; k < grid.start[gc + 1]; ++k) {  // last containing triangle wins (UL:2098)
    const int *t = stris + 3 * grid.list[k];
    double A_tri, A1, A2, A3;
    if (tri_contains(SX + 3 * t[0], SX + 3 * t[1], SX + 3 * t[2], sX, A_tri, A1, A2, A3)) {
      for (int c = 0; c < 3; ++c)
        subs_u[c] = (A1 / A_tri) * (Sx[3 * t[0] + c] - SX[3 * t[0] + c]) +
                    (A2 / A_tri) * (Sx[3 * t[1] + c] - SX[3 * t[1] + c]) +
                    (A3 / A_tri) * (Sx[3 * t[2] + c] - SX[3 * t[2] + c]);
      located = true;
    }
  }
  if (!located) atomicAdd(n_unlocated, 1);
  double u[3];
  for (int c = 0; c < 3; ++c) u[c] = fau[3 * i + c] + cell_u[c] - subs_u[c];
  double force;
  if (law == 0) {
    force = kint * sqrt(u[0] * u[0] + u[1] * u[1]);  // 2D norm, UL1:2077
  } else {
    const double D = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    const double f = md_force(D);
    double fs[3];
    for (int c = 0; c < 3; ++c) fs[c] = D == 0.0 ? 0.0 : f * (u[c] / D);
    force = sqrt(fs[0] * fs[0] + fs[1] * fs[1] + fs[2] * fs[2]);
  }
  const double k_on = 1.0, ka = 0.004, kb = 10.0, Fa = 15.0, Fb = 15.0;  // UL:2139-2143
  const double k_off = ka * shared_exp(force / Fa) + kb * shared_exp(-force / Fb);
  const double cold = cC[i];
  double cnew = cold + dt * k_on * (1.0 - cold) - dt * k_off * cold;
  if (cnew < 0.0) cnew = 0.0;
  if (cA[i] < 0.0) cA[i] = 0.0;
  cC[i] = cnew;
  const double keep = cold - dt * k_off * cold;
  for (int c = 0; c < 3; ++c) fau[3 * i + c] = (keep > 0.0 && cnew > 0.0) ? u[c] * keep / cnew : 0.0 * u[c];
}

void launch_ref_update_fa(hipStream_t s, int n_nodes, const double *X, const double *x, double *cA, double *cC,
                          double *fau, double dt, const PointGrid &grid, const int *stris, const double *SX,
                          const double *Sx, int law, double kint, int *n_unlocated) {
  if (n_nodes <= 0) return;
  hipLaunchKernelGGL(k_ref_update_fa, dim3((n_nodes + 255) / 256), dim3(256), 0, s, n_nodes, X, x, cA, cC, fau, dt,
                     grid, stris, SX, Sx, law, kint, n_unlocated);
}

// calculateForce_FA_subs (UL:1305-1475, law 1 MD table; UL1:1297, law 0
// linear kint), one thread per cell node: the node's integrin force at the
// substrate point x_cell - fa_u, for every substrate triangle containing it
// in the substrate's current positions, split by barycentric weight.  Each
// hit's 9 contributions are written out (hits in increasing triangle order,
// at most SUBS_MAX_HITS per node) and summed on the host in the reference's
// order.
__global__ __launch_bounds__(256) void k_ref_force_fa_subs(int n_nodes, const int *ring_start, const int *ring,
                                                           const int *tris, const double *x_cell, const double *cC,
                                                           const double *fau, PointGrid grid, const int *stris,
                                                           const double *Sx, int law, double kint, int *hit_tri,
                                                           double *hit_val, int *n_hits, int *overflow) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_nodes) return;
  const double p[3] = {x_cell[3 * i] - fau[3 * i], x_cell[3 * i + 1] - fau[3 * i + 1],
                       x_cell[3 * i + 2] - fau[3 * i + 2]};
  const int gc = grid_cell(grid, p[0], p[1]);
  int h = 0;
  double Asum = -1.0, fs[3] = {0.0, 0.0, 0.0};
  for (int k = grid.start[gc]; k < grid.start[gc + 1]; ++k) {
    const int e = grid.list[k];
    const int *t = stris + 3 * e;
    double A_tri, A1, A2, A3;
    if (!tri_contains(Sx + 3 * t[0], Sx + 3 * t[1], Sx + 3 * t[2], p, A_tri, A1, A2, A3)) continue;
    if (Asum < 0.0) {  // first hit: the node's one-ring area / 3 and its spring force
      Asum = 0.0;
      for (int r = ring_start[i]; r < ring_start[i + 1]; ++r) {
        const int *ct = tris + 3 * ring[r];
        const double *p0 = x_cell + 3 * ct[0], *p1 = x_cell + 3 * ct[1], *p2 = x_cell + 3 * ct[2];
        const double E1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
        const double E2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
        const double nx = E1[1] * E2[2] - E1[2] * E2[1], ny = E1[2] * E2[0] - E1[0] * E2[2],
                     nz = E1[0] * E2[1] - E1[1] * E2[0];
        Asum += 0.5 * sqrt(nx * nx + ny * ny + nz * nz) / 3;
      }
      const double *u = fau + 3 * i;
      if (law == 0) {
        for (int c = 0; c < 3; ++c) fs[c] = kint * u[c];
      } else {
        const double D = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        const double f = md_force(D);
        for (int c = 0; c < 3; ++c) fs[c] = D == 0.0 ? 0.0 : f * (u[c] / D);
      }
    }
    if (h == SUBS_MAX_HITS) {
      atomicAdd(overflow, 1);
      break;
    }
    const double w[3] = {A1 / A_tri, A2 / A_tri, A3 / A_tri};
    hit_tri[i * SUBS_MAX_HITS + h] = e;
    double *v = hit_val + 9LL * (i * SUBS_MAX_HITS + h);
    const double rho_i_max = 100.0;  // UL:1391
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c) v[3 * a + c] = cC[i] * fs[c] * w[a] * Asum * rho_i_max;
    ++h;
  }
  n_hits[i] = h;
}
void launch_ref_force_fa_subs(hipStream_t s, int n_nodes, const int *ring_start, const int *ring, const int *tris,
                              const double *x_cell, const double *cC, const double *fau, const PointGrid &grid,
                              const int *stris, const double *Sx, int law, double kint, int *hit_tri,
                              double *hit_val, int *n_hits, int *overflow) {
  if (n_nodes <= 0) return;
  hipLaunchKernelGGL(k_ref_force_fa_subs, dim3((n_nodes + 255) / 256), dim3(256), 0, s, n_nodes, ring_start, ring,
                     tris, x_cell, cC, fau, grid, stris, Sx, law, kint, hit_tri, hit_val, n_hits, overflow);
}

// r = b - A x in FP64 for dense systems (one workgroup per system, one row per
// thread); systems whose refinement has converged are skipped.
__global__ __launch_bounds__(256) void k_dense_residual(const CellParams *cp, const CellState *cs, const int *cells,
                                                        const double *A64, const long long *a_off, const double *b,
                                                        const double *x, double *r) {
  const int c = cells[blockIdx.x];
  if (cs[c].ir_done) return;
  const CellParams p = cp[c];
  const double *A = A64 + a_off[c];
  const double *bb = b + p.vec_off, *xx = x + p.vec_off;
  double *rr = r + p.vec_off;
  for (int i = threadIdx.x; i < p.ndof; i += 256) {
    double sacc = 0.0;
    const double *row = A + (long long)i * p.ndof;
    for (int j = 0; j < p.ndof; ++j) sacc += row[j] * xx[j];
    rr[i] = bb[i] - sacc;
  }
}
void launch_dense_residual(hipStream_t s, const CellParams *cp, const CellState *cs, const int *cells, int ncell,
                           const double *A64, const long long *a_off, const double *b, const double *x, double *r) {
  if (ncell <= 0) return;
  hipLaunchKernelGGL(k_dense_residual, dim3(ncell), dim3(256), 0, s, cp, cs, cells, A64, a_off, b, x, r);
}

// The shared FP64 exp over an array (diagnostics: bit-for-bit comparison with
// the same function compiled for the CPU).
__global__ __launch_bounds__(256) void k_eval_exp(long long n, const double *x, double *y) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = shared_exp(x[i]);
}
void launch_eval_exp(hipStream_t s, long long n, const double *x, double *y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_eval_exp, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, x, y);
}

}  // namespace wcfe
