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

// updateFA 11-argument form (UL:2026-2174, law 1; UL1:2005-2115, law 0), one
// thread per cell node.  The substrate point X - fa_u is located by testing
// every substrate triangle in its reference configuration with the
// reference's z-sign tests; the last containing triangle wins (UL:2098).
__global__ __launch_bounds__(256) void k_ref_update_fa(int n_nodes, const double *X, const double *x, double *cA,
                                                       double *cC, double *fau, double dt, int n_stris,
                                                       const int *stris, const double *SX, const double *Sx, int law,
                                                       double kint, int *n_unlocated) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_nodes) return;
  double cell_u[3], sX[3], subs_u[3] = {0.0, 0.0, 0.0};
  for (int c = 0; c < 3; ++c) {
    cell_u[c] = x[3 * i + c] - X[3 * i + c];
    sX[c] = X[3 * i + c] - fau[3 * i + c];
  }
  bool located = false;
  for (int e = 0; e < n_stris; ++e) {
    const int *t = stris + 3 * e;
    const double *q0 = SX + 3 * t[0], *q1 = SX + 3 * t[1], *q2 = SX + 3 * t[2];
    const double e1[3] = {q1[0] - q0[0], q1[1] - q0[1], q1[2] - q0[2]};
    const double e2[3] = {q2[0] - q0[0], q2[1] - q0[1], q2[2] - q0[2]};
    const double ep[3] = {sX[0] - q0[0], sX[1] - q0[1], sX[2] - q0[2]};
    const double nn[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    const double A_tri = 0.5 * sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
    const double a3[3] = {e1[1] * ep[2] - e1[2] * ep[1], e1[2] * ep[0] - e1[0] * ep[2], e1[0] * ep[1] - e1[1] * ep[0]};
    double A3 = -1.0;
    if (a3[2] >= 0.0) A3 = 0.5 * sqrt(a3[0] * a3[0] + a3[1] * a3[1] + a3[2] * a3[2]);
    const double a2[3] = {ep[1] * e2[2] - ep[2] * e2[1], ep[2] * e2[0] - ep[0] * e2[2], ep[0] * e2[1] - ep[1] * e2[0]};
    double A2 = -1.0;
    if (a2[2] >= 0.0) A2 = 0.5 * sqrt(a2[0] * a2[0] + a2[1] * a2[1] + a2[2] * a2[2]);
    const double A1 = A_tri - A2 - A3;
    if (A1 >= 0.0 && A2 >= 0.0 && A3 >= 0.0) {
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
  const double k_off = ka * exp(force / Fa) + kb * exp(-force / Fb);
  const double cold = cC[i];
  double cnew = cold + dt * k_on * (1.0 - cold) - dt * k_off * cold;
  if (cnew < 0.0) cnew = 0.0;
  if (cA[i] < 0.0) cA[i] = 0.0;
  cC[i] = cnew;
  const double keep = cold - dt * k_off * cold;
  for (int c = 0; c < 3; ++c) fau[3 * i + c] = (keep > 0.0 && cnew > 0.0) ? u[c] * keep / cnew : 0.0 * u[c];
}

void launch_ref_update_fa(hipStream_t s, int n_nodes, const double *X, const double *x, double *cA, double *cC,
                          double *fau, double dt, int n_stris, const int *stris, const double *SX, const double *Sx,
                          int law, double kint, int *n_unlocated) {
  if (n_nodes <= 0) return;
  hipLaunchKernelGGL(k_ref_update_fa, dim3((n_nodes + 255) / 256), dim3(256), 0, s, n_nodes, X, x, cA, cC, fau, dt,
                     n_stris, stris, SX, Sx, law, kint, n_unlocated);
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

}  // namespace wcfe
