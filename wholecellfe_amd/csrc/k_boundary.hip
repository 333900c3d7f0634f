// Boundary and body forces of the reference cell model.
//
// Reference (wholeCellFE_kMD/cellBP/UL_growth_explicit.cpp):
//   calculateForce_AP              :1477-1568  actin polymerisation, outward, stochastic
//   calculateBendingForce2D        :1571-1660  curvature of the boundary curve
//   calculateAreaTot               :1662-1688
//   calculateAreaConstraintForce   :1691-1784  pressure-like penalty on the total area
//   viscous drag (updateForces_Explicit, :568-574)
// Boundary vertices follow El Topo's rule (an edge with an odd number of
// incident triangles is a boundary edge, nondestructivetrimesh.cpp:527-543);
// each boundary vertex carries the list of (nj, nint) candidates the
// reference's loops visit, in their order, built on the host.
//
// Two forms:
//  * reference-shaped stateless kernels (3 components per node), one thread
//    per vertex, arithmetic in the reference's order without contraction;
//  * node terms of the implicit Newton residual G(x) (planar, batched): one
//    thread per node writes only its own rows of G and K, after the element
//    assembly, so the scatter needs no colouring and no atomics.
#include "wcfe_internal.hpp"

namespace wcfe {

namespace {

struct P3 {
  double x, y, z;
};

struct Pos3 {  // 3 components per node (reference layout)
  const double *x;
  __device__ P3 operator()(int i) const { return P3{x[3 * i], x[3 * i + 1], x[3 * i + 2]}; }
};
struct Pos2 {  // planar batch positions
  const double2 *x;
  __device__ P3 operator()(int i) const {
    const double2 v = x[i];
    return P3{v.x, v.y, 0.0};
  }
};

__device__ __forceinline__ double norm3(double a, double b, double c) {
#pragma clang fp contract(off)
  return sqrt(a * a + (b * b + c * c));
}

// j1 (start) and j2 (end) of the boundary curve segment through vertex i:
// (x_i - x_nj) x (x_nint - x_i) . e3 > 0 marks j1, anything else j2; a later
// candidate overwrites an earlier one (UL:1497-1533).  False when either is
// missing (the reference then reads an uninitialised vector).
template <class Pos>
__device__ bool curve_neighbours(const Pos &P, int i, int k0, int k1, const int2 *cand, int &j1, int &j2) {
#pragma clang fp contract(off)
  const P3 xi = P(i);
  bool h1 = false, h2 = false;
  for (int k = k0; k < k1; ++k) {
    const int2 c = cand[k];
    const P3 pb = P(c.x), pn = P(c.y);
    const double ax = xi.x - pb.x, ay = xi.y - pb.y;
    const double bx = pn.x - xi.x, by = pn.y - xi.y;
    if (ax * by - ay * bx > 0) {
      j1 = c.x;
      h1 = true;
    } else {
      j2 = c.x;
      h2 = true;
    }
  }
  return h1 && h2;
}

// normal_i = l1/L n1 + l2/L n2 with n1 = -e3 x (x_i - x_j1) / |.|,
// n2 = -e3 x (x_j2 - x_i) / |.| (UL:1537-1545); -e3 x v = (v_y, -v_x, 0)
struct CurveGeom {
  double n[3];
  double l1, l2;
  double d1[3], d2[3];
};
__device__ CurveGeom curve_geom(const P3 &xi, const P3 &x1, const P3 &x2) {
#pragma clang fp contract(off)
  CurveGeom g;
  g.d1[0] = xi.x - x1.x; g.d1[1] = xi.y - x1.y; g.d1[2] = xi.z - x1.z;
  g.d2[0] = x2.x - xi.x; g.d2[1] = x2.y - xi.y; g.d2[2] = x2.z - xi.z;
  const double m1[3] = {g.d1[1], -g.d1[0], 0.0}, m2[3] = {g.d2[1], -g.d2[0], 0.0};
  const double nm1 = norm3(m1[0], m1[1], m1[2]), nm2 = norm3(m2[0], m2[1], m2[2]);
  g.l1 = norm3(g.d1[0], g.d1[1], g.d1[2]);
  g.l2 = norm3(g.d2[0], g.d2[1], g.d2[2]);
  const double L = g.l1 + g.l2;
  for (int c = 0; c < 3; ++c) g.n[c] = g.l1 / L * (m1[c] / nm1) + g.l2 / L * (m2[c] / nm2);
  return g;
}

// UL:1609-1624: turning angle (cos - 1 for a left turn, 1 - cos otherwise)
__device__ double bend_scale(const CurveGeom &g, double k_bend) {
#pragma clang fp contract(off)
  const double a = norm3(g.d1[0], g.d1[1], g.d1[2]), b = norm3(g.d2[0], g.d2[1], g.d2[2]);
  const double e1[3] = {g.d1[0] / a, g.d1[1] / a, g.d1[2] / a}, e2[3] = {g.d2[0] / b, g.d2[1] / b, g.d2[2] / b};
  const double costurn = e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2];
  const double turna = (e1[0] * e2[1] - e1[1] * e2[0]) > 0 ? costurn - 1 : 1 - costurn;
  const double kappa = turna * 2 / (g.l1 + g.l2);
  return k_bend * kappa;  // F = k_bend kappa normal_i (UL:1644-1647)
}

// UL:1549-1562: the vertex's single draw u fires every one of the
// density * L / 2 trials (integers ii with ii < density L / 2, i.e.
// ceil(density L / 2) of them) when u < rate dt exp(-rate dt).  The trial
// count is capped at 2^24 so that non-finite positions cannot spin the loop.
__device__ double ap_scale(const CurveGeom &g, double u, double dt, double rate, double quantum, double density) {
#pragma clang fp contract(off)
  const double L = g.l1 + g.l2, lim = density * L / 2, pf = rate * dt * shared_exp(-rate * dt);
  const int trials = lim > 0 ? (int)fmin(ceil(lim), 16777216.0) : 0;
  double f = 0;
  if (u < pf)
    for (int ii = 0; ii < trials; ii++) f += quantum;
  return f;
}

__device__ __forceinline__ double area_scale(const CurveGeom &g, double tot_area, double k_area, double area_0) {
#pragma clang fp contract(off)
  return -k_area * (tot_area - area_0) / 2 * (g.l1 + g.l2);  // UL:1776-1781
}

}  // namespace

// ---------------------------------------------------------------------------
// reference-shaped stateless calls
// ---------------------------------------------------------------------------
template <int WHICH>
__global__ __launch_bounds__(256) void k_ref_boundary(int n_nodes, const int *cst, const int2 *cand, const double *x,
                                                      RefBoundaryArgs a, const double *u, double *F, int *n_open) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_nodes || cst[i] == cst[i + 1]) return;  // interior vertices keep F = 0
  const Pos3 P{x};
  int j1 = -1, j2 = -1;
  if (!curve_neighbours(P, i, cst[i], cst[i + 1], cand, j1, j2)) {
    atomicAdd(n_open, 1);
    return;
  }
  const CurveGeom g = curve_geom(P(i), P(j1), P(j2));
  double s;
  if (WHICH == REF_BEND) {
    s = bend_scale(g, a.k_bend);
  } else if (WHICH == REF_AREA) {
    s = area_scale(g, a.tot_area, a.k_area, a.area_0);
  } else {
    const double ui = u ? u[i] : philox_uniform(a.draw, (unsigned)i, 0u, 2u, a.key_lo, a.key_hi);
    s = ap_scale(g, ui, a.dt, a.rate, a.quantum, a.density);
  }
  for (int c = 0; c < 3; ++c) F[3 * i + c] = s * g.n[c];
}

void launch_ref_boundary(hipStream_t s, int which, int n_nodes, const int *cst, const int2 *cand, const double *x,
                         const RefBoundaryArgs &a, const double *u, double *F, int *n_open) {
  if (n_nodes <= 0) return;
  dim3 grid((n_nodes + 255) / 256), block(256);
  if (which == REF_BEND)
    hipLaunchKernelGGL(k_ref_boundary<REF_BEND>, grid, block, 0, s, n_nodes, cst, cand, x, a, u, F, n_open);
  else if (which == REF_AREA)
    hipLaunchKernelGGL(k_ref_boundary<REF_AREA>, grid, block, 0, s, n_nodes, cst, cand, x, a, u, F, n_open);
  else
    hipLaunchKernelGGL(k_ref_boundary<REF_AP>, grid, block, 0, s, n_nodes, cst, cand, x, a, u, F, n_open);
}

// calculateAreaTot (UL:1662-1688): area and area-weighted centroid, one
// workgroup, fixed-order tree reduction.
__global__ __launch_bounds__(256) void k_ref_area_tot(int n_tris, const int *tris, const double *x, double *out) {
  __shared__ double red[4][256];
  double acc[4] = {0, 0, 0, 0};
  for (int e = threadIdx.x; e < n_tris; e += 256) {
    const double *p0 = x + 3 * tris[3 * e], *p1 = x + 3 * tris[3 * e + 1], *p2 = x + 3 * tris[3 * e + 2];
    const double a[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    const double b[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    const double ar = norm3(0.5 * (a[1] * b[2] - a[2] * b[1]), 0.5 * (a[2] * b[0] - a[0] * b[2]),
                            0.5 * (a[0] * b[1] - a[1] * b[0]));
    for (int c = 0; c < 3; ++c) acc[1 + c] += (1. / 3. * (p0[c] + p1[c] + p2[c])) * ar;
    acc[0] += ar;
  }
  for (int q = 0; q < 4; ++q) red[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int q = 0; q < 4; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = red[0][0];
    for (int c = 0; c < 3; ++c) out[1 + c] = red[1 + c][0] / red[0][0];
  }
}

void launch_ref_area_tot(hipStream_t s, int n_tris, const int *tris, const double *x, double *out) {
  hipLaunchKernelGGL(k_ref_area_tot, dim3(1), dim3(256), 0, s, n_tris, tris, x, out);
}

// ---------------------------------------------------------------------------
// node terms of the implicit residual
// ---------------------------------------------------------------------------

// Per boundary vertex: curve neighbours and dA/dx = R (x_j2 - x_j1) / 2 at the
// current positions.
__global__ __launch_bounds__(256) void k_bnd_frames(BndView bv, const double2 *x, int b0, int b1) {
  const int b = b0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= b1) return;
  int j1 = -1, j2 = -1;
  if (!curve_neighbours(Pos2{x}, bv.node[b], bv.cst[b], bv.cst[b + 1], bv.cand, j1, j2)) {
    bv.frame[b] = make_int2(-1, -1);
    bv.g[b] = make_double2(0.0, 0.0);
    return;
  }
  bv.frame[b] = make_int2(j1, j2);
  bv.g[b] = make_double2(0.5 * (x[j2].y - x[j1].y), -0.5 * (x[j2].x - x[j1].x));
}

void launch_bnd_frames(hipStream_t s, const BndView &bv, const double2 *x, int b0, int b1) {
  if (b1 <= b0) return;
  hipLaunchKernelGGL(k_bnd_frames, dim3((b1 - b0 + 255) / 256), dim3(256), 0, s, bv, x, b0, b1);
}

// Per cell with the area term: sum of triangle areas at the current positions
// (the cell's elements are its colour ranges crange[2k], crange[2k+1]).
__global__ __launch_bounds__(256) void k_cell_area(const CellParams *cp, const int *cells, const int *crange,
                                                   int ncolor, const int *el_tri, const double2 *x, double *area) {
  __shared__ double red[256];
  const int c = cells[blockIdx.x];
  if (!(cp[c].bterm & BT_AREA)) return;
  const int *cr = crange + 2LL * ncolor * c;
  double acc = 0;
  for (int k = 0; k < ncolor; ++k)
    for (int e = cr[2 * k] + threadIdx.x; e < cr[2 * k + 1]; e += 256) {
      const double2 p0 = x[el_tri[3 * e]], p1 = x[el_tri[3 * e + 1]], p2 = x[el_tri[3 * e + 2]];
      acc += 0.5 * fabs((p1.x - p0.x) * (p2.y - p0.y) - (p1.y - p0.y) * (p2.x - p0.x));
    }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) area[c] = red[0];
}

void launch_cell_area(hipStream_t s, const CellParams *cp, const int *cells, int ncell, const int *crange, int ncolor,
                      const int *el_tri, const double2 *x, double *area) {
  if (ncell <= 0) return;
  hipLaunchKernelGGL(k_cell_area, dim3(ncell), dim3(256), 0, s, cp, cells, crange, ncolor, el_tri, x, area);
}

// Per cell with the area term: gd = sum over its boundary vertices of
// g_b . d_b (the rank-one part of the area tangent in the refinement matvec).
__global__ __launch_bounds__(256) void k_cell_gdot(const CellParams *cp, const int *cells, BndView bv,
                                                   const double *delta) {
  __shared__ double red[256];
  const int c = cells[blockIdx.x];
  const CellParams &p = cp[c];
  if (!(p.bterm & BT_AREA)) return;
  const double2 *d2 = reinterpret_cast<const double2 *>(delta);
  double acc = 0;
  for (int b = p.bnd_off + threadIdx.x; b < p.bnd_off + p.n_bnd; b += 256) {
    if (bv.frame[b].x < 0) continue;
    const double2 g = bv.g[b], d = d2[bv.node[b]];
    acc += g.x * d.x + g.y * d.y;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) bv.gd[c] = red[0];
}

void launch_cell_gdot(hipStream_t s, const CellParams *cp, const int *cells, int ncell, const BndView &bv,
                      const double *delta) {
  if (ncell <= 0) return;
  hipLaunchKernelGGL(k_cell_gdot, dim3(ncell), dim3(256), 0, s, cp, cells, bv, delta);
}

// One thread per node of the scope: G_n += drag (x_n - X_n)/dt - F_bend(x) -
// F_area(x) - F_AP(X); FULL adds -dF/dx to the node's rows of K (and, for a
// dense cell with the area term, the rank-one coupling k g_n g_m^T to every
// boundary vertex m of the cell) and keeps the 2 x 6 block for the FP64
// refinement matvec.
template <class T, bool FULL>
__global__ __launch_bounds__(256) void k_node_terms(View v, BndView bv, const int *node_cell, int n0, int n1, T *K,
                                                    T *Kb, const int *row_ptr, const int *bcol, const int *diag_pos,
                                                    double *G) {
  const int n = n0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= n1) return;
  const int c = node_cell[n];
  const CellParams &p = v.cp[c];
  if (!p.bterm) return;
  const double2 xn = v.x[n];
  double gx = 0.0, gy = 0.0, cd = 0.0;
  if (p.bterm & BT_EXT) {  // constant load: the cell's integrin forces on a substrate (UL:467-470)
    gx -= bv.fext[2LL * n];
    gy -= bv.fext[2LL * n + 1];
  }
  if (p.bterm & BT_DRAG) {  // F_vd = -drag v, v = (x - X) / dt (UL:568-574)
    cd = p.drag / bv.dt;
    const double2 Xn = v.X[n];
    gx += cd * (xn.x - Xn.x);
    gy += cd * (xn.y - Xn.y);
  }
  const int b = bv.bidx[n];
  double J[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};  // -dF/d(x_n, x_j1, x_j2)
  int2 fr = make_int2(-1, -1);
  if (b >= 0) {
    fr = bv.frame[b];
    if (fr.x >= 0 && (p.bterm & (BT_BEND | BT_AREA))) {
      const Pos2 P{v.x};
      const CurveGeom g = curve_geom(P(n), P(fr.x), P(fr.y));
      double s = 0.0;
      if (p.bterm & BT_BEND) s += bend_scale(g, p.k_bend);
      if (p.bterm & BT_AREA) s += area_scale(g, bv.area[c], p.k_area, p.area_0);
      gx -= s * g.n[0];
      gy -= s * g.n[1];
      if (FULL) {
        const double2 x1 = v.x[fr.x], x2 = v.x[fr.y];
        const double Rs[2] = {x2.y - x1.y, -(x2.x - x1.x)};  // R (x_j2 - x_j1), R v = (v_y, -v_x)
        const double R[2][2] = {{0.0, 1.0}, {-1.0, 0.0}};
        const double dsm[3] = {0.0, -1.0, 1.0};              // d(x_j2 - x_j1) / d(x_n, x_j1, x_j2)
        double dF[3][2][2] = {};
        if (p.bterm & BT_BEND) {
          // F = 2 k t R s / L^2, t = sg (cos - 1), L = |u| + |w|
          const double u[2] = {xn.x - x1.x, xn.y - x1.y}, w[2] = {x2.x - xn.x, x2.y - xn.y};
          const double lu = sqrt(u[0] * u[0] + u[1] * u[1]), lw = sqrt(w[0] * w[0] + w[1] * w[1]);
          const double e1[2] = {u[0] / lu, u[1] / lu}, e2[2] = {w[0] / lw, w[1] / lw};
          const double cth = e1[0] * e2[0] + e1[1] * e2[1];
          const double sg = (e1[0] * e2[1] - e1[1] * e2[0]) > 0 ? 1.0 : -1.0;
          const double t = sg * (cth - 1), L = lu + lw, iL2 = 1.0 / (L * L), iL3 = iL2 / L;
          const double cu[2] = {(e2[0] - cth * e1[0]) / lu, (e2[1] - cth * e1[1]) / lu};
          const double cw[2] = {(e1[0] - cth * e2[0]) / lw, (e1[1] - cth * e2[1]) / lw};
          const double dc[3][2] = {{cu[0] - cw[0], cu[1] - cw[1]}, {-cu[0], -cu[1]}, {cw[0], cw[1]}};
          const double dL[3][2] = {{e1[0] - e2[0], e1[1] - e2[1]}, {-e1[0], -e1[1]}, {e2[0], e2[1]}};
          const double k2 = 2.0 * p.k_bend;
#pragma unroll
          for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int r = 0; r < 2; ++r)
#pragma unroll
              for (int q = 0; q < 2; ++q)
                dF[m][r][q] += k2 * (Rs[r] * (sg * dc[m][q] * iL2 - 2.0 * t * dL[m][q] * iL3) +
                                     t * iL2 * dsm[m] * R[r][q]);
        }
        if (p.bterm & BT_AREA) {
          const double h = -0.5 * p.k_area * (bv.area[c] - p.area_0);
#pragma unroll
          for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int r = 0; r < 2; ++r)
#pragma unroll
              for (int q = 0; q < 2; ++q) dF[m][r][q] += h * dsm[m] * R[r][q];
        }
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
          for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int q = 0; q < 2; ++q) J[6 * r + 2 * m + q] = -dF[m][r][q];
      }
    }
    if (p.bterm & BT_AP) {  // explicit: start-of-step positions, one draw per vertex and load step
      const Pos2 P{v.X};
      int j1 = -1, j2 = -1;
      if (curve_neighbours(P, n, bv.cst[b], bv.cst[b + 1], bv.cand, j1, j2)) {
        const CurveGeom g = curve_geom(P(n), P(j1), P(j2));
        const double u = philox_uniform((unsigned)bv.replica[c], (unsigned)(n - p.node_off),
                                        (unsigned)v.cs[c].step, 1u, bv.key_lo, bv.key_hi);
        const double s = ap_scale(g, u, bv.dt, p.ap_rate, p.ap_quantum, p.ap_density);
        gx -= s * g.n[0];
        gy -= s * g.n[1];
      }
    }
  }
  G[2 * n] += gx;
  G[2 * n + 1] += gy;
  if (!FULL) return;
  if (b >= 0) {
    double *Jb = bv.J + 12LL * b;
#pragma unroll
    for (int k = 0; k < 12; ++k) Jb[k] = J[k];
  }
  J[0] += cd;
  J[7] += cd;
  const bool has_bnd = fr.x >= 0 && (p.bterm & (BT_BEND | BT_AREA));
  const int cols[3] = {n, fr.x, fr.y};
  if (p.csr) {
    // block-CSR cells carry no bending or area term (their tangents need the
    // dense solver): the drag diagonal only
    T *B = Kb + 4LL * diag_pos[n];
    B[0] += T(J[0]);
    B[3] += T(J[7]);
    return;
  }
  T *A = K + p.k_off;
  const long long r0 = 2LL * v.kpos[n];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    if (m > 0 && !has_bnd) break;
    const long long cc = 2LL * v.kpos[cols[m]];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      A[(r0 + r) * p.npad + cc] += T(J[6 * r + 2 * m]);
      A[(r0 + r) * p.npad + cc + 1] += T(J[6 * r + 2 * m + 1]);
    }
  }
  if (has_bnd && (p.bterm & BT_AREA)) {
    // rank one: dG_n/dx_m += k g_n g_m^T over the cell's boundary vertices
    const double2 gn = bv.g[b];
    for (int bb = p.bnd_off; bb < p.bnd_off + p.n_bnd; ++bb) {
      if (bv.frame[bb].x < 0) continue;
      const double2 gm = bv.g[bb];
      const long long cc = 2LL * v.kpos[bv.node[bb]];
      A[r0 * p.npad + cc] += T(p.k_area * gn.x * gm.x);
      A[r0 * p.npad + cc + 1] += T(p.k_area * gn.x * gm.y);
      A[(r0 + 1) * p.npad + cc] += T(p.k_area * gn.y * gm.x);
      A[(r0 + 1) * p.npad + cc + 1] += T(p.k_area * gn.y * gm.y);
    }
  }
}

template <class T>
void launch_node_terms(hipStream_t s, bool full, const View &v, const BndView &bv, const int *node_cell, int n0, int n1,
                       T *K, T *Kb, const int *row_ptr, const int *bcol, const int *diag_pos, double *G) {
  if (n1 <= n0) return;
  dim3 grid((n1 - n0 + 255) / 256), block(256);
  if (full)
    hipLaunchKernelGGL((k_node_terms<T, true>), grid, block, 0, s, v, bv, node_cell, n0, n1, K, Kb, row_ptr, bcol,
                       diag_pos, G);
  else
    hipLaunchKernelGGL((k_node_terms<T, false>), grid, block, 0, s, v, bv, node_cell, n0, n1, K, Kb, row_ptr, bcol,
                       diag_pos, G);
}
template void launch_node_terms<float>(hipStream_t, bool, const View &, const BndView &, const int *, int, int, float *,
                                       float *, const int *, const int *, const int *, double *);
template void launch_node_terms<double>(hipStream_t, bool, const View &, const BndView &, const int *, int, int,
                                        double *, double *, const int *, const int *, const int *, double *);

}  // namespace wcfe
