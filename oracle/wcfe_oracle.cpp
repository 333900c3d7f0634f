// FP64 CPU oracle for the wholeCellFE per-step mechanics hot path.
//
// TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py as the checker.  The product never links it.
//
// Part 1 restates the reference functions (file:line cited per function).
// Part 2 is the planar Newton model the reference lacks; its tangent is the
// forward-mode derivative (dual numbers) of the Part 1 residual, so it is an
// independent derivation from the GPU kernels' hand-derived tangents.
//
// Compiled with -ffp-contract=off: the reference build had no FMA.
//
// exp: the kinetics (mean-field k_off, Bell probabilities, actin
// polymerisation) use the library's shared exp (wcfe_exp.hpp: Cody-Waite
// reduction and a Taylor polynomial in fused multiply-adds), the one the GPU
// kinetics use, so both round every bond decision identically; it is pinned
// against the C library's exp to 1 ulp (tests/test_oracle_pins.py).
#include "wcfe_oracle.h"

#include "../wholecellfe_amd/csrc/wcfe_exp.hpp"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <map>
#include <vector>

namespace {

// ------------------------------------------------------------------------
// Scalars: double, or a forward-mode dual number carrying ND derivatives.
// ------------------------------------------------------------------------
template <int ND> struct Dual {
  double v;
  double d[ND];
  Dual() : v(0) { for (int i = 0; i < ND; ++i) d[i] = 0; }
  Dual(double a) : v(a) { for (int i = 0; i < ND; ++i) d[i] = 0; }
};
template <int ND> Dual<ND> operator+(const Dual<ND> &a, const Dual<ND> &b) {
  Dual<ND> r(a.v + b.v);
  for (int i = 0; i < ND; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
template <int ND> Dual<ND> operator-(const Dual<ND> &a, const Dual<ND> &b) {
  Dual<ND> r(a.v - b.v);
  for (int i = 0; i < ND; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
template <int ND> Dual<ND> operator-(const Dual<ND> &a) {
  Dual<ND> r(-a.v);
  for (int i = 0; i < ND; ++i) r.d[i] = -a.d[i];
  return r;
}
template <int ND> Dual<ND> operator*(const Dual<ND> &a, const Dual<ND> &b) {
  Dual<ND> r(a.v * b.v);
  for (int i = 0; i < ND; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int ND> Dual<ND> operator/(const Dual<ND> &a, const Dual<ND> &b) {
  Dual<ND> r(a.v / b.v);
  for (int i = 0; i < ND; ++i) r.d[i] = (a.d[i] * b.v - a.v * b.d[i]) / (b.v * b.v);
  return r;
}
template <int ND> Dual<ND> operator+(const Dual<ND> &a, double b) { return a + Dual<ND>(b); }
template <int ND> Dual<ND> operator+(double a, const Dual<ND> &b) { return Dual<ND>(a) + b; }
template <int ND> Dual<ND> operator-(const Dual<ND> &a, double b) { return a - Dual<ND>(b); }
template <int ND> Dual<ND> operator-(double a, const Dual<ND> &b) { return Dual<ND>(a) - b; }
template <int ND> Dual<ND> operator*(const Dual<ND> &a, double b) { return a * Dual<ND>(b); }
template <int ND> Dual<ND> operator*(double a, const Dual<ND> &b) { return Dual<ND>(a) * b; }
template <int ND> Dual<ND> operator/(const Dual<ND> &a, double b) { return a / Dual<ND>(b); }
template <int ND> Dual<ND> operator/(double a, const Dual<ND> &b) { return Dual<ND>(a) / b; }
template <int ND> Dual<ND> sqrt(const Dual<ND> &a) {
  double s = ::sqrt(a.v);
  Dual<ND> r(s);
  for (int i = 0; i < ND; ++i) r.d[i] = a.d[i] / (2.0 * s);
  return r;
}
inline double sqrt(double a) { return ::sqrt(a); }
inline double val(double a) { return a; }
template <int ND> double val(const Dual<ND> &a) { return a.v; }

typedef Dual<6> D6;

// 3-vectors and 3x3 matrices (row-major m[3*i+j]) over a scalar S.
template <class S> struct V3 { S a[3]; };
template <class S> struct M3 { S m[9]; };

template <class S> V3<S> cross(const V3<S> &u, const V3<S> &w) {
  V3<S> r;
  r.a[0] = u.a[1] * w.a[2] - u.a[2] * w.a[1];
  r.a[1] = u.a[2] * w.a[0] - u.a[0] * w.a[2];
  r.a[2] = u.a[0] * w.a[1] - u.a[1] * w.a[0];
  return r;
}
template <class S> S norm(const V3<S> &u) { return sqrt(u.a[0] * u.a[0] + (u.a[1] * u.a[1] + u.a[2] * u.a[2])); }
template <class S> S dot3(const V3<S> &u, const V3<S> &w) { return u.a[0] * w.a[0] + (u.a[1] * w.a[1] + u.a[2] * w.a[2]); }

// Eigen 3.3.4 LU/Determinant.h:18-23,61-69 (row-0 expansion).
template <class S> S det3(const M3<S> &M) {
  const S *m = M.m;
  return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) + m[2] * (m[3] * m[7] - m[4] * m[6]);
}
// Eigen 3.3.4 LU/InverseImpl.h:126-170: cofactors of column 0, det from
// column 0, invdet, adjugate * invdet.
template <class S> S cof(const M3<S> &M, int i, int j) {
  int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
  return M.m[3 * i1 + j1] * M.m[3 * i2 + j2] - M.m[3 * i1 + j2] * M.m[3 * i2 + j1];
}
template <class S> M3<S> inv3(const M3<S> &M) {
  S c0 = cof(M, 0, 0), c1 = cof(M, 1, 0), c2 = cof(M, 2, 0);
  S det = c0 * M.m[0] + (c1 * M.m[3] + c2 * M.m[6]);
  S invdet = S(1.0) / det;
  M3<S> R;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R.m[3 * i + j] = cof(M, j, i) * invdet;
  return R;
}
template <class S> M3<S> matmul(const M3<S> &A, const M3<S> &B) {
  M3<S> R;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      R.m[3 * i + j] = A.m[3 * i] * B.m[j] + (A.m[3 * i + 1] * B.m[3 + j] + A.m[3 * i + 2] * B.m[6 + j]);
  return R;
}
template <class S> M3<S> transpose(const M3<S> &A) {
  M3<S> R;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R.m[3 * i + j] = A.m[3 * j + i];
  return R;
}

// ------------------------------------------------------------------------
// UL:710-928 evalElementRe (cell overload), generic in the scalar of the
// current positions x.
// ------------------------------------------------------------------------
// Material constants come from ul_constants() for the reference restatement
// (hard-coded there) and from the cell parameters for the planar model.
void ul_constants(int mat, double time, double *mu, double *t_act, double *thick) {
  *mu = 0;  // UL:721-740
  if (mat == 1) *mu = 1000;
  else if (mat == 2) *mu = 1000000;
  *t_act = 0;  // UL:743-763
  if (mat == 1) *t_act = (time < 2) ? time * 200.0 : 400;
  *thick = 1;  // UL:768
}

template <class S>
void ul_element(const double *Xd, const S *x, const double *ea, const double *cA, const double *cC, double dt_real,
                double mu, double t_act, double thick, const double *centroid, S *Re, S *Re_ea, S *MMe, S *ReA,
                S *ReC) {
  const double thetag_dot = 0.0;  // UL:714

  V3<S> X0, X1, X2, x0, x1, x2;
  for (int c = 0; c < 3; ++c) {
    X0.a[c] = S(Xd[c]); X1.a[c] = S(Xd[3 + c]); X2.a[c] = S(Xd[6 + c]);
    x0.a[c] = x[c]; x1.a[c] = x[3 + c]; x2.a[c] = x[6 + c];
  }
  // reference metric and normal UL:771-774
  V3<S> G1, G2;
  for (int c = 0; c < 3; ++c) { G1.a[c] = X1.a[c] - X0.a[c]; G2.a[c] = X2.a[c] - X0.a[c]; }
  V3<S> G1xG2 = cross(G1, G2);
  S nG = norm(G1xG2);
  V3<S> N;
  for (int c = 0; c < 3; ++c) N.a[c] = G1xG2.a[c] / nG;
  // centroid strain and be_t UL:778-792
  double third = 1.0 / 3.0;
  double eav[6];
  for (int q = 0; q < 6; ++q) eav[q] = (third * ea[q] + third * ea[6 + q]) + third * ea[12 + q];
  M3<S> aux;
  const int vo[9] = {0, 3, 4, 3, 1, 5, 4, 5, 2};
  for (int k = 0; k < 9; ++k) aux.m[k] = S(((k % 4) == 0 ? 1.0 : 0.0) - 2.0 * eav[vo[k]]);
  M3<S> be_t = inv3(aux);
  M3<S> NN, be_t_s;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      NN.m[3 * i + j] = N.a[i] * N.a[j];
      be_t_s.m[3 * i + j] = be_t.m[3 * i + j] - NN.m[3 * i + j];
    }
  // deformed metric UL:802-806
  V3<S> g1, g2;
  for (int c = 0; c < 3; ++c) { g1.a[c] = x1.a[c] - x0.a[c]; g2.a[c] = x2.a[c] - x0.a[c]; }
  V3<S> g12 = cross(g1, g2);
  S ng = norm(g12);
  V3<S> n;
  for (int c = 0; c < 3; ++c) n.a[c] = g12.a[c] / ng;
  // G, g, inverses, F, C UL:810-832
  M3<S> G, g;
  for (int r = 0; r < 3; ++r) {
    G.m[3 * r + 0] = G1.a[r]; G.m[3 * r + 1] = G2.a[r]; G.m[3 * r + 2] = N.a[r];
    g.m[3 * r + 0] = g1.a[r]; g.m[3 * r + 1] = g2.a[r]; g.m[3 * r + 2] = n.a[r];
  }
  M3<S> Ginv = inv3(G);
  V3<S> G1i, G2i;
  for (int c = 0; c < 3; ++c) { G1i.a[c] = Ginv.m[c]; G2i.a[c] = Ginv.m[3 + c]; }
  M3<S> F;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      F.m[3 * i + j] = (g1.a[i] * G1i.a[j] + g2.a[i] * G2i.a[j]) + n.a[i] * N.a[j];
  M3<S> Ft = transpose(F);
  M3<S> C = matmul(Ft, F);
  M3<S> Cinv = inv3(C);
  M3<S> Cinv_s;
  for (int k = 0; k < 9; ++k) Cinv_s.m[k] = Cinv.m[k] - NN.m[k];
  // concentration gradients UL:835-845
  M3<S> ginv = inv3(g);
  V3<S> g1i, g2i;
  for (int c = 0; c < 3; ++c) { g1i.a[c] = ginv.m[c]; g2i.a[c] = ginv.m[3 + c]; }
  V3<S> gradA, gradC;
  for (int c = 0; c < 3; ++c) {
    S m12 = -g1i.a[c] - g2i.a[c];
    gradA.a[c] = (cA[0] * g1i.a[c] + cA[1] * g2i.a[c]) + cA[2] * m12;
    gradC.a[c] = (cC[0] * g1i.a[c] + cC[1] * g2i.a[c]) + cC[2] * m12;
  }
  double c_A = (1. / 3.) * (cA[0] + cA[1] + cA[2]);
  double c_C = (1. / 3.) * (cC[0] + cC[1] + cC[2]);
  // area changes UL:848-862
  S theta_e_t = sqrt(det3(be_t));
  double theta_g = 1.0 + thetag_dot * dt_real;
  S theta = det3(F);
  S theta_e = theta * theta_e_t / theta_g;
  S lamda_N = S(1.0) / theta_e;
  S p = -mu * (lamda_N * lamda_N);
  // stresses UL:865-880
  M3<S> Spas;
  for (int k = 0; k < 9; ++k) Spas.m[k] = (mu / theta_g) * be_t_s.m[k] + theta_g * p * Cinv_s.m[k];
  double kappa = 0.5;
  V3<S> a0;
  for (int c = 0; c < 3; ++c) a0.a[c] = (1. / 3.) * ((x0.a[c] + x1.a[c]) + x2.a[c]) - centroid[c];
  S na0 = norm(a0);
  for (int c = 0; c < 3; ++c) a0.a[c] = a0.a[c] / na0;
  M3<S> Ssum;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      S sact = t_act * ((i == j ? kappa : 0.0) + (1 - 2 * kappa) * (a0.a[i] * a0.a[j]));
      Ssum.m[3 * i + j] = sact + Spas.m[3 * i + j];
    }
  // updated strain UL:884-887
  M3<S> be = matmul(matmul(F, be_t), Ft);
  for (int k = 0; k < 9; ++k) be.m[k] = (1.0 / (theta_g * theta_g)) * be.m[k];
  M3<S> bei = inv3(be);
  S eas[6];
  {
    M3<S> es;
    for (int k = 0; k < 9; ++k) es.m[k] = 0.5 * (((k % 4) == 0 ? 1.0 : 0.0) - bei.m[k]);
    eas[0] = es.m[0]; eas[1] = es.m[4]; eas[2] = es.m[8];
    eas[3] = es.m[1]; eas[4] = es.m[2]; eas[5] = es.m[5];
  }
  // residual UL:893-927
  S Sv[6] = {Ssum.m[0], Ssum.m[4], Ssum.m[8], Ssum.m[1], Ssum.m[2], Ssum.m[5]};
  for (int i = 0; i < 3; ++i) {
    for (int pp = 0; pp < 3; ++pp) {
      double vg1[3] = {0, 0, 0}, vg2[3] = {0, 0, 0};
      if (i == 0) { vg1[pp] = -1.0; vg2[pp] = -1.0; }
      if (i == 1) vg1[pp] = 1.0;
      if (i == 2) vg2[pp] = 1.0;
      M3<S> varF;
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) varF.m[3 * r + c] = vg1[r] * G1i.a[c] + vg2[r] * G2i.a[c];
      M3<S> vC1 = matmul(transpose(varF), F), vC2 = matmul(Ft, varF);
      M3<S> varE;
      for (int k = 0; k < 9; ++k) varE.m[k] = 0.5 * (vC1.m[k] + vC2.m[k]);
      S vEv[6] = {varE.m[0], varE.m[4], varE.m[8], 2 * varE.m[1], 2 * varE.m[2], 2 * varE.m[5]};
      S dotv = Sv[0] * vEv[0];
      for (int q = 1; q < 6; ++q) dotv = dotv + Sv[q] * vEv[q];
      Re[i * 3 + pp] = dotv * nG / 2.0 * thick;
    }
    for (int q = 0; q < 6; ++q) Re_ea[i * 6 + q] = (1. / 3.) * (eas[q] - ea[6 * i + q]) * nG / 2.0 * thick;
    MMe[i] = (1. / 3.) * nG / 2.0 * thick;
    V3<S> gradNi;
    for (int c = 0; c < 3; ++c) {
      S m12 = -g1i.a[c] - g2i.a[c];
      gradNi.a[c] = (i == 0 ? g1i.a[c] : (i == 1 ? g2i.a[c] : m12));
    }
    const double D_A = 1., D_C = 1e-10;
    ReA[i] = (c_A * (1. / 3.) - dt_real * D_A * dot3(gradA, gradNi)) * nG / 2.0 * thick;
    ReC[i] = (c_C * (1. / 3.) - dt_real * D_C * dot3(gradC, gradNi)) * nG / 2.0 * thick;
  }
}

// MD table UL:1207-1214 (kMD variant, 6 knots).
const int MD_N = 6;
const double MD_PX[MD_N] = {0., 2.51250583e-03, 4.40265201e-03, 5.91408228e-03, 11.76817302e-03, 17.001e-03};
const double MD_PY[MD_N] = {-7.15514497, 71.80501308, 356.62420075, 356.96656556, 126.21811585, 528.40481124};
const double MD_K[MD_N] = {31426.8556, 99485.86235, 67080.67086, -7908.39043, 21978.66923, 76858.39733};

// Table lookup exactly as UL:1258-1264 (later matches overwrite earlier ones).
template <class S>
S table_force(int npts, const double *px, const double *py, const double *k, const S &D, bool *found) {
  S f(0.0);
  *found = false;
  for (int j = 0; j < npts - 1; j++) {
    if (val(D) < px[j + 1] && val(D) >= px[j]) {
      f = py[j] + k[j] * (D - px[j]);
      *found = true;
    } else if (val(D) >= px[npts - 1]) {
      f = py[npts - 1] + k[npts - 1] * (D - px[npts - 1]);
      *found = true;
    }
  }
  return f;
}

}  // namespace

// ==========================================================================
// Part 1: reference restatement
// ==========================================================================
extern "C" void orc_eval_element_re(const double *X, const double *x, const double *ea, const double *cA,
                                    const double *cC, double dt, double time, const double *centroid, int mat,
                                    double *Re, double *Re_ea, double *MMe, double *ReA, double *ReC) {
  double mu, t_act, thick;
  ul_constants(mat, time, &mu, &t_act, &thick);
  ul_element<double>(X, x, ea, cA, cC, dt, mu, t_act, thick, centroid, Re, Re_ea, MMe, ReA, ReC);
}

extern "C" void orc_area_tot(int n_tris, const int32_t *tris, const double *x, double *out) {
  // UL:1662-1688
  double area_tot = 0, cen[3] = {0, 0, 0};
  for (int e = 0; e < n_tris; ++e) {
    const double *p0 = x + 3 * tris[3 * e], *p1 = x + 3 * tris[3 * e + 1], *p2 = x + 3 * tris[3 * e + 2];
    V3<double> a, b;
    for (int c = 0; c < 3; ++c) { a.a[c] = p1[c] - p0[c]; b.a[c] = p2[c] - p0[c]; }
    V3<double> cr = cross(a, b);
    for (int c = 0; c < 3; ++c) cr.a[c] *= 0.5;
    double ar = norm(cr);
    for (int c = 0; c < 3; ++c) cen[c] += (1. / 3. * (p0[c] + p1[c] + p2[c])) * ar;
    area_tot += ar;
  }
  out[0] = area_tot;
  for (int c = 0; c < 3; ++c) out[1 + c] = cen[c] / area_tot;
}

extern "C" void orc_calculate_res(int n_nodes, int n_tris, const int32_t *tris, const double *X, const double *x,
                                  const double *ea, const double *cA, const double *cC, double dt, double time,
                                  const double *centroid, int mat, double *Res, double *Res_ea, double *MM,
                                  double *ResA, double *ResC) {
  // UL:632-707: zero, element loop, element-order += scatter.
  memset(Res, 0, sizeof(double) * 3 * n_nodes);
  memset(Res_ea, 0, sizeof(double) * 6 * n_nodes);
  memset(MM, 0, sizeof(double) * n_nodes);
  memset(ResA, 0, sizeof(double) * n_nodes);
  memset(ResC, 0, sizeof(double) * n_nodes);
  for (int e = 0; e < n_tris; ++e) {
    int v[3] = {tris[3 * e], tris[3 * e + 1], tris[3 * e + 2]};
    double Xe[9], xe[9], eae[18], ca[3], cc[3];
    for (int a = 0; a < 3; ++a) {
      for (int c = 0; c < 3; ++c) { Xe[3 * a + c] = X[3 * v[a] + c]; xe[3 * a + c] = x[3 * v[a] + c]; }
      for (int q = 0; q < 6; ++q) eae[6 * a + q] = ea[6 * v[a] + q];
      ca[a] = cA[v[a]];
      cc[a] = cC[v[a]];
    }
    double Re[9], Rea[18], MMe[3], RA[3], RC[3], mu, t_act, thick;
    ul_constants(mat, time, &mu, &t_act, &thick);
    ul_element<double>(Xe, xe, eae, ca, cc, dt, mu, t_act, thick, centroid, Re, Rea, MMe, RA, RC);
    for (int a = 0; a < 3; ++a) {
      for (int c = 0; c < 3; ++c) Res[3 * v[a] + c] += Re[3 * a + c];
      for (int q = 0; q < 6; ++q) Res_ea[6 * v[a] + q] += Rea[6 * a + q];
      MM[v[a]] += MMe[a];
      ResA[v[a]] += RA[a];
      ResC[v[a]] += RC[a];
    }
  }
}

namespace {
// Node one-ring lists in increasing triangle order (nondestructivetrimesh.cpp:489-506).
void one_rings(int n_nodes, int n_tris, const int32_t *tris, std::vector<int> &start, std::vector<int> &list) {
  start.assign(n_nodes + 1, 0);
  for (int e = 0; e < n_tris; ++e)
    for (int a = 0; a < 3; ++a) start[tris[3 * e + a] + 1]++;
  for (int i = 0; i < n_nodes; ++i) start[i + 1] += start[i];
  list.assign(start[n_nodes], 0);
  std::vector<int> fill(start.begin(), start.end() - 1);
  for (int e = 0; e < n_tris; ++e)
    for (int a = 0; a < 3; ++a) list[fill[tris[3 * e + a]]++] = e;
}
}  // namespace

extern "C" void orc_force_fa(int n_nodes, int n_tris, const int32_t *tris, const double *x, const double *cC,
                             const double *fau, int law, double kint, double *F) {
  // UL:1198-1303 (law 1) and UL1:1196-1295 (law 0)
  const double rho_i_max = 100;
  std::vector<int> st, li;
  one_rings(n_nodes, n_tris, tris, st, li);
  for (int ni = 0; ni < n_nodes; ++ni) {
    double Asum = 0.0;
    for (int k = st[ni]; k < st[ni + 1]; ++k) {
      int e = li[k];
      const double *p0 = x + 3 * tris[3 * e], *p1 = x + 3 * tris[3 * e + 1], *p2 = x + 3 * tris[3 * e + 2];
      V3<double> E1, E2;
      for (int c = 0; c < 3; ++c) { E1.a[c] = p1[c] - p0[c]; E2.a[c] = p2[c] - p0[c]; }
      double area = 0.5 * norm(cross(E1, E2));
      Asum += area / 3;
    }
    const double *u = fau + 3 * ni;
    double fs[3] = {0, 0, 0};
    if (law == 0) {
      for (int c = 0; c < 3; ++c) fs[c] = kint * u[c];
    } else {
      double D = ::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
      double uD[3] = {0, 0, 0};
      if (D != 0)
        for (int c = 0; c < 3; ++c) uD[c] = u[c] / D;
      bool found;
      double f = table_force<double>(MD_N, MD_PX, MD_PY, MD_K, D, &found);
      for (int c = 0; c < 3; ++c) fs[c] = f * uD[c];
    }
    for (int c = 0; c < 3; ++c) F[3 * ni + c] = -cC[ni] * fs[c] * Asum * rho_i_max;
  }
}

extern "C" void orc_update_fa(int n_nodes, int n_tris, const int32_t *tris, const double *X, const double *x,
                              double *cA, double *cC, double *fau, double dt, int n_snodes, int n_stris,
                              const int32_t *stris, const double *SX, const double *Sx, int law, double kint,
                              int *n_unlocated) {
  // UL:2026-2174 (law 1) and UL1:2005-2115 (law 0)
  (void)n_tris; (void)tris; (void)n_snodes;
  int unloc = 0;
  for (int ni = 0; ni < n_nodes; ++ni) {
    double cell_u[3], sX[3], subs_u[3] = {0, 0, 0};
    bool located = false;
    for (int c = 0; c < 3; ++c) {
      cell_u[c] = x[3 * ni + c] - X[3 * ni + c];
      sX[c] = X[3 * ni + c] - fau[3 * ni + c];
    }
    // brute-force point location, last containing triangle wins (UL:2061-2100)
    for (int e = 0; e < n_stris; ++e) {
      const int *t = stris + 3 * e;
      V3<double> e1, e2, ep;
      for (int c = 0; c < 3; ++c) {
        e1.a[c] = SX[3 * t[1] + c] - SX[3 * t[0] + c];
        e2.a[c] = SX[3 * t[2] + c] - SX[3 * t[0] + c];
        ep.a[c] = sX[c] - SX[3 * t[0] + c];
      }
      double A_tri = 0.5 * norm(cross(e1, e2));
      V3<double> A3n = cross(e1, ep);
      double A3 = -1;
      if (A3n.a[2] >= 0) A3 = 0.5 * norm(A3n);
      double A2 = -1;
      V3<double> A2n = cross(ep, e2);
      if (A2n.a[2] >= 0) A2 = 0.5 * norm(A2n);
      double A1 = A_tri - A2 - A3;
      if (A1 >= 0 && A2 >= 0 && A3 >= 0) {
        for (int c = 0; c < 3; ++c)
          subs_u[c] = ((A1 / A_tri) * (Sx[3 * t[0] + c] - SX[3 * t[0] + c]) +
                       (A2 / A_tri) * (Sx[3 * t[1] + c] - SX[3 * t[1] + c])) +
                      (A3 / A_tri) * (Sx[3 * t[2] + c] - SX[3 * t[2] + c]);
        located = true;
      }
    }
    if (!located) ++unloc;  // the reference leaves subs_u uninitialised here; defined as 0
    for (int c = 0; c < 3; ++c) fau[3 * ni + c] = (fau[3 * ni + c] + cell_u[c]) - subs_u[c];
    double force_FA;
    const double *u = fau + 3 * ni;
    if (law == 0) {
      force_FA = kint * ::sqrt(u[0] * u[0] + u[1] * u[1]);  // 2D norm, UL1:2077
    } else {
      double D = ::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
      double uD[3] = {0, 0, 0};
      if (D != 0)
        for (int c = 0; c < 3; ++c) uD[c] = u[c] / D;
      bool found;
      double f = table_force<double>(MD_N, MD_PX, MD_PY, MD_K, D, &found);
      double fs[3] = {f * uD[0], f * uD[1], f * uD[2]};
      force_FA = ::sqrt(fs[0] * fs[0] + fs[1] * fs[1] + fs[2] * fs[2]);
    }
    const double k_on = 1, ka = 0.004, kb = 10, Fa = 15, Fb = 15;  // UL:2139-2143
    double k_off = ka * wcfe::shared_exp(force_FA / Fa) + kb * wcfe::shared_exp(-force_FA / Fb);
    double c_C_old = cC[ni];
    cC[ni] = cC[ni] + dt * k_on * (1 - cC[ni]) - dt * k_off * cC[ni];
    if (cC[ni] < 0) cC[ni] = 0;
    if (cA[ni] < 0) cA[ni] = 0;
    if (c_C_old - dt * k_off * c_C_old > 0 && cC[ni] > 0) {
      for (int c = 0; c < 3; ++c) fau[3 * ni + c] = fau[3 * ni + c] * (c_C_old - dt * k_off * c_C_old) / cC[ni];
    } else {
      for (int c = 0; c < 3; ++c) fau[3 * ni + c] = 0.0 * fau[3 * ni + c];
    }
  }
  if (n_unlocated) *n_unlocated = unloc;
}

extern "C" void orc_force_fa_subs(int n_nodes, int n_tris, const int32_t *tris, const double *x_cell,
                                  const double *cC, const double *fau, int n_snodes, int n_stris,
                                  const int32_t *stris, const double *Sx, int law, double kint, double *F) {
  // UL:1305-1475 (law 1, 6-knot MD table) / UL1:1297 (law 0, linear kint):
  // the integrin force of every cell node, applied to every substrate
  // triangle containing x_cell - fa_u in the substrate's current positions
  // (no early exit: a point on a shared edge loads both triangles), split by
  // barycentric weight.
  const double rho_i_max = 100;
  for (int i = 0; i < 3 * n_snodes; ++i) F[i] = 0.0;
  std::vector<int> st, li;
  one_rings(n_nodes, n_tris, tris, st, li);
  for (int ni = 0; ni < n_nodes; ++ni) {
    double p[3];
    for (int c = 0; c < 3; ++c) p[c] = x_cell[3 * ni + c] - fau[3 * ni + c];
    for (int e = 0; e < n_stris; ++e) {
      const int *t = stris + 3 * e;
      V3<double> e1, e2, ep;
      for (int c = 0; c < 3; ++c) {
        e1.a[c] = Sx[3 * t[1] + c] - Sx[3 * t[0] + c];
        e2.a[c] = Sx[3 * t[2] + c] - Sx[3 * t[0] + c];
        ep.a[c] = p[c] - Sx[3 * t[0] + c];
      }
      const double A_tri = 0.5 * norm(cross(e1, e2));
      V3<double> A3n = cross(e1, ep);
      double A3 = -1;
      if (A3n.a[2] >= 0) A3 = 0.5 * norm(A3n);
      double A2 = -1;
      V3<double> A2n = cross(ep, e2);
      if (A2n.a[2] >= 0) A2 = 0.5 * norm(A2n);
      const double A1 = A_tri - A2 - A3;
      if (!(A1 >= 0 && A2 >= 0 && A3 >= 0)) continue;
      double Asum = 0.0;  // one-ring area / 3 of the cell node (UL:1358-1371)
      for (int k = st[ni]; k < st[ni + 1]; ++k) {
        const int *ct = tris + 3 * li[k];
        V3<double> E1, E2;
        for (int c = 0; c < 3; ++c) {
          E1.a[c] = x_cell[3 * ct[1] + c] - x_cell[3 * ct[0] + c];
          E2.a[c] = x_cell[3 * ct[2] + c] - x_cell[3 * ct[0] + c];
        }
        Asum += 0.5 * norm(cross(E1, E2)) / 3;
      }
      const double *u = fau + 3 * ni;
      double fs[3];
      if (law == 0) {
        for (int c = 0; c < 3; ++c) fs[c] = kint * u[c];
      } else {
        const double D = ::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        double uD[3] = {0, 0, 0};
        if (D != 0)
          for (int c = 0; c < 3; ++c) uD[c] = u[c] / D;
        bool found;
        const double f = table_force<double>(MD_N, MD_PX, MD_PY, MD_K, D, &found);
        for (int c = 0; c < 3; ++c) fs[c] = f * uD[c];
      }
      const double w[3] = {A1 / A_tri, A2 / A_tri, A3 / A_tri};
      for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) F[3 * t[a] + c] += cC[ni] * fs[c] * w[a] * Asum * rho_i_max;
    }
  }
}

// --------------------------------------------------------------------------
// Boundary forces (UL:1477-1568 calculateForce_AP, UL:1571-1660
// calculateBendingForce2D, UL:1691-1784 calculateAreaConstraintForce)
// --------------------------------------------------------------------------
extern "C" int orc_boundary_candidates(int n_nodes, int n_tris, const int32_t *tris, int32_t *is_bnd,
                                       int32_t *cand_start, int32_t *cand) {
  // El Topo's rule (nondestructivetrimesh.cpp:527-543): an edge with an odd
  // number of incident triangles is a boundary edge, its ends are boundary
  // vertices.  Triangles with t0 == t1 are deleted ones and skipped (:497).
  std::map<std::pair<int, int>, int> edge_count;
  for (int e = 0; e < n_tris; ++e) {
    const int32_t *t = tris + 3 * e;
    if (t[0] == t[1]) continue;
    for (int j = 0; j < 3; ++j) {
      int a = t[j], b = t[(j + 1) % 3];
      edge_count[std::make_pair(std::min(a, b), std::max(a, b))]++;
    }
  }
  for (int i = 0; i < n_nodes; ++i) is_bnd[i] = 0;
  for (const auto &kv : edge_count)
    if (kv.second % 2 == 1) is_bnd[kv.first.first] = is_bnd[kv.first.second] = 1;
  // Candidates in the order the reference's loops visit them (UL:1497-1528):
  // one-ring triangles by increasing index (the vertex-to-triangle map is
  // filled in triangle order, nondestructivetrimesh.cpp:503-506), corners in
  // slot order; nint is the first corner that is neither ni nor nj.
  std::vector<int> st, li;
  one_rings(n_nodes, n_tris, tris, st, li);
  int m = 0;
  for (int ni = 0; ni < n_nodes; ++ni) {
    cand_start[ni] = m;
    if (!is_bnd[ni]) continue;
    for (int k = st[ni]; k < st[ni + 1]; ++k) {
      const int32_t *t = tris + 3 * li[k];
      if (t[0] == t[1]) continue;
      for (int enj = 0; enj < 3; ++enj) {
        int nj = t[enj];
        if (!is_bnd[nj] || nj == ni) continue;
        int nint = -1;
        for (int enk = 0; enk < 3; ++enk)
          if (t[enk] != nj && t[enk] != ni) { nint = t[enk]; break; }
        cand[2 * m] = nj;
        cand[2 * m + 1] = nint;
        ++m;
      }
    }
  }
  cand_start[n_nodes] = m;
  return m;
}

namespace {

struct BFrame {
  V3<double> xi, x1, x2;  // the vertex, the start and the end of its boundary curve segment
  bool ok;
};

// UL:1497-1533: j1 is the neighbour with (x_i - x_j) x (x_int - x_i) . e3 > 0,
// j2 any other; later candidates overwrite earlier ones.
BFrame boundary_frame(const double *x, int ni, const int32_t *cs, const int32_t *cand) {
  BFrame f;
  bool h1 = false, h2 = false;
  for (int c = 0; c < 3; ++c) f.xi.a[c] = x[3 * ni + c];
  for (int k = cs[ni]; k < cs[ni + 1]; ++k) {
    const int nj = cand[2 * k], nint = cand[2 * k + 1];
    V3<double> a, b;
    for (int c = 0; c < 3; ++c) {
      a.a[c] = f.xi.a[c] - x[3 * nj + c];
      b.a[c] = x[3 * nint + c] - f.xi.a[c];
    }
    if (cross(a, b).a[2] > 0) {
      for (int c = 0; c < 3; ++c) f.x1.a[c] = x[3 * nj + c];
      h1 = true;
    } else {
      for (int c = 0; c < 3; ++c) f.x2.a[c] = x[3 * nj + c];
      h2 = true;
    }
  }
  f.ok = h1 && h2;
  return f;
}

// normal_i and the segment lengths of UL:1537-1545 (also :1634-1641, :1764-1771)
void boundary_normal(const BFrame &f, double n[3], double *len1, double *len2) {
  V3<double> d1, d2;
  for (int c = 0; c < 3; ++c) {
    d1.a[c] = f.xi.a[c] - f.x1.a[c];
    d2.a[c] = f.x2.a[c] - f.xi.a[c];
  }
  // -e3 x v = (v_y, -v_x, 0)
  V3<double> n1{{d1.a[1], -d1.a[0], 0.0}}, n2{{d2.a[1], -d2.a[0], 0.0}};
  const double nn1 = norm(n1), nn2 = norm(n2);
  for (int c = 0; c < 3; ++c) { n1.a[c] /= nn1; n2.a[c] /= nn2; }
  const double l1 = norm(d1), l2 = norm(d2), L = l1 + l2;
  for (int c = 0; c < 3; ++c) n[c] = l1 / L * n1.a[c] + l2 / L * n2.a[c];
  *len1 = l1;
  *len2 = l2;
}

struct BTable {
  std::vector<int32_t> is_bnd, cs, cand;
  BTable(int n_nodes, int n_tris, const int32_t *tris)
      : is_bnd(n_nodes), cs(n_nodes + 1), cand(12 * (size_t)std::max(n_tris, 1)) {
    orc_boundary_candidates(n_nodes, n_tris, tris, is_bnd.data(), cs.data(), cand.data());
  }
};

}  // namespace

extern "C" int orc_bending_force_2d(int n_nodes, int n_tris, const int32_t *tris, const double *x, double k_bend,
                                    double *F) {
  BTable bt(n_nodes, n_tris, tris);
  int open = 0;
  for (int ni = 0; ni < n_nodes; ++ni) {
    if (!bt.is_bnd[ni]) continue;
    BFrame f = boundary_frame(x, ni, bt.cs.data(), bt.cand.data());
    if (!f.ok) { ++open; for (int c = 0; c < 3; ++c) F[3 * ni + c] = 0.0; continue; }
    // UL:1609-1624 turning angle and curvature
    V3<double> e1, e2;
    for (int c = 0; c < 3; ++c) {
      e1.a[c] = f.xi.a[c] - f.x1.a[c];
      e2.a[c] = f.x2.a[c] - f.xi.a[c];
    }
    const double ne1 = norm(e1), ne2 = norm(e2);
    for (int c = 0; c < 3; ++c) { e1.a[c] /= ne1; e2.a[c] /= ne2; }
    const double costurn = e1.a[0] * e2.a[0] + e1.a[1] * e2.a[1] + e1.a[2] * e2.a[2];
    const double turna = cross(e1, e2).a[2] > 0 ? costurn - 1 : 1 - costurn;
    double n[3], l1, l2;
    boundary_normal(f, n, &l1, &l2);
    const double kappa = turna * 2 / (l1 + l2);
    for (int c = 0; c < 3; ++c) F[3 * ni + c] = k_bend * kappa * n[c];  // UL:1644-1647
  }
  return open;
}

extern "C" int orc_area_constraint_force(int n_nodes, int n_tris, const int32_t *tris, const double *x,
                                         double tot_area, double k_area, double area_0, double *F) {
  BTable bt(n_nodes, n_tris, tris);
  int open = 0;
  for (int ni = 0; ni < n_nodes; ++ni) {
    if (!bt.is_bnd[ni]) continue;
    BFrame f = boundary_frame(x, ni, bt.cs.data(), bt.cand.data());
    if (!f.ok) { ++open; for (int c = 0; c < 3; ++c) F[3 * ni + c] = 0.0; continue; }
    double n[3], l1, l2;
    boundary_normal(f, n, &l1, &l2);
    const double L = l1 + l2;
    for (int c = 0; c < 3; ++c) F[3 * ni + c] = -k_area * (tot_area - area_0) / 2 * L * n[c];  // UL:1776-1781
  }
  return open;
}

extern "C" int orc_force_ap(int n_nodes, int n_tris, const int32_t *tris, const double *x, double dt, double rate,
                            double quantum, double density, const double *u, double *F) {
  BTable bt(n_nodes, n_tris, tris);
  int open = 0;
  for (int ni = 0; ni < n_nodes; ++ni) {
    if (!bt.is_bnd[ni]) continue;
    BFrame f = boundary_frame(x, ni, bt.cs.data(), bt.cand.data());
    if (!f.ok) { ++open; for (int c = 0; c < 3; ++c) F[3 * ni + c] = 0.0; continue; }
    double n[3], l1, l2;
    boundary_normal(f, n, &l1, &l2);
    const double L = l1 + l2;
    // UL:1549-1562: one draw per vertex; every one of the density L / 2 trials
    // adds a quantum when the draw is below rate dt exp(-rate dt)
    double factin = 0;
    for (int ii = 0; ii < density * L / 2; ii++)
      if (u[ni] < rate * dt * wcfe::shared_exp(-rate * dt)) factin += quantum;
    for (int c = 0; c < 3; ++c) F[3 * ni + c] = factin * n[c];
  }
  return open;
}

// ==========================================================================
// Part 2: planar Newton model
// ==========================================================================
namespace {

double t_act_of(double time) { return (time < 2) ? time * 200.0 : 400.0; }  // UL:743-757
double cell_t_act(const orc_cell *c) { return c->passive ? 0.0 : t_act_of(c->time); }  // mat 2: UL:763

// Spring force vector for extension u (2D): (D == 0) ? k0 * u : f(D) * u / D.
template <class S> void spring_vec(const orc_cell *c, const S *u, S *out) {
  if (c->spring_law == ORC_SPRING_LINEAR) {
    out[0] = c->k_lin * u[0];
    out[1] = c->k_lin * u[1];
    return;
  }
  S D2 = u[0] * u[0] + u[1] * u[1];
  if (val(D2) == 0.0) {
    out[0] = c->tab_k[0] * u[0];
    out[1] = c->tab_k[0] * u[1];
    return;
  }
  S D = sqrt(D2);
  bool found;
  S f = table_force<S>(c->tab_n, c->tab_px, c->tab_py, c->tab_k, D, &found);
  out[0] = f * u[0] / D;
  out[1] = f * u[1] / D;
}

// Element equilibrium contribution G_e(x_e) (6 entries), generic in the scalar.
template <class S>
void element_G(const orc_cell *c, int e, const S *xe /*6*/, const std::vector<int> &sst, const std::vector<int> &sli,
               S *Ge) {
  const int32_t *t = c->tris + 3 * e;
  for (int k = 0; k < 6; ++k) Ge[k] = S(0.0);
  if (c->law == ORC_LAW_UL) {
    double Xe[9], ea6[18], zero3[3] = {0, 0, 0};
    S x9[9];
    for (int a = 0; a < 3; ++a) {
      Xe[3 * a] = c->X[2 * t[a]];
      Xe[3 * a + 1] = c->X[2 * t[a] + 1];
      Xe[3 * a + 2] = 0;
      x9[3 * a] = xe[2 * a];
      x9[3 * a + 1] = xe[2 * a + 1];
      x9[3 * a + 2] = S(0.0);
      const double *E3 = c->ea + 3 * t[a];
      double *E6 = ea6 + 6 * a;
      E6[0] = E3[0]; E6[1] = E3[1]; E6[2] = 0; E6[3] = E3[2]; E6[4] = 0; E6[5] = 0;
    }
    // far-away centroid: (1 - 2 kappa) = 0 removes a0 (UL:872-878); avoids the
    // 0/0 of an element centroid coinciding with the cell centroid.
    const double far[3] = {1e30, 1e30, 0};
    S Re[9], Rea[18], MMe[3], RA[3], RC[3];
    ul_element<S>(Xe, x9, ea6, zero3, zero3, 0.0, c->mu, cell_t_act(c), c->thick, far, Re, Rea, MMe, RA, RC);
    for (int a = 0; a < 3; ++a) {
      Ge[2 * a] = Re[3 * a];
      Ge[2 * a + 1] = Re[3 * a + 1];
    }
  } else {
    // Linear plane-stress CST (small strain), active isotropic stress 0.5 t_act.
    double X0 = c->X[2 * t[0]], Y0 = c->X[2 * t[0] + 1];
    double X1 = c->X[2 * t[1]], Y1 = c->X[2 * t[1] + 1];
    double X2 = c->X[2 * t[2]], Y2 = c->X[2 * t[2] + 1];
    double det = (X1 - X0) * (Y2 - Y0) - (X2 - X0) * (Y1 - Y0);
    double A = 0.5 * fabs(det);
    // shape-function gradients
    double b[3] = {(Y1 - Y2) / det, (Y2 - Y0) / det, (Y0 - Y1) / det};
    double d[3] = {(X2 - X1) / det, (X0 - X2) / det, (X1 - X0) / det};
    S ux[3], uy[3];
    for (int a = 0; a < 3; ++a) {
      ux[a] = xe[2 * a] - c->X[2 * t[a]];
      uy[a] = xe[2 * a + 1] - c->X[2 * t[a] + 1];
    }
    S exx = b[0] * ux[0] + b[1] * ux[1] + b[2] * ux[2];
    S eyy = d[0] * uy[0] + d[1] * uy[1] + d[2] * uy[2];
    S gxy = (d[0] * ux[0] + d[1] * ux[1] + d[2] * ux[2]) + (b[0] * uy[0] + b[1] * uy[1] + b[2] * uy[2]);
    double Dc = c->E / (1.0 - c->nu * c->nu);
    double sa = 0.5 * cell_t_act(c);
    S sxx = Dc * (exx + c->nu * eyy) + sa;
    S syy = Dc * (c->nu * exx + eyy) + sa;
    S sxy = Dc * 0.5 * (1.0 - c->nu) * gxy;
    double At = A * c->thick;
    for (int a = 0; a < 3; ++a) {
      Ge[2 * a] = At * (b[a] * sxx + d[a] * sxy);
      Ge[2 * a + 1] = At * (d[a] * syy + b[a] * sxy);
    }
  }
  // barycentric springs hosted by this element: G += b_a * spring_vec(p(x) - anchor) * state
  for (int k = sst[e]; k < sst[e + 1]; ++k) {
    int s = sli[k];
    double st = c->spr_state[s];
    if (st == 0.0) continue;
    const double *bb = c->spr_bary + 3 * s;
    S u[2];
    for (int q = 0; q < 2; ++q)
      u[q] = (bb[0] * xe[q] + bb[1] * xe[2 + q] + bb[2] * xe[4 + q]) - c->spr_anchor[2 * s + q];
    S f[2];
    spring_vec<S>(c, u, f);
    for (int a = 0; a < 3; ++a)
      for (int q = 0; q < 2; ++q) Ge[2 * a + q] = Ge[2 * a + q] + (bb[a] * st) * f[q];
  }
  // node-attached springs (UL:1226-1269): G_a += (area_e / 3) * cC rho f(u_a)
  if (c->node_springs) {
    S cz = (xe[2] - xe[0]) * (xe[5] - xe[1]) - (xe[3] - xe[1]) * (xe[4] - xe[0]);
    S area = 0.5 * sqrt(cz * cz);
    for (int a = 0; a < 3; ++a) {
      int v = t[a];
      S u[2];
      for (int q = 0; q < 2; ++q) u[q] = c->fau0[2 * v + q] + (xe[2 * a + q] - c->X[2 * v + q]);
      S f[2];
      spring_vec<S>(c, u, f);
      for (int q = 0; q < 2; ++q) Ge[2 * a + q] = Ge[2 * a + q] + (area / 3.0) * (c->cC[v] * c->rho_max) * f[q];
    }
  }
}

void spring_lists(const orc_cell *c, std::vector<int> &st, std::vector<int> &li) {
  st.assign(c->n_tris + 1, 0);
  for (int s = 0; s < c->n_spr; ++s) st[c->spr_tri[s] + 1]++;
  for (int e = 0; e < c->n_tris; ++e) st[e + 1] += st[e];
  li.assign(c->n_spr, 0);
  std::vector<int> fill(st.begin(), st.end() - 1);
  for (int s = 0; s < c->n_spr; ++s) li[fill[c->spr_tri[s]]++] = s;
}

void element_system_impl(const orc_cell *c, int e, const std::vector<int> &sst, const std::vector<int> &sli,
                         double *Ge, double *Ke) {
  const int32_t *t = c->tris + 3 * e;
  D6 xe[6];
  for (int a = 0; a < 3; ++a)
    for (int q = 0; q < 2; ++q) {
      xe[2 * a + q] = D6(c->x[2 * t[a] + q]);
      xe[2 * a + q].d[2 * a + q] = 1.0;
    }
  D6 G[6];
  element_G<D6>(c, e, xe, sst, sli, G);
  for (int r = 0; r < 6; ++r) {
    Ge[r] = G[r].v;
    if (Ke)
      for (int k = 0; k < 6; ++k) Ke[6 * r + k] = G[r].d[k];
  }
}

}  // namespace

extern "C" void orc_spring_law(const orc_cell *c, double D, double *f, double *df) {
  if (c->spring_law == ORC_SPRING_LINEAR) {
    *f = c->k_lin * D;
    *df = c->k_lin;
    return;
  }
  Dual<1> d(D);
  d.d[0] = 1.0;
  bool found;
  Dual<1> r = table_force<Dual<1>>(c->tab_n, c->tab_px, c->tab_py, c->tab_k, d, &found);
  *f = r.v;
  *df = r.d[0];
}

extern "C" int orc_color_elements(int n_nodes, int n_tris, const int32_t *tris, int32_t *color) {
  std::vector<uint64_t> used(n_nodes, 0);
  int nc = 0;
  for (int e = 0; e < n_tris; ++e) {
    uint64_t forb = used[tris[3 * e]] | used[tris[3 * e + 1]] | used[tris[3 * e + 2]];
    int col = 0;
    while (col < 64 && ((forb >> col) & 1ull)) ++col;
    if (col == 64) return -1;
    color[e] = col;
    for (int a = 0; a < 3; ++a) used[tris[3 * e + a]] |= (1ull << col);
    nc = std::max(nc, col + 1);
  }
  return nc;
}

extern "C" void orc_element_system(const orc_cell *c, int e, const int32_t *spr_start, double *Ge, double *Ke) {
  (void)spr_start;
  std::vector<int> sst, sli;
  spring_lists(c, sst, sli);
  element_system_impl(c, e, sst, sli, Ge, Ke);
}

namespace {

// j1 / j2 of boundary vertex ni by the rule of boundary_frame (indices)
bool frame_indices(const double *x3, int ni, const int32_t *cs, const int32_t *cand, int *j1, int *j2) {
  bool h1 = false, h2 = false;
  for (int k = cs[ni]; k < cs[ni + 1]; ++k) {
    const int nj = cand[2 * k], nint = cand[2 * k + 1];
    V3<double> a, b;
    for (int q = 0; q < 3; ++q) {
      a.a[q] = x3[3 * ni + q] - x3[3 * nj + q];
      b.a[q] = x3[3 * nint + q] - x3[3 * ni + q];
    }
    if (cross(a, b).a[2] > 0) { *j1 = nj; h1 = true; }
    else { *j2 = nj; h2 = true; }
  }
  return h1 && h2;
}

std::vector<double> planar3(const double *x2, int n) {
  std::vector<double> x3(3 * (size_t)n, 0.0);
  for (int i = 0; i < n; ++i) { x3[3 * i] = x2[2 * i]; x3[3 * i + 1] = x2[2 * i + 1]; }
  return x3;
}

// Node terms of the implicit residual and their tangent (rows of node i only,
// plus the area term's rank-one coupling of boundary vertices).
void add_node_terms(const orc_cell *c, double *G, double *K) {
  const int N = c->n_nodes, n = 2 * N;
  const bool bnd = c->k_bend != 0 || c->k_area != 0 || c->ap_u;
  if (c->f_ext)
    for (int i = 0; i < n; ++i) G[i] -= c->f_ext[i];
  if (c->drag != 0) {
    const double cd = c->drag / c->dt;
    for (int i = 0; i < n; ++i) {
      G[i] += cd * (c->x[i] - c->X[i]);
      if (K) K[(size_t)i * n + i] += cd;
    }
  }
  if (!bnd) return;
  std::vector<double> x3 = planar3(c->x, N), X3 = planar3(c->X, N), F(3 * (size_t)N, 0.0);
  double A = 0;
  if (c->k_bend != 0) {
    orc_bending_force_2d(N, c->n_tris, c->tris, x3.data(), c->k_bend, F.data());
    for (int i = 0; i < N; ++i) for (int q = 0; q < 2; ++q) G[2 * i + q] -= F[3 * i + q];
  }
  if (c->k_area != 0) {
    double a4[4];
    orc_area_tot(c->n_tris, c->tris, x3.data(), a4);
    A = a4[0];
    std::fill(F.begin(), F.end(), 0.0);
    orc_area_constraint_force(N, c->n_tris, c->tris, x3.data(), A, c->k_area, c->area_0, F.data());
    for (int i = 0; i < N; ++i) for (int q = 0; q < 2; ++q) G[2 * i + q] -= F[3 * i + q];
  }
  if (c->ap_u) {
    std::fill(F.begin(), F.end(), 0.0);
    orc_force_ap(N, c->n_tris, c->tris, X3.data(), c->dt, c->ap_rate, c->ap_quantum, c->ap_density, c->ap_u,
                 F.data());
    for (int i = 0; i < N; ++i) for (int q = 0; q < 2; ++q) G[2 * i + q] -= F[3 * i + q];
  }
  if (!K || (c->k_bend == 0 && c->k_area == 0)) return;
  BTable bt(N, c->n_tris, c->tris);
  std::vector<int> J1(N, -1), J2(N, -1);
  std::vector<double> g(2 * (size_t)N, 0.0);  // dA/dx at boundary vertices
  for (int i = 0; i < N; ++i) {
    if (!bt.is_bnd[i]) continue;
    int j1 = -1, j2 = -1;
    if (!frame_indices(x3.data(), i, bt.cs.data(), bt.cand.data(), &j1, &j2)) continue;
    J1[i] = j1;
    J2[i] = j2;
    g[2 * i] = 0.5 * (c->x[2 * j2 + 1] - c->x[2 * j1 + 1]);
    g[2 * i + 1] = -0.5 * (c->x[2 * j2] - c->x[2 * j1]);
  }
  // -dF/dx added to K; R = [[0, 1], [-1, 0]]
  const double R[2][2] = {{0, 1}, {-1, 0}};
  for (int i = 0; i < N; ++i) {
    if (J1[i] < 0) continue;
    const int j1 = J1[i], j2 = J2[i];
    const int nodes[3] = {i, j1, j2};
    const double dsm[3] = {0, -1, 1};
    double Rs[2] = {c->x[2 * j2 + 1] - c->x[2 * j1 + 1], -(c->x[2 * j2] - c->x[2 * j1])};
    double dF[3][2][2] = {};
    if (c->k_bend != 0) {
      const double u[2] = {c->x[2 * i] - c->x[2 * j1], c->x[2 * i + 1] - c->x[2 * j1 + 1]};
      const double w[2] = {c->x[2 * j2] - c->x[2 * i], c->x[2 * j2 + 1] - c->x[2 * i + 1]};
      const double lu = ::sqrt(u[0] * u[0] + u[1] * u[1]), lw = ::sqrt(w[0] * w[0] + w[1] * w[1]);
      const double e1[2] = {u[0] / lu, u[1] / lu}, e2[2] = {w[0] / lw, w[1] / lw};
      const double cth = e1[0] * e2[0] + e1[1] * e2[1];
      const double sg = (e1[0] * e2[1] - e1[1] * e2[0]) > 0 ? 1.0 : -1.0;
      const double t = sg * (cth - 1), L = lu + lw;
      // d cos / du, d cos / dw
      const double cu[2] = {(e2[0] - cth * e1[0]) / lu, (e2[1] - cth * e1[1]) / lu};
      const double cw[2] = {(e1[0] - cth * e2[0]) / lw, (e1[1] - cth * e2[1]) / lw};
      const double dc[3][2] = {{cu[0] - cw[0], cu[1] - cw[1]}, {-cu[0], -cu[1]}, {cw[0], cw[1]}};
      const double dL[3][2] = {{e1[0] - e2[0], e1[1] - e2[1]}, {-e1[0], -e1[1]}, {e2[0], e2[1]}};
      const double k2 = 2 * c->k_bend;
      for (int m = 0; m < 3; ++m)
        for (int r = 0; r < 2; ++r)
          for (int q = 0; q < 2; ++q)
            dF[m][r][q] += k2 * (Rs[r] * (sg * dc[m][q] / (L * L) - 2 * t * dL[m][q] / (L * L * L)) +
                                 t / (L * L) * dsm[m] * R[r][q]);
    }
    if (c->k_area != 0) {
      const double h = -0.5 * c->k_area * (A - c->area_0);
      for (int m = 0; m < 3; ++m)
        for (int r = 0; r < 2; ++r)
          for (int q = 0; q < 2; ++q) dF[m][r][q] += h * dsm[m] * R[r][q];
      // rank one: dF_i/dx_m = -k g_i g_m^T over boundary vertices m
      for (int m = 0; m < N; ++m) {
        if (J1[m] < 0) continue;
        for (int r = 0; r < 2; ++r)
          for (int q = 0; q < 2; ++q)
            K[(size_t)(2 * i + r) * n + 2 * m + q] += c->k_area * g[2 * i + r] * g[2 * m + q];
      }
    }
    for (int m = 0; m < 3; ++m)
      for (int r = 0; r < 2; ++r)
        for (int q = 0; q < 2; ++q) K[(size_t)(2 * i + r) * n + 2 * nodes[m] + q] -= dF[m][r][q];
  }
}

}  // namespace

extern "C" void orc_assemble(const orc_cell *c, const int32_t *color, int n_colors, double *G, double *K) {
  const int n = 2 * c->n_nodes;
  memset(G, 0, sizeof(double) * n);
  if (K) memset(K, 0, sizeof(double) * (size_t)n * n);
  std::vector<int> sst, sli;
  spring_lists(c, sst, sli);
  int passes = color ? n_colors : 1;
  for (int pc = 0; pc < passes; ++pc) {
    for (int e = 0; e < c->n_tris; ++e) {
      if (color && color[e] != pc) continue;
      double Ge[6], Ke[36];
      element_system_impl(c, e, sst, sli, Ge, K ? Ke : nullptr);
      const int32_t *t = c->tris + 3 * e;
      for (int r = 0; r < 6; ++r) {
        int gr = 2 * t[r / 2] + (r % 2);
        G[gr] += Ge[r];
        if (K)
          for (int k = 0; k < 6; ++k) K[(size_t)gr * n + 2 * t[k / 2] + (k % 2)] += Ke[6 * r + k];
      }
    }
  }
  add_node_terms(c, G, K);
}

template <class T> static int lu_factor_t(int n, T *A, int32_t *ipiv, int pivoting) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    if (pivoting) {
      T best = fabs(A[(size_t)k * n + k]);
      for (int i = k + 1; i < n; ++i) {
        T v = fabs(A[(size_t)i * n + k]);
        if (v > best) { best = v; p = i; }
      }
      ipiv[k] = p;
      if (p != k)
        for (int j = 0; j < n; ++j) std::swap(A[(size_t)k * n + j], A[(size_t)p * n + j]);
    } else if (ipiv) {
      ipiv[k] = k;
    }
    T piv = A[(size_t)k * n + k];
    if (piv == 0) return k + 1;
    for (int i = k + 1; i < n; ++i) {
      T l = A[(size_t)i * n + k] / piv;
      A[(size_t)i * n + k] = l;
      if (l != 0)
        for (int j = k + 1; j < n; ++j) A[(size_t)i * n + j] -= l * A[(size_t)k * n + j];
    }
  }
  return 0;
}

extern "C" int orc_lu_factor(int n, double *A, int32_t *ipiv, int pivoting) { return lu_factor_t<double>(n, A, ipiv, pivoting); }
extern "C" int orc_lu_factor_f32(int n, float *A, int32_t *ipiv, int pivoting) { return lu_factor_t<float>(n, A, ipiv, pivoting); }

extern "C" void orc_lu_solve(int n, const double *LU, const int32_t *ipiv, int pivoting, double *b) {
  if (pivoting)
    for (int k = 0; k < n; ++k)
      if (ipiv[k] != k) std::swap(b[k], b[ipiv[k]]);
  for (int i = 0; i < n; ++i) {
    double s = b[i];
    for (int j = 0; j < i; ++j) s -= LU[(size_t)i * n + j] * b[j];
    b[i] = s;
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = b[i];
    for (int j = i + 1; j < n; ++j) s -= LU[(size_t)i * n + j] * b[j];
    b[i] = s / LU[(size_t)i * n + i];
  }
}

extern "C" int orc_pcg_csr(int n, const int32_t *rowptr, const int32_t *col, const double *val_, const double *b,
                           double *x, double rtol, int maxit) {
  std::vector<double> r(n), z(n), p(n), q(n), dinv(n);
  for (int i = 0; i < n; ++i) {
    double d = 0;
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k)
      if (col[k] == i) d = val_[k];
    dinv[i] = d != 0 ? 1.0 / d : 1.0;
  }
  auto spmv = [&](const double *v, double *out) {
    for (int i = 0; i < n; ++i) {
      double s = 0;
      for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) s += val_[k] * v[col[k]];
      out[i] = s;
    }
  };
  spmv(x, q.data());
  double bn = 0;
  for (int i = 0; i < n; ++i) { r[i] = b[i] - q[i]; bn += b[i] * b[i]; }
  bn = ::sqrt(bn);
  if (bn == 0) bn = 1;
  double rz = 0;
  for (int i = 0; i < n; ++i) { z[i] = dinv[i] * r[i]; p[i] = z[i]; rz += r[i] * z[i]; }
  for (int it = 0; it < maxit; ++it) {
    double rn = 0;
    for (int i = 0; i < n; ++i) rn += r[i] * r[i];
    if (::sqrt(rn) <= rtol * bn) return it;
    spmv(p.data(), q.data());
    double pq = 0;
    for (int i = 0; i < n; ++i) pq += p[i] * q[i];
    double alpha = rz / pq;
    for (int i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
    double rz2 = 0;
    for (int i = 0; i < n; ++i) { z[i] = dinv[i] * r[i]; rz2 += r[i] * z[i]; }
    double beta = rz2 / rz;
    rz = rz2;
    for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
  }
  return -1;
}

extern "C" int orc_newton_solve(orc_cell *c, double *x, double rtol, double atol, int maxit, double *res_hist) {
  const int n = 2 * c->n_nodes;
  std::vector<double> G(n), K((size_t)n * n);
  std::vector<int32_t> ipiv(n);
  std::vector<int32_t> color(c->n_tris);
  int nc = orc_color_elements(c->n_nodes, c->n_tris, c->tris, color.data());
  const double *xsave = c->x;
  c->x = x;
  double r0 = -1;
  int ret = -1;
  for (int it = 0; it <= maxit; ++it) {
    orc_assemble(c, color.data(), nc, G.data(), K.data());
    double rn = 0;
    for (int i = 0; i < n; ++i) rn += G[i] * G[i];
    rn = ::sqrt(rn);
    if (res_hist) res_hist[it] = rn;
    if (r0 < 0) r0 = rn;
    if (rn <= rtol * r0 || rn <= atol) { ret = it; break; }
    if (it == maxit) break;
    if (orc_lu_factor(n, K.data(), ipiv.data(), 1)) break;
    for (int i = 0; i < n; ++i) G[i] = -G[i];
    orc_lu_solve(n, K.data(), ipiv.data(), 1, G.data());
    for (int i = 0; i < n; ++i) x[i] += G[i];
  }
  c->x = xsave;
  return ret;
}

// Philox4x32-10, Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11: multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key increments
// 0x9E3779B9 / 0xBB67AE85, 10 rounds.
extern "C" void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c[4] = {ctr_in[0], ctr_in[1], ctr_in[2], ctr_in[3]};
  uint32_t k[2] = {key_in[0], key_in[1]};
  for (int r = 0; r < 10; ++r) {
    if (r > 0) { k[0] += 0x9E3779B9u; k[1] += 0xBB67AE85u; }
    uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c[1] ^ k[0];
    uint32_t n2 = hi0 ^ c[3] ^ k[1];
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
  }
  for (int i = 0; i < 4; ++i) out[i] = c[i];
}

extern "C" void orc_exp_array(int n, const double *x, double *y) {
  for (int i = 0; i < n; ++i) y[i] = wcfe::shared_exp(x[i]);
}

extern "C" double orc_uniform(uint32_t key_lo, uint32_t key_hi, uint32_t a, uint32_t b, uint32_t cc) {
  uint32_t ctr[4] = {a, b, cc, 0}, key[2] = {key_lo, key_hi}, o[4];
  orc_philox4x32_10(ctr, key, o);
  return ((double)o[0] + 0.5) * (1.0 / 4294967296.0);
}

extern "C" int orc_bell_update(const orc_cell *c, double *state, double *anchor, double dt, double k_on, double k_off0,
                               double f_b, uint32_t key_lo, uint32_t key_hi, uint32_t replica, uint32_t step,
                               double *min_margin) {
  int transitions = 0;
  double mm = 1.0;
  for (int s = 0; s < c->n_spr; ++s) {
    const int32_t *t = c->tris + 3 * c->spr_tri[s];
    const double *bb = c->spr_bary + 3 * s;
    double p[2];
    for (int q = 0; q < 2; ++q)
      p[q] = bb[0] * c->x[2 * t[0] + q] + bb[1] * c->x[2 * t[1] + q] + bb[2] * c->x[2 * t[2] + q];
    double u = orc_uniform(key_lo, key_hi, replica, (uint32_t)s, step);
    double prob;
    if (state[s] != 0.0) {
      double d0 = p[0] - anchor[2 * s], d1 = p[1] - anchor[2 * s + 1];
      double D = ::sqrt(d0 * d0 + d1 * d1), f, df;
      orc_spring_law(c, D, &f, &df);
      double koff = k_off0 * wcfe::shared_exp(fabs(f) / f_b);
      prob = 1.0 - wcfe::shared_exp(-koff * dt);
      if (u < prob) { state[s] = 0.0; ++transitions; }
    } else {
      prob = 1.0 - wcfe::shared_exp(-k_on * dt);
      if (u < prob) {
        state[s] = 1.0;
        anchor[2 * s] = p[0];
        anchor[2 * s + 1] = p[1];
        ++transitions;
      }
    }
    mm = std::min(mm, fabs(u - prob));
  }
  if (min_margin) *min_margin = mm;
  return transitions;
}

extern "C" void orc_commit_ul_strain(const orc_cell *c, double *ea_out) {
  // UL:607-612: ea += Res_ea / MM, Res_ea/MM assembled in element order (UL:692-705).
  const int N = c->n_nodes;
  std::vector<double> Rea(6 * N, 0.0), MM(N, 0.0);
  const double far[3] = {1e30, 1e30, 0}, zero3[3] = {0, 0, 0};
  for (int e = 0; e < c->n_tris; ++e) {
    const int32_t *t = c->tris + 3 * e;
    double Xe[9], xe[9], ea6[18];
    for (int a = 0; a < 3; ++a) {
      Xe[3 * a] = c->X[2 * t[a]]; Xe[3 * a + 1] = c->X[2 * t[a] + 1]; Xe[3 * a + 2] = 0;
      xe[3 * a] = c->x[2 * t[a]]; xe[3 * a + 1] = c->x[2 * t[a] + 1]; xe[3 * a + 2] = 0;
      const double *E3 = c->ea + 3 * t[a];
      double *E6 = ea6 + 6 * a;
      E6[0] = E3[0]; E6[1] = E3[1]; E6[2] = 0; E6[3] = E3[2]; E6[4] = 0; E6[5] = 0;
    }
    double Re[9], Re_ea[18], MMe[3], RA[3], RC[3];
    ul_element<double>(Xe, xe, ea6, zero3, zero3, 0.0, c->mu, cell_t_act(c), c->thick, far, Re, Re_ea, MMe, RA,
                       RC);
    for (int a = 0; a < 3; ++a) {
      for (int q = 0; q < 6; ++q) Rea[6 * t[a] + q] += Re_ea[6 * a + q];
      MM[t[a]] += MMe[a];
    }
  }
  for (int i = 0; i < N; ++i) {
    ea_out[3 * i + 0] = c->ea[3 * i + 0] + Rea[6 * i + 0] / MM[i];
    ea_out[3 * i + 1] = c->ea[3 * i + 1] + Rea[6 * i + 1] / MM[i];
    ea_out[3 * i + 2] = c->ea[3 * i + 2] + Rea[6 * i + 3] / MM[i];
  }
}

// ==========================================================================
// Part 2, sparse forms: CSR assembly, Newton with Jacobi-PCG, and one Newton
// step with a reverse Cuthill-McKee band LU (the CPU baseline's solver: the
// same numbering idea and factorization the GPU's band LU uses, in FP64)
// ==========================================================================
namespace {

std::vector<std::vector<int>> node_neighbours(int n_nodes, int n_tris, const int32_t *tris) {
  std::vector<std::vector<int>> nb(n_nodes);
  for (int i = 0; i < n_nodes; ++i) nb[i].push_back(i);
  for (int e = 0; e < n_tris; ++e)
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) nb[tris[3 * e + a]].push_back(tris[3 * e + b]);
  for (auto &v : nb) {
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
  }
  return nb;
}

bool csr_supported(const orc_cell *c) { return c->k_bend == 0 && c->k_area == 0; }

void csr_add(const int32_t *rowptr, const int32_t *col, double *val, int r, int cidx, double v) {
  const int32_t *b = col + rowptr[r], *e = col + rowptr[r + 1];
  const int32_t *p = std::lower_bound(b, e, cidx);
  val[p - col] += v;
}

}  // namespace

extern "C" int orc_csr_pattern(int n_nodes, int n_tris, const int32_t *tris, int32_t *rowptr, int32_t *col, int cap) {
  std::vector<std::vector<int>> nb = node_neighbours(n_nodes, n_tris, tris);
  long long nnz = 0;
  for (auto &v : nb) nnz += 4LL * (long long)v.size();
  if (nnz > 0x7fffffff) return -1;
  if (!rowptr || !col || cap < nnz) return (int)nnz;
  int k = 0;
  rowptr[0] = 0;
  for (int i = 0; i < n_nodes; ++i) {
    for (int q = 0; q < 2; ++q) {
      for (int j : nb[i]) {
        col[k++] = 2 * j;
        col[k++] = 2 * j + 1;
      }
      rowptr[2 * i + q + 1] = k;
    }
  }
  return (int)nnz;
}

extern "C" int orc_assemble_csr(const orc_cell *c, const int32_t *rowptr, const int32_t *col, double *val, double *G) {
  if (!csr_supported(c)) return -1;
  const int n = 2 * c->n_nodes;
  memset(G, 0, sizeof(double) * n);
  memset(val, 0, sizeof(double) * rowptr[n]);
  std::vector<int> sst, sli;
  spring_lists(c, sst, sli);
  for (int e = 0; e < c->n_tris; ++e) {
    double Ge[6], Ke[36];
    element_system_impl(c, e, sst, sli, Ge, Ke);
    const int32_t *t = c->tris + 3 * e;
    for (int r = 0; r < 6; ++r) {
      const int gr = 2 * t[r / 2] + (r % 2);
      G[gr] += Ge[r];
      for (int k = 0; k < 6; ++k) csr_add(rowptr, col, val, gr, 2 * t[k / 2] + (k % 2), Ke[6 * r + k]);
    }
  }
  add_node_terms(c, G, nullptr);  // G parts (external load, drag, actin polymerisation)
  if (c->drag != 0) {
    const double cd = c->drag / c->dt;
    for (int i = 0; i < n; ++i) csr_add(rowptr, col, val, i, i, cd);
  }
  return 0;
}

extern "C" int orc_newton_solve_pcg(orc_cell *c, double *x, double rtol, double atol, int maxit, double cg_rtol,
                                    int cg_maxit, double *res_hist) {
  if (!csr_supported(c)) return -2;
  const int n = 2 * c->n_nodes;
  const int nnz = orc_csr_pattern(c->n_nodes, c->n_tris, c->tris, nullptr, nullptr, 0);
  if (nnz < 0) return -2;
  std::vector<int32_t> rowptr(n + 1), col(nnz);
  orc_csr_pattern(c->n_nodes, c->n_tris, c->tris, rowptr.data(), col.data(), nnz);
  std::vector<double> val(nnz), G(n), dx(n);
  const double *xsave = c->x;
  c->x = x;
  double r0 = -1;
  int ret = -1;
  for (int it = 0; it <= maxit; ++it) {
    orc_assemble_csr(c, rowptr.data(), col.data(), val.data(), G.data());
    double rn = 0;
    for (int i = 0; i < n; ++i) rn += G[i] * G[i];
    rn = ::sqrt(rn);
    if (r0 < 0) r0 = rn;
    if (res_hist) res_hist[it] = rn;
    if (rn <= rtol * r0 || rn <= atol) { ret = it; break; }
    if (it == maxit) break;
    for (int i = 0; i < n; ++i) { G[i] = -G[i]; dx[i] = 0; }
    if (orc_pcg_csr(n, rowptr.data(), col.data(), val.data(), G.data(), dx.data(), cg_rtol, cg_maxit) < 0) break;
    for (int i = 0; i < n; ++i) x[i] += dx[i];
  }
  c->x = xsave;
  return ret;
}

namespace {

// Reverse Cuthill-McKee node order (George and Liu, "Computer Solution of
// Large Sparse Positive Definite Systems", 1981): each connected component
// starts at a pseudo-peripheral node (repeated breadth-first searches from a
// node of least degree in the last level), is numbered level by level with
// unnumbered neighbours in increasing degree, and the order is reversed.
std::vector<int> rcm_order(const std::vector<std::vector<int>> &nb) {
  const int N = (int)nb.size();
  std::vector<int> deg(N), order, level(N, -1);
  for (int i = 0; i < N; ++i) deg[i] = (int)nb[i].size() - 1;
  std::vector<char> done(N, 0);
  order.reserve(N);
  auto bfs_last_level = [&](int s, int &depth) {
    std::vector<int> cur{s}, nxt, seen;
    level[s] = 0;
    seen.push_back(s);
    depth = 0;
    std::vector<int> last = cur;
    while (!cur.empty()) {
      nxt.clear();
      for (int v : cur)
        for (int w : nb[v])
          if (level[w] < 0 && !done[w]) { level[w] = depth + 1; nxt.push_back(w); seen.push_back(w); }
      if (nxt.empty()) break;
      last = nxt;
      cur.swap(nxt);
      ++depth;
    }
    for (int v : seen) level[v] = -1;
    return last;
  };
  for (int s0 = 0; s0 < N; ++s0) {
    if (done[s0]) continue;
    // least-degree node of this component
    int s = s0;
    {
      std::vector<int> stack{s0}, comp;
      std::vector<char> mark(N, 0);
      mark[s0] = 1;
      while (!stack.empty()) {
        int v = stack.back();
        stack.pop_back();
        comp.push_back(v);
        for (int w : nb[v])
          if (!mark[w] && !done[w]) { mark[w] = 1; stack.push_back(w); }
      }
      for (int v : comp)
        if (deg[v] < deg[s] || (deg[v] == deg[s] && v < s)) s = v;
    }
    int depth = 0;
    for (int iter = 0; iter < 8; ++iter) {
      std::vector<int> last = bfs_last_level(s, depth);
      int best = last[0];
      for (int v : last)
        if (deg[v] < deg[best] || (deg[v] == deg[best] && v < best)) best = v;
      int d2 = 0;
      bfs_last_level(best, d2);
      if (d2 <= depth) break;
      s = best;
    }
    // Cuthill-McKee from s
    const size_t first = order.size();
    order.push_back(s);
    done[s] = 1;
    for (size_t h = first; h < order.size(); ++h) {
      const int v = order[h];
      std::vector<int> nx;
      for (int w : nb[v])
        if (!done[w]) { done[w] = 1; nx.push_back(w); }
      std::sort(nx.begin(), nx.end(), [&](int a, int b) { return deg[a] != deg[b] ? deg[a] < deg[b] : a < b; });
      order.insert(order.end(), nx.begin(), nx.end());
    }
  }
  std::reverse(order.begin(), order.end());
  return order;
}

// One cell's band system in RCM numbering (2 unknowns per node): row r of
// the permuted matrix holds columns r - w .. r + w at band[r (2w+1) + (c - r + w)].
struct BandPlan {
  int n = 0, w = 0;
  std::vector<int> pos;  // node -> position in the RCM order
  std::vector<int32_t> rowptr, col;
  std::vector<double> val, G, band, rhs;
};

BandPlan band_plan(const orc_cell *c) {
  BandPlan p;
  const int N = c->n_nodes;
  p.n = 2 * N;
  std::vector<std::vector<int>> nb = node_neighbours(N, c->n_tris, c->tris);
  std::vector<int> ord = rcm_order(nb);
  p.pos.assign(N, 0);
  for (int k = 0; k < N; ++k) p.pos[ord[k]] = k;
  int wn = 0;
  for (int i = 0; i < N; ++i)
    for (int j : nb[i]) wn = std::max(wn, std::abs(p.pos[i] - p.pos[j]));
  p.w = 2 * wn + 1;
  const int nnz = orc_csr_pattern(N, c->n_tris, c->tris, nullptr, nullptr, 0);
  p.rowptr.resize(p.n + 1);
  p.col.resize(nnz);
  orc_csr_pattern(N, c->n_tris, c->tris, p.rowptr.data(), p.col.data(), nnz);
  p.val.resize(nnz);
  p.G.resize(p.n);
  p.band.resize((size_t)p.n * (2 * p.w + 1));
  p.rhs.resize(p.n);
  return p;
}

// Newton iteration: assemble, band LU without pivoting (FP64), solve, update.
int band_newton_step(orc_cell *c, BandPlan &p, double *x) {
  const double *xsave = c->x;
  c->x = x;
  int rc = orc_assemble_csr(c, p.rowptr.data(), p.col.data(), p.val.data(), p.G.data());
  c->x = xsave;
  if (rc) return rc;
  const int n = p.n, w = p.w, W = 2 * w + 1;
  std::fill(p.band.begin(), p.band.end(), 0.0);
  auto P = [&](int r) { return 2 * p.pos[r / 2] + (r % 2); };
  for (int r = 0; r < n; ++r) {
    const int pr = P(r);
    for (int k = p.rowptr[r]; k < p.rowptr[r + 1]; ++k) {
      const int pc = P(p.col[k]);
      p.band[(size_t)pr * W + (pc - pr + w)] += p.val[k];
    }
    p.rhs[pr] = -p.G[r];
  }
  double *B = p.band.data();
  for (int k = 0; k < n; ++k) {
    const double piv = B[(size_t)k * W + w];
    if (piv == 0.0) return k + 1;
    const int iend = std::min(n - 1, k + w);
    for (int i = k + 1; i <= iend; ++i) {
      double *Ri = B + (size_t)i * W;
      const double l = Ri[k - i + w] / piv;
      Ri[k - i + w] = l;
      if (l == 0.0) continue;
      const double *Rk = B + (size_t)k * W;
      for (int j = k + 1; j <= iend; ++j) Ri[j - i + w] -= l * Rk[j - k + w];
    }
  }
  double *y = p.rhs.data();
  for (int i = 0; i < n; ++i) {
    double s = y[i];
    const double *Ri = B + (size_t)i * W;
    for (int j = std::max(0, i - w); j < i; ++j) s -= Ri[j - i + w] * y[j];
    y[i] = s;
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = y[i];
    const double *Ri = B + (size_t)i * W;
    for (int j = i + 1; j <= std::min(n - 1, i + w); ++j) s -= Ri[j - i + w] * y[j];
    y[i] = s / Ri[w];
  }
  for (int r = 0; r < n; ++r) x[r] += y[P(r)];
  return 0;
}

}  // namespace

extern "C" int orc_band_width(int n_nodes, int n_tris, const int32_t *tris, int32_t *order) {
  std::vector<std::vector<int>> nb = node_neighbours(n_nodes, n_tris, tris);
  std::vector<int> ord = rcm_order(nb);
  std::vector<int> pos(n_nodes);
  for (int k = 0; k < n_nodes; ++k) pos[ord[k]] = k;
  if (order)
    for (int k = 0; k < n_nodes; ++k) order[k] = ord[k];
  int wn = 0;
  for (int i = 0; i < n_nodes; ++i)
    for (int j : nb[i]) wn = std::max(wn, std::abs(pos[i] - pos[j]));
  return wn;
}

extern "C" int orc_band_newton_step(orc_cell *c, double *x) {
  BandPlan p = band_plan(c);
  return band_newton_step(c, p, x);
}

extern "C" double orc_batch_band_newton_step(int n_cells, orc_cell *cells, double *const *x, int n_threads, int *err) {
  std::vector<BandPlan> plans(n_cells);
#pragma omp parallel for schedule(dynamic) num_threads(n_threads)
  for (int i = 0; i < n_cells; ++i) plans[i] = band_plan(&cells[i]);
  int bad = 0;
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  // largest cells first: their plans are ordered by the caller's list, the
  // dynamic schedule hands out one cell at a time
#pragma omp parallel for schedule(dynamic) num_threads(n_threads) reduction(+ : bad)
  for (int i = 0; i < n_cells; ++i) bad += band_newton_step(&cells[i], plans[i], x[i]) != 0;
  clock_gettime(CLOCK_MONOTONIC, &t1);
  if (err) *err = bad;
  return (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
}
