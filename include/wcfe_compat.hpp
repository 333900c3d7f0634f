// wcfe_compat.hpp -- adapters with the reference's call signatures.
//
// Drop-in for the hot-path call sites of wholeCellFE's advance_sim
// (wholeCellFE_kMD/cellBP/main.cpp:356-363) and for updateForces_Explicit's
// inner calls (UL_growth_explicit.cpp:527, :543, :562, :565, :582, :587).  Each adapter has the
// signature and argument meaning of the reference function it replaces and
// forwards to the C ABI in wcfe.h (GPU kernels for gfx950); El Topo meshes,
// VTK output and everything else stay in the host application.
//
// The adapters are templates over the reference's types so this header needs
// none of the reference's headers: Vec3 is anything with operator[] (Vec3d,
// vec.h:23), VecX anything with operator()(i) and resize(n) (Eigen::VectorXd),
// Mesh anything with num_triangles() and get_triangle(i)[k]
// (NonDestructiveTriMesh, nondestructivetrimesh.h:49-50).
//
// Error behaviour: the reference functions return void and never fail; the
// adapters throw wcfe_compat::error (a std::runtime_error carrying the wcfe
// error code) when the library reports one, e.g. an out-of-plane node.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "wcfe.h"

namespace wcfe_compat {

struct error : std::runtime_error {
  int code;
  error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// One library context per host thread (the reference is single threaded).
inline wcfe_ctx *context() {
  thread_local struct Holder {
    wcfe_ctx *ctx = nullptr;
    ~Holder() {
      if (ctx) wcfe_destroy(ctx);
    }
  } h;
  if (!h.ctx) {
    int rc = wcfe_create(nullptr, &h.ctx);
    if (rc != WCFE_OK) {
      std::string m = h.ctx ? wcfe_last_error(h.ctx) : "no gfx950 device";
      if (h.ctx) wcfe_destroy(h.ctx);
      h.ctx = nullptr;
      throw error(rc, m);
    }
  }
  return h.ctx;
}

inline void check(int rc) {
  if (rc != WCFE_OK) throw error(rc, wcfe_last_error(context()));
}

// Force law of calculateForce_FA / updateFA: the reference selects it by
// compiling a different folder (k1, k31: linear kint, UL1:1259-1261;
// kMD: 6-knot MD table, UL:1204-1264).  0 = linear, 1 = MD table.
inline int &force_law() {
  thread_local int law = 1;
  return law;
}

template <class Vec3> std::vector<double> flat3(const std::vector<Vec3> &v) {
  std::vector<double> o(3 * v.size());
  for (size_t i = 0; i < v.size(); ++i)
    for (int c = 0; c < 3; ++c) o[3 * i + c] = v[i][c];
  return o;
}

template <class Mesh> std::vector<int32_t> flat_tris(const Mesh &mesh) {
  const size_t T = mesh.num_triangles();
  std::vector<int32_t> t(3 * T);
  for (size_t e = 0; e < T; ++e)
    for (int a = 0; a < 3; ++a) t[3 * e + a] = (int32_t)mesh.get_triangle(e)[a];
  return t;
}

// calculateRes, cell overload (UL_growth_explicit.cpp:632): FP64 residual,
// strain increment and lumped mass of the cell law (mat = 1) or the
// substrate law (mat = 2) on the GPU.
// Res_Adif / Res_Cdif (diffusion terms the reference computes but never uses,
// UL:617-618) are returned as zero vectors.  `centroid` takes the reference's
// Eigen::Vector3d (UL:632) unchanged; its value is not read: it only sets the
// fibre direction a0 of the active stress t_act (kappa I + (1 - 2 kappa) a0 a0),
// and kappa = 0.5 (UL:872) makes that term vanish.
template <class Vec3, class VecX, class Mesh, class Centroid>
void calculateRes(const std::vector<Vec3> &X_t, const std::vector<Vec3> &x_t, std::vector<VecX> &ea_t,
                  const std::vector<double> &c_A, const std::vector<double> &c_C, const Mesh &mesh, double dt_real,
                  VecX &Res, VecX &Res_ea, VecX &MM, VecX &Res_Adif, VecX &Res_Cdif, int mat, const Centroid &centroid,
                  double time) {
  (void)c_A; (void)c_C; (void)dt_real; (void)centroid;
  const int n = (int)X_t.size();
  std::vector<double> X = flat3(X_t), x = flat3(x_t), ea(6 * (size_t)n);
  for (int i = 0; i < n; ++i)
    for (int q = 0; q < 6; ++q) ea[6 * i + q] = ea_t[i](q);
  std::vector<int32_t> tris = flat_tris(mesh);
  std::vector<double> R(3 * (size_t)n), Rea(6 * (size_t)n), M(n);
  check(wcfe_calculate_res(context(), n, (int32_t)(tris.size() / 3), tris.data(), X.data(), x.data(), ea.data(), mat,
                           time, R.data(), Rea.data(), M.data()));
  Res.resize(3 * n);
  Res_ea.resize(6 * n);
  MM.resize(n);
  Res_Adif.resize(n);
  Res_Cdif.resize(n);
  for (int i = 0; i < 3 * n; ++i) Res(i) = R[i];
  for (int i = 0; i < 6 * n; ++i) Res_ea(i) = Rea[i];
  for (int i = 0; i < n; ++i) {
    MM(i) = M[i];
    Res_Adif(i) = 0.0;
    Res_Cdif(i) = 0.0;
  }
}

// calculateRes, 10-argument overload (UL_growth_explicit.cpp:931), as the
// substrate's updateForces_Explicit calls it with mat = 2 (UL:451).
template <class Vec3, class VecX, class Mesh>
void calculateRes(const std::vector<Vec3> &X_t, const std::vector<Vec3> &x_t, std::vector<VecX> &ea_t,
                  const Mesh &mesh, double dt_real, VecX &Res, VecX &Res_ea, VecX &MM, int mat, double time) {
  VecX RA, RC;
  const std::vector<double> none(X_t.size(), 0.0);
  const double no_centroid[3] = {0.0, 0.0, 0.0};
  calculateRes(X_t, x_t, ea_t, none, none, mesh, dt_real, Res, Res_ea, MM, RA, RC, mat, no_centroid, time);
}

// calculateForce_FA (UL_growth_explicit.cpp:1198 / UL1:1196).
template <class VecX, class Vec3, class Mesh>
void calculateForce_FA(VecX &F_constraint, const std::vector<Vec3> &X_t, const std::vector<Vec3> &x_t,
                       const std::vector<double> &c_C, std::vector<Vec3> &fa_u, const Mesh &mesh, double kint) {
  (void)X_t;
  const int n = (int)x_t.size();
  std::vector<double> x = flat3(x_t), fau = flat3(fa_u), F(3 * (size_t)n);
  std::vector<int32_t> tris = flat_tris(mesh);
  check(wcfe_force_fa(context(), n, (int32_t)(tris.size() / 3), tris.data(), x.data(), c_C.data(), fau.data(),
                      force_law(), kint, F.data()));
  F_constraint.resize(3 * n);
  for (int i = 0; i < 3 * n; ++i) F_constraint(i) = F[i];
}

// calculateForce_FA_subs (UL_growth_explicit.cpp:1305 / UL1:1297), as called
// by the substrate's updateForces_Explicit (UL:467).  F_constraint receives 3
// components per substrate node.
template <class VecX, class Vec3, class Mesh>
void calculateForce_FA_subs(VecX &F_constraint, const std::vector<Vec3> &X_t, const std::vector<Vec3> &x_t,
                            const Mesh &mesh, const std::vector<double> &c_C, const std::vector<Vec3> &x_cell,
                            std::vector<Vec3> &fa_u, const Mesh &mesh_cell, double kint) {
  (void)X_t;
  const int m = (int)x_t.size(), n = (int)x_cell.size();
  std::vector<double> Sx = flat3(x_t), xc = flat3(x_cell), fau = flat3(fa_u), F(3 * (size_t)m);
  std::vector<int32_t> st = flat_tris(mesh), ct = flat_tris(mesh_cell);
  check(wcfe_force_fa_subs(context(), n, (int32_t)(ct.size() / 3), ct.data(), xc.data(), c_C.data(), fau.data(), m,
                           (int32_t)(st.size() / 3), st.data(), Sx.data(), force_law(), kint, F.data()));
  F_constraint.resize(3 * m);
  for (int i = 0; i < 3 * m; ++i) F_constraint(i) = F[i];
}

// updateFA, 11-argument form (UL_growth_explicit.cpp:2026 / UL1:2005), as
// called at main.cpp:356.
template <class Vec3, class Mesh>
void updateFA(std::vector<double> &c_A, std::vector<double> &c_C, std::vector<Vec3> &fa_u, double dt_real,
              const std::vector<Vec3> &X_t_cell, const std::vector<Vec3> &x_t_cell, const Mesh &mesh_cell,
              const std::vector<Vec3> &X_t_subs, const std::vector<Vec3> &x_t_subs, const Mesh &mesh_subs,
              double kint) {
  (void)mesh_cell;
  const int n = (int)X_t_cell.size();
  std::vector<double> X = flat3(X_t_cell), x = flat3(x_t_cell), fau = flat3(fa_u);
  std::vector<double> SX = flat3(X_t_subs), Sx = flat3(x_t_subs);
  std::vector<int32_t> st = flat_tris(mesh_subs);
  int32_t unlocated = 0;
  check(wcfe_update_fa(context(), n, X.data(), x.data(), c_A.data(), c_C.data(), fau.data(), dt_real,
                       (int32_t)X_t_subs.size(), (int32_t)(st.size() / 3), st.data(), SX.data(), Sx.data(),
                       force_law(), kint, &unlocated));
  for (int i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) fa_u[i][c] = fau[3 * i + c];
}

// calculateAreaTot (UL_growth_explicit.cpp:1662): {area, centroid} as the
// caller's 4-vector type, e.g. calculateAreaTot<Eigen::Vector4d>(x_t, mesh).
template <class Vec4, class Vec3, class Mesh> Vec4 calculateAreaTot(const std::vector<Vec3> &x_t, const Mesh &mesh) {
  std::vector<double> x = flat3(x_t);
  std::vector<int32_t> tris = flat_tris(mesh);
  double out[4];
  check(wcfe_area_tot(context(), (int32_t)x_t.size(), (int32_t)(tris.size() / 3), tris.data(), x.data(), out));
  Vec4 v;
  for (int k = 0; k < 4; ++k) v(k) = out[k];
  return v;
}

// Boundary vertices whose neighbour search found only one side of the curve
// (the reference reads an uninitialised vector there; the adapters apply no
// force and count them here).
inline int &open_boundary_vertices() {
  thread_local int n = 0;
  return n;
}

// calculateBendingForce2D (UL_growth_explicit.cpp:1571), k_bend = 20 pN/um
// (UL:1653).  The boundary is found from the triangles (El Topo's rule).
template <class VecX, class Vec3, class Mesh>
void calculateBendingForce2D(VecX &F_bending, const std::vector<Vec3> &X_t, const std::vector<Vec3> &x_t,
                             const Mesh &mesh) {
  (void)X_t;
  const int n = (int)x_t.size();
  std::vector<double> x = flat3(x_t), F(3 * (size_t)n);
  std::vector<int32_t> tris = flat_tris(mesh);
  check(wcfe_bending_force_2d(context(), n, (int32_t)(tris.size() / 3), tris.data(), x.data(), 20.0, F.data(),
                              &open_boundary_vertices()));
  F_bending.resize(3 * n);
  for (int i = 0; i < 3 * n; ++i) F_bending(i) = F[i];
}

// calculateAreaConstraintForce (UL_growth_explicit.cpp:1691), k_area = 1,
// area_0 = 10*10*3.1416 (UL:1776-1777).
template <class VecX, class Vec3, class Mesh>
void calculateAreaConstraintForce(VecX &F_areaConstraint, const std::vector<Vec3> &X_t, const std::vector<Vec3> &x_t,
                                  const Mesh &mesh, double tot_area) {
  (void)X_t;
  const int n = (int)x_t.size();
  std::vector<double> x = flat3(x_t), F(3 * (size_t)n);
  std::vector<int32_t> tris = flat_tris(mesh);
  check(wcfe_area_constraint_force(context(), n, (int32_t)(tris.size() / 3), tris.data(), x.data(), tot_area, 1.0,
                                   10 * 10 * 3.1416, F.data(), &open_boundary_vertices()));
  F_areaConstraint.resize(3 * n);
  for (int i = 0; i < 3 * n; ++i) F_areaConstraint(i) = F[i];
}

// calculateForce_AP (UL_growth_explicit.cpp:1477): rate 10 1/s, 5 pN per
// trial, 100 trials per um (UL:1552-1561).  The reference draws from rand()
// reseeded with the clock on every call; the adapter draws from the
// library's Philox stream, a new draw index per call.
template <class VecX, class Vec3, class Mesh>
void calculateForce_AP(VecX &F_constraint, const std::vector<Vec3> &X_t, const std::vector<Vec3> &x_t,
                       const std::vector<double> &c_A, const Mesh &mesh, double dt_real) {
  (void)X_t; (void)c_A;
  thread_local uint32_t draw = 0;
  const int n = (int)x_t.size();
  std::vector<double> x = flat3(x_t), F(3 * (size_t)n);
  std::vector<int32_t> tris = flat_tris(mesh);
  check(wcfe_force_ap(context(), n, (int32_t)(tris.size() / 3), tris.data(), x.data(), dt_real, 10.0, 5.0, 100.0,
                      nullptr, draw++, F.data(), &open_boundary_vertices()));
  F_constraint.resize(3 * n);
  for (int i = 0; i < 3 * n; ++i) F_constraint(i) = F[i];
}

}  // namespace wcfe_compat
