// Exercises include/wcfe_compat.hpp the way the reference's call sites would:
// std::vector<Vec3d>-like positions, VectorXd-like strains, a
// NonDestructiveTriMesh-like mesh, and the reference function names.
// Stand-in types with the same member functions as the reference's
// (vec.h Vec3d operator[], Eigen::VectorXd operator()/resize,
// NonDestructiveTriMesh num_triangles/get_triangle) keep this file free of
// the reference's headers.
//
// Usage: compat_check <case_dir> <mode>   (mode: calcres | calcressubs | forcefa |
//                                          forcesubs | updatefa | boundary)
// Reads <name>.bin arrays written by tests/test_gpu_compat.py and writes
// out_<name>.bin results.
#include <array>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "../include/wcfe_compat.hpp"

struct Vec3 {
  double v[3];
  double &operator[](int i) { return v[i]; }
  const double &operator[](int i) const { return v[i]; }
};
struct VecX {
  std::vector<double> d;
  void resize(size_t n) { d.resize(n); }
  double &operator()(size_t i) { return d[i]; }
  const double &operator()(size_t i) const { return d[i]; }
};
struct Vec4 {
  double v[4];
  double &operator()(int i) { return v[i]; }
};
struct Mesh {
  std::vector<std::array<size_t, 3>> tris;
  size_t num_triangles() const { return tris.size(); }
  const std::array<size_t, 3> &get_triangle(size_t i) const { return tris[i]; }
};

static std::string dir;
template <class T> std::vector<T> load(const std::string &n) {
  std::ifstream f(dir + "/" + n + ".bin", std::ios::binary | std::ios::ate);
  size_t sz = f.tellg();
  f.seekg(0);
  std::vector<T> v(sz / sizeof(T));
  f.read(reinterpret_cast<char *>(v.data()), sz);
  return v;
}
static void store(const std::string &n, const std::vector<double> &v) {
  std::ofstream f(dir + "/out_" + n + ".bin", std::ios::binary);
  f.write(reinterpret_cast<const char *>(v.data()), v.size() * sizeof(double));
}
static std::vector<Vec3> vec3(const std::vector<double> &a) {
  std::vector<Vec3> o(a.size() / 3);
  for (size_t i = 0; i < o.size(); ++i) o[i] = Vec3{{a[3 * i], a[3 * i + 1], a[3 * i + 2]}};
  return o;
}
static Mesh mesh(const std::vector<int> &t) {
  Mesh m;
  for (size_t i = 0; i + 2 < t.size(); i += 3) m.tris.push_back({(size_t)t[i], (size_t)t[i + 1], (size_t)t[i + 2]});
  return m;
}

int main(int argc, char **argv) {
  if (argc != 3) return 2;
  dir = argv[1];
  std::string mode = argv[2];
  try {
    std::vector<double> meta = load<double>("meta");  // time, dt, kint, law
    wcfe_compat::force_law() = (int)meta[3];
    Mesh m = mesh(load<int>("tris"));
    std::vector<Vec3> X = vec3(load<double>("X")), x = vec3(load<double>("x"));
    if (mode == "calcres") {
      std::vector<double> e = load<double>("ea");
      std::vector<VecX> ea(X.size());
      for (size_t i = 0; i < X.size(); ++i) {
        ea[i].resize(6);
        for (int q = 0; q < 6; ++q) ea[i](q) = e[6 * i + q];
      }
      std::vector<double> cA(X.size(), 0.0), cC(X.size(), 0.0);
      VecX Res, Rea, MM, RA, RC;
      Vec3 centroid = {{0.0, 0.0, 0.0}};  // stands for the reference's Eigen::Vector3d lvalue (UL:543)
      wcfe_compat::calculateRes(X, x, ea, cA, cC, m, meta[1], Res, Rea, MM, RA, RC, 1, centroid, meta[0]);
      store("Res", Res.d);
      store("Res_ea", Rea.d);
      store("MM", MM.d);
    } else if (mode == "calcressubs") {
      // the substrate's call (UL:451): 10-argument overload, mat = 2
      std::vector<double> e = load<double>("ea");
      std::vector<VecX> ea(X.size());
      for (size_t i = 0; i < X.size(); ++i) {
        ea[i].resize(6);
        for (int q = 0; q < 6; ++q) ea[i](q) = e[6 * i + q];
      }
      VecX Res, Rea, MM;
      wcfe_compat::calculateRes(X, x, ea, m, meta[1], Res, Rea, MM, 2, meta[0]);
      store("Res", Res.d);
      store("Res_ea", Rea.d);
      store("MM", MM.d);
    } else if (mode == "boundary") {
      // updateForces_Explicit's boundary calls (UL:527, :565, :582, :587)
      Vec4 area = wcfe_compat::calculateAreaTot<Vec4>(x, m);
      VecX Fb, Fa, Fp;
      wcfe_compat::calculateBendingForce2D(Fb, X, x, m);
      wcfe_compat::calculateAreaConstraintForce(Fa, X, x, m, area(0));
      std::vector<double> cA(X.size(), 100.0);
      wcfe_compat::calculateForce_AP(Fp, X, x, cA, m, meta[1]);
      store("areatot", std::vector<double>(area.v, area.v + 4));
      store("F_bend", Fb.d);
      store("F_area", Fa.d);
      store("F_ap", Fp.d);
    } else if (mode == "forcesubs") {
      // the substrate's updateForces_Explicit (UL:467): X, x are the
      // substrate's, the cell enters through x_cell, c_C, fa_u, mesh_cell
      std::vector<double> cC = load<double>("cC");
      std::vector<Vec3> fau = vec3(load<double>("fau")), xc = vec3(load<double>("x_cell"));
      Mesh cm = mesh(load<int>("cell_tris"));
      VecX F;
      wcfe_compat::calculateForce_FA_subs(F, X, x, m, cC, xc, fau, cm, meta[2]);
      store("F_subs", F.d);
    } else if (mode == "forcefa") {
      std::vector<double> cC = load<double>("cC");
      std::vector<Vec3> fau = vec3(load<double>("fau"));
      VecX F;
      wcfe_compat::calculateForce_FA(F, X, x, cC, fau, m, meta[2]);
      store("F_FA", F.d);
    } else {
      std::vector<double> cA = load<double>("cA"), cC = load<double>("cC");
      std::vector<Vec3> fau = vec3(load<double>("fau"));
      std::vector<Vec3> SX = vec3(load<double>("SX")), Sx = vec3(load<double>("Sx"));
      Mesh sm = mesh(load<int>("Stris"));
      wcfe_compat::updateFA(cA, cC, fau, meta[1], X, x, m, SX, Sx, sm, meta[2]);
      std::vector<double> f(3 * fau.size());
      for (size_t i = 0; i < fau.size(); ++i)
        for (int c = 0; c < 3; ++c) f[3 * i + c] = fau[i][c];
      store("cA", cA);
      store("cC", cC);
      store("fau", f);
    }
  } catch (const wcfe_compat::error &e) {
    std::fprintf(stderr, "compat_check: wcfe error %d: %s\n", e.code, e.what());
    return 3;
  }
  return 0;
}
