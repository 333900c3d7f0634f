// Golden-vector harness for the reference mechanics hot path.
//
// This file is test infrastructure.  It is compiled (by oracle/Makefile, target
// `ref`) together with the reference's own UL_growth_explicit.cpp and
// nondestructivetrimesh.cpp, straight from the reference tree, into
// oracle/_ref/refharness_<variant>.  It only marshals raw little-endian arrays
// in and out of the reference functions; it contains no mechanics of its own.
//
// Usage:  refharness_<variant> <mode> <case_dir>
//   modes: elem | calcres | forcefa | updatefa | areatot | bend | areacon | ap | forcefasubs | calcressubs |
//          vtkcell | vtksubs | time (per-call seconds of the per-step residual work)
// Inputs are read from <case_dir>/<name>.bin (float64 or int64, row-major),
// outputs written to <case_dir>/out_<name>.bin.  Scalars come from
// <case_dir>/meta.txt ("key value" lines).
//
// Reference entry points called (wholeCellFE_*/cellBP/UL_growth_explicit.cpp):
//   evalElementRe (cell overload)      UL:710
//   calculateRes  (cell overload)      UL:632
//   calculateForce_FA                  UL:1198 (kMD) / UL1:1196 (k1, k31)
//   updateFA (11-argument form)        UL:2026 (kMD) / UL1:2005 (k1, k31)
//   calculateAreaTot                   UL:1662
//   calculateBendingForce2D            UL:1571
//   calculateAreaConstraintForce       UL:1691 (tot_area from calculateAreaTot at x)
//   calculateForce_AP                  UL:1477 (its own srand(time(NULL)) / rand() draws)
//   calculateForce_FA_subs             UL:1305 (kMD) / UL1:1297 (k1, k31)
//   calculateRes (substrate overload)  UL:931 with mat = 2
//   write_vtk (cell frame)             iomesh.cpp:844 (mode vtkcell, writes <case_dir>/out.vtk)
//   write_vtk (substrate frame)        iomesh.cpp:730 (mode vtksubs)
#include "vec.h"
#include "nondestructivetrimesh.h"
#include <Eigen/Dense>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <string>
#include <vector>

void calculateRes(const std::vector<Vec3d> &X_t, const std::vector<Vec3d> &x_t, std::vector<Eigen::VectorXd> &ea_t,
                  const std::vector<double> &c_A, const std::vector<double> &c_C, const NonDestructiveTriMesh &mesh,
                  double dt_real, Eigen::VectorXd &Res, Eigen::VectorXd &Res_ea, Eigen::VectorXd &MM,
                  Eigen::VectorXd &Res_Adif, Eigen::VectorXd &Res_Cdif, int mat, Eigen::Vector3d &centroid,
                  double time);
void evalElementRe(Eigen::Vector3d &node1_X, Eigen::Vector3d &node2_X, Eigen::Vector3d &node3_X,
                   Eigen::Vector3d &node1_x, Eigen::Vector3d &node2_x, Eigen::Vector3d &node3_x,
                   Eigen::VectorXd &node1_ea, Eigen::VectorXd &node2_ea, Eigen::VectorXd &node3_ea,
                   std::vector<double> &node_c_A, std::vector<double> &node_c_C, double dt_real, Eigen::VectorXd &Re,
                   Eigen::VectorXd &Re_ea, Eigen::VectorXd &Re_Adif, Eigen::VectorXd &Re_Cdif, Eigen::Vector3d &MMe,
                   int mat, Eigen::Vector3d &centroid, double time);
void calculateForce_FA(Eigen::VectorXd &F_constraint, const std::vector<Vec3d> &X_t, const std::vector<Vec3d> &x_t,
                       const std::vector<double> &c_C, std::vector<Vec3d> &fa_u, const NonDestructiveTriMesh &mesh,
                       double kint);
void updateFA(std::vector<double> &c_A, std::vector<double> &c_C, std::vector<Vec3d> &fa_u, double dt_real,
              const std::vector<Vec3d> &X_t_cell, const std::vector<Vec3d> &x_t_cell,
              const NonDestructiveTriMesh &mesh_cell, const std::vector<Vec3d> &X_t_subs,
              const std::vector<Vec3d> &x_t_subs, const NonDestructiveTriMesh &mesh_subs, double kint);
Eigen::Vector4d calculateAreaTot(const std::vector<Vec3d> &x_t, const NonDestructiveTriMesh &mesh);
void calculateForce_FA_subs(Eigen::VectorXd &F_constraint, const std::vector<Vec3d> &X_t, const std::vector<Vec3d> &x_t,
                            const NonDestructiveTriMesh &mesh, const std::vector<double> &c_C,
                            const std::vector<Vec3d> &x_cell, std::vector<Vec3d> &fa_u,
                            const NonDestructiveTriMesh &mesh_cell, double kint);
void calculateRes(const std::vector<Vec3d> &X_t, const std::vector<Vec3d> &x_t, std::vector<Eigen::VectorXd> &ea_t,
                  const NonDestructiveTriMesh &mesh, double dt_real, Eigen::VectorXd &Res, Eigen::VectorXd &Res_ea,
                  Eigen::VectorXd &MM, int mat, double time);
void calculateForce_AP(Eigen::VectorXd &F_constraint, const std::vector<Vec3d> &X_t, const std::vector<Vec3d> &x_t,
                       const std::vector<double> &c_A, const NonDestructiveTriMesh &mesh, double dt_real);
void calculateBendingForce2D(Eigen::VectorXd &F_bending, const std::vector<Vec3d> &X_t, const std::vector<Vec3d> &x_t,
                             const NonDestructiveTriMesh &mesh);
void calculateAreaConstraintForce(Eigen::VectorXd &F_areaConstraint, const std::vector<Vec3d> &X_t,
                                  const std::vector<Vec3d> &x_t, const NonDestructiveTriMesh &mesh, double tot_area);

void write_vtk(const NonDestructiveTriMesh &mesh, const std::vector<Vec3d> &x, const std::vector<Eigen::VectorXd> &ea_t,
               const std::vector<Vec3d> &fa_u, double current_time, const char *filename);
// the cell writer takes the force law's kint in the kMD variant only
#ifdef REF_VARIANT_kMD
void write_vtk(const NonDestructiveTriMesh &mesh, const std::vector<Vec3d> &x, const std::vector<double> &c_A,
               const std::vector<double> &c_B, const std::vector<double> &c_C, const std::vector<Vec3d> &fa_u,
               const std::vector<Eigen::VectorXd> &ea_t, double current_time, const char *filename, double kint);
#else
void write_vtk(const NonDestructiveTriMesh &mesh, const std::vector<Vec3d> &x, const std::vector<double> &c_A,
               const std::vector<double> &c_B, const std::vector<double> &c_C, const std::vector<Vec3d> &fa_u,
               const std::vector<Eigen::VectorXd> &ea_t, double current_time, const char *filename);
#endif

namespace {

std::string g_dir;

template <class T> std::vector<T> load(const std::string &name) {
  std::string p = g_dir + "/" + name + ".bin";
  std::ifstream f(p, std::ios::binary | std::ios::ate);
  if (!f) { std::fprintf(stderr, "refharness: cannot open %s\n", p.c_str()); std::exit(2); }
  std::streamsize sz = f.tellg();
  f.seekg(0);
  std::vector<T> v(sz / sizeof(T));
  f.read(reinterpret_cast<char *>(v.data()), sz);
  return v;
}

void store(const std::string &name, const double *p, size_t n) {
  std::string path = g_dir + "/out_" + name + ".bin";
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char *>(p), n * sizeof(double));
}

std::map<std::string, double> load_meta() {
  std::map<std::string, double> m;
  std::ifstream f(g_dir + "/meta.txt");
  std::string k;
  double v;
  while (f >> k >> v) m[k] = v;
  return m;
}

std::vector<Vec3d> vec3(const std::vector<double> &a) {
  std::vector<Vec3d> v(a.size() / 3);
  for (size_t i = 0; i < v.size(); ++i) v[i] = Vec3d(a[3 * i], a[3 * i + 1], a[3 * i + 2]);
  return v;
}

void build_mesh(NonDestructiveTriMesh &mesh, const std::vector<long long> &t) {
  std::vector<Vec3st> tris(t.size() / 3);
  for (size_t i = 0; i < tris.size(); ++i) tris[i] = Vec3st(t[3 * i], t[3 * i + 1], t[3 * i + 2]);
  mesh.replace_all_triangles(tris);
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 3) { std::fprintf(stderr, "usage: %s mode case_dir\n", argv[0]); return 2; }
  std::string mode = argv[1];
  g_dir = argv[2];
  std::map<std::string, double> meta = load_meta();

  std::vector<Vec3d> X = vec3(load<double>("X"));
  std::vector<Vec3d> x = vec3(load<double>("x"));
  std::vector<long long> tri = load<long long>("tris");
  NonDestructiveTriMesh mesh;
  build_mesh(mesh, tri);
  const size_t n = X.size();

  if (mode == "areatot") {
    Eigen::Vector4d a = calculateAreaTot(x, mesh);
    store("areatot", a.data(), 4);
    return 0;
  }
  if (mode == "bend" || mode == "areacon" || mode == "ap") {
    Eigen::VectorXd F(3 * n);
    F.setZero();  // the caller zeroes the force vector (UL:561, :578, :585)
    if (mode == "bend") {
      calculateBendingForce2D(F, X, x, mesh);
    } else if (mode == "areacon") {
      Eigen::Vector4d a = calculateAreaTot(x, mesh);
      calculateAreaConstraintForce(F, X, x, mesh, a(0));
      store("areatot", a.data(), 4);
    } else {
      std::vector<double> cA(n, 100.0);
      calculateForce_AP(F, X, x, cA, mesh, meta["dt"]);
    }
    store("F", F.data(), 3 * n);
    return 0;
  }
  if (mode == "elem" || mode == "calcres") {
    std::vector<double> ea_raw = load<double>("ea");
    std::vector<double> cA = load<double>("cA"), cC = load<double>("cC");
    std::vector<Eigen::VectorXd> ea(n);
    for (size_t i = 0; i < n; ++i) {
      ea[i].resize(6);
      for (int q = 0; q < 6; ++q) ea[i](q) = ea_raw[6 * i + q];
    }
    double dt = meta["dt"], time = meta["time"];
    int mat = 1;
    Eigen::Vector4d area = calculateAreaTot(x, mesh);
    Eigen::Vector3d centroid(area(1), area(2), area(3));
    if (mode == "calcres") {
      Eigen::VectorXd Res(3 * n), Res_ea(6 * n), MM(n), Ad(n), Cd(n);
      calculateRes(X, x, ea, cA, cC, mesh, dt, Res, Res_ea, MM, Ad, Cd, mat, centroid, time);
      store("Res", Res.data(), 3 * n);
      store("Res_ea", Res_ea.data(), 6 * n);
      store("MM", MM.data(), n);
      store("Res_Adif", Ad.data(), n);
      store("Res_Cdif", Cd.data(), n);
      store("centroid", centroid.data(), 3);
    } else {
      size_t T = tri.size() / 3;
      std::vector<double> Re_all(9 * T), Rea_all(18 * T), MMe_all(3 * T), Ad_all(3 * T), Cd_all(3 * T);
      Eigen::VectorXd Re(9), Re_ea(18), Re_A(3), Re_C(3);
      Eigen::Vector3d MMe;
      for (size_t e = 0; e < T; ++e) {
        size_t v[3] = {(size_t)tri[3 * e], (size_t)tri[3 * e + 1], (size_t)tri[3 * e + 2]};
        Eigen::Vector3d nX[3], nx[3];
        Eigen::VectorXd nea[3];
        std::vector<double> ncA(3), ncC(3);
        for (int a = 0; a < 3; ++a) {
          nX[a] << X[v[a]][0], X[v[a]][1], X[v[a]][2];
          nx[a] << x[v[a]][0], x[v[a]][1], x[v[a]][2];
          nea[a] = ea[v[a]];
          ncA[a] = cA[v[a]];
          ncC[a] = cC[v[a]];
        }
        Re.setZero(); Re_ea.setZero(); MMe.setZero(); Re_A.setZero(); Re_C.setZero();
        evalElementRe(nX[0], nX[1], nX[2], nx[0], nx[1], nx[2], nea[0], nea[1], nea[2], ncA, ncC, dt, Re, Re_ea,
                      Re_A, Re_C, MMe, mat, centroid, time);
        for (int k = 0; k < 9; ++k) Re_all[9 * e + k] = Re(k);
        for (int k = 0; k < 18; ++k) Rea_all[18 * e + k] = Re_ea(k);
        for (int k = 0; k < 3; ++k) {
          MMe_all[3 * e + k] = MMe(k);
          Ad_all[3 * e + k] = Re_A(k);
          Cd_all[3 * e + k] = Re_C(k);
        }
      }
      store("Re", Re_all.data(), Re_all.size());
      store("Re_ea", Rea_all.data(), Rea_all.size());
      store("MMe", MMe_all.data(), MMe_all.size());
      store("Re_Adif", Ad_all.data(), Ad_all.size());
      store("Re_Cdif", Cd_all.data(), Cd_all.size());
    }
    return 0;
  }
  if (mode == "forcefa") {
    std::vector<double> cC = load<double>("cC");
    std::vector<Vec3d> fau = vec3(load<double>("fau"));
    Eigen::VectorXd F(3 * n);
    calculateForce_FA(F, X, x, cC, fau, mesh, meta["kint"]);
    store("F_FA", F.data(), 3 * n);
    return 0;
  }
  if (mode == "forcefasubs") {
    std::vector<double> cC = load<double>("cC");
    std::vector<Vec3d> fau = vec3(load<double>("fau"));
    std::vector<Vec3d> SX = vec3(load<double>("SX")), Sx = vec3(load<double>("Sx"));
    NonDestructiveTriMesh smesh;
    build_mesh(smesh, load<long long>("Stris"));
    Eigen::VectorXd F(3 * SX.size());
    calculateForce_FA_subs(F, SX, Sx, smesh, cC, x, fau, mesh, meta["kint"]);
    store("F_subs", F.data(), 3 * SX.size());
    return 0;
  }
  if (mode == "vtkcell" || mode == "vtksubs") {
    std::vector<double> ea_raw = load<double>("ea");
    std::vector<Eigen::VectorXd> ea(n);
    for (size_t i = 0; i < n; ++i) {
      ea[i].resize(6);
      for (int q = 0; q < 6; ++q) ea[i](q) = ea_raw[6 * i + q];
    }
    std::vector<Vec3d> fau = vec3(load<double>("fau"));
    const std::string out = g_dir + "/out.vtk";
    if (mode == "vtksubs") {
      write_vtk(mesh, x, ea, fau, meta["time"], out.c_str());
    } else {
      std::vector<double> cA = load<double>("cA"), cC = load<double>("cC"), cB(n, 0.0);
#ifdef REF_VARIANT_kMD
      write_vtk(mesh, x, cA, cB, cC, fau, ea, meta["time"], out.c_str(), meta["kint"]);
#else
      write_vtk(mesh, x, cA, cB, cC, fau, ea, meta["time"], out.c_str());
#endif
    }
    return 0;
  }
  if (mode == "calcressubs") {
    std::vector<double> ea_raw = load<double>("ea");
    std::vector<Eigen::VectorXd> ea(n);
    for (size_t i = 0; i < n; ++i) {
      ea[i].resize(6);
      for (int q = 0; q < 6; ++q) ea[i](q) = ea_raw[6 * i + q];
    }
    Eigen::VectorXd Res(3 * n), Res_ea(6 * n), MM(n);
    calculateRes(X, x, ea, mesh, meta["dt"], Res, Res_ea, MM, 2, meta["time"]);
    store("Res", Res.data(), 3 * n);
    store("Res_ea", Res_ea.data(), 6 * n);
    store("MM", MM.data(), n);
    return 0;
  }
  if (mode == "updatefa") {
    std::vector<double> cA = load<double>("cA"), cC = load<double>("cC");
    std::vector<Vec3d> fau = vec3(load<double>("fau"));
    std::vector<Vec3d> SX = vec3(load<double>("SX")), Sx = vec3(load<double>("Sx"));
    NonDestructiveTriMesh smesh;
    build_mesh(smesh, load<long long>("Stris"));
    updateFA(cA, cC, fau, meta["dt"], X, x, mesh, SX, Sx, smesh, meta["kint"]);
    std::vector<double> fo(3 * n);
    for (size_t i = 0; i < n; ++i)
      for (int c = 0; c < 3; ++c) fo[3 * i + c] = fau[i][c];
    store("cA", cA.data(), n);
    store("cC", cC.data(), n);
    store("fau", fo.data(), 3 * n);
    return 0;
  }
  if (mode == "time") {
    // the reference's per-step residual work on these inputs, `reps` times:
    // calculateAreaTot + calculateRes (cell law), calculateForce_FA, updateFA
    // against the substrate (each timed on its own; wall clock, one thread)
    std::vector<double> ea_raw = load<double>("ea");
    std::vector<double> cA = load<double>("cA"), cC = load<double>("cC");
    std::vector<Vec3d> fau = vec3(load<double>("fau"));
    std::vector<Vec3d> SX = vec3(load<double>("SX")), Sx = vec3(load<double>("Sx"));
    NonDestructiveTriMesh smesh;
    build_mesh(smesh, load<long long>("Stris"));
    std::vector<Eigen::VectorXd> ea(n);
    for (size_t i = 0; i < n; ++i) {
      ea[i].resize(6);
      for (int q = 0; q < 6; ++q) ea[i](q) = ea_raw[6 * i + q];
    }
    const int reps = (int)meta["reps"];
    const double dt = meta["dt"], time = meta["time"], kint = meta["kint"];
    double t_res = 0, t_fa = 0, t_upd = 0, sink = 0;
    for (int r = 0; r < reps; ++r) {
      auto t0 = std::chrono::steady_clock::now();
      Eigen::Vector4d area = calculateAreaTot(x, mesh);
      Eigen::Vector3d centroid(area(1), area(2), area(3));
      Eigen::VectorXd Res(3 * n), Res_ea(6 * n), MM(n), Ad(n), Cd(n);
      calculateRes(X, x, ea, cA, cC, mesh, dt, Res, Res_ea, MM, Ad, Cd, 1, centroid, time);
      auto t1 = std::chrono::steady_clock::now();
      Eigen::VectorXd F(3 * n);
      calculateForce_FA(F, X, x, cC, fau, mesh, kint);
      auto t2 = std::chrono::steady_clock::now();
      std::vector<double> cA2(cA), cC2(cC);
      std::vector<Vec3d> fau2(fau);
      updateFA(cA2, cC2, fau2, dt, X, x, mesh, SX, Sx, smesh, kint);
      auto t3 = std::chrono::steady_clock::now();
      t_res += std::chrono::duration<double>(t1 - t0).count();
      t_fa += std::chrono::duration<double>(t2 - t1).count();
      t_upd += std::chrono::duration<double>(t3 - t2).count();
      sink += Res(0) + F(0) + cC2[0];
    }
    double out[4] = {t_res / reps, t_fa / reps, t_upd / reps, sink};
    store("time", out, 4);
    return 0;
  }
  std::fprintf(stderr, "refharness: unknown mode %s\n", mode.c_str());
  return 2;
}
