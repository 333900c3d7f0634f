// Per-frame legacy-VTK output in the reference's format (host code).
//
// The reference writes one ASCII VTK file per cell and per substrate at the
// end of every frame (main.cpp:447-454) with write_vtk (iomesh.cpp:844-1047
// cell frame, iomesh.cpp:730-842 substrate frame).  The file layout is an
// interface of the reference's post-processing, so it is kept: the same
// header, POINTS / CELLS / CELL_TYPES blocks, the same point data and the
// same stream formatting (default 6 significant digits; fixed notation in
// the substrate frame).  The point data are derived here:
//   theta_e_t = sqrt(det (I - 2 ea)^-1)            (elastic area change)
//   cell frame, 4th scalar: law 1 (kMD) the MD-table force at |fa_u|
//   (UL:1204-1214 table, iomesh.cpp kMD:924-956); law 0 (k1, k31)
//   |kint * fa_u_xy| (iomesh.cpp k1/k31:923-927).
#include <cmath>
#include <fstream>
#include <string>

#include "../../include/wcfe.h"
#include "wcfe_internal.hpp"

using namespace wcfe;

namespace {

// the 6-knot MD force table (UL:1207-1214): extension [um], force [pN],
// stiffness [pN/um]; the last segment extrapolates
const double MD_PX[6] = {0., 2.51250583e-03, 4.40265201e-03, 5.91408228e-03, 11.76817302e-03, 17.001e-03};
const double MD_PY[6] = {-7.15514497, 71.80501308, 356.62420075, 356.96656556, 126.21811585, 528.40481124};
const double MD_K[6] = {31426.8556, 99485.86235, 67080.67086, -7908.39043, 21978.66923, 76858.39733};

double md_table_force(double D) {
  int j = 0;
  while (j < 5 && D >= MD_PX[j + 1]) ++j;
  return MD_PY[j] + MD_K[j] * (D - MD_PX[j]);
}

// symmetric strain from the input layout: 6 = Voigt [xx,yy,zz,xy,xz,yz],
// 3 = planar [xx,yy,xy]
void strain3(const VtkInput &in, int i, double E[3][3]) {
  const double *e = in.ea ? in.ea + (size_t)in.ea_kind * i : nullptr;
  double xx = 0, yy = 0, zz = 0, xy = 0, xz = 0, yz = 0;
  if (e && in.ea_kind == 6) { xx = e[0]; yy = e[1]; zz = e[2]; xy = e[3]; xz = e[4]; yz = e[5]; }
  if (e && in.ea_kind == 3) { xx = e[0]; yy = e[1]; xy = e[2]; }
  E[0][0] = xx; E[0][1] = xy; E[0][2] = xz;
  E[1][0] = xy; E[1][1] = yy; E[1][2] = yz;
  E[2][0] = xz; E[2][1] = yz; E[2][2] = zz;
}

double theta_e(const VtkInput &in, int i) {
  double E[3][3], A[3][3];
  strain3(in, i, E);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) A[r][c] = (r == c ? 1.0 : 0.0) - 2.0 * E[r][c];
  const double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
                     A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
                     A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
  return std::sqrt(1.0 / det);  // det of the inverse
}

double comp(const double *a, int stride, int i, int c) { return (a && c < stride) ? a[(size_t)stride * i + c] : 0.0; }

}  // namespace

int wcfe::write_vtk_file(const VtkInput &in, std::string &err) {
  std::ofstream f(in.path);
  if (!f) {
    err = std::string("cannot open ") + in.path + " for writing";
    return WCFE_ERR_ARG;
  }
  const int n = in.n_nodes, T = in.n_tris;
  f << "# vtk DataFile Version 2.0\nCellBP Time: " << in.time << "\nASCII\nDATASET UNSTRUCTURED_GRID\n";
  f << "POINTS " << n << " double\n";
  if (in.style == WCFE_VTK_SUBSTRATE) f << std::fixed;
  for (int i = 0; i < n; ++i)
    f << comp(in.x, in.x_stride, i, 0) << " " << comp(in.x, in.x_stride, i, 1) << " " << comp(in.x, in.x_stride, i, 2)
      << "\n";
  f << "CELLS " << T << " " << 4 * T << "\n";
  for (int e = 0; e < T; ++e)
    f << "3 " << in.tris[3 * e] << " " << in.tris[3 * e + 1] << " " << in.tris[3 * e + 2] << "\n";
  f << "CELL_TYPES " << T << "\n";
  for (int e = 0; e < T; ++e) f << "5\n";
  f << "POINT_DATA " << n << "\n";
  if (in.style == WCFE_VTK_SUBSTRATE) {
    f << "SCALARS theta_e float 1\nLOOKUP_TABLE default\n";
    for (int i = 0; i < n; ++i) f << theta_e(in, i) << "\n";
  } else {
    f << "SCALARS concentrations float 4\nLOOKUP_TABLE default\n";
    for (int i = 0; i < n; ++i) {
      const double ux = comp(in.fau, in.fau_stride, i, 0), uy = comp(in.fau, in.fau_stride, i, 1),
                   uz = comp(in.fau, in.fau_stride, i, 2);
      double s4;
      if (in.law == 1) {
        s4 = md_table_force(std::sqrt(ux * ux + uy * uy + uz * uz));
      } else {
        const double fx = in.kint * ux, fy = in.kint * uy;
        s4 = std::sqrt(fx * fx + fy * fy);
      }
      const double cA = in.cA ? in.cA[i] : in.cA_default;
      const double cC = in.cC ? in.cC[i] : 0.0;
      f << cA << " " << theta_e(in, i) << " " << cC << " " << s4 << "\n";
    }
  }
  f << "VECTORS fa_u float\n";
  for (int i = 0; i < n; ++i)
    f << comp(in.fau, in.fau_stride, i, 0) << " " << comp(in.fau, in.fau_stride, i, 1) << " "
      << comp(in.fau, in.fau_stride, i, 2) << "\n";
  f.close();
  if (!f) {
    err = std::string("writing ") + in.path + " failed";
    return WCFE_ERR_ARG;
  }
  return WCFE_OK;
}

extern "C" int wcfe_write_vtk(const char *path, int32_t style, int32_t n_nodes, int32_t n_tris, const int32_t *tris,
                              const double *x, const double *ea, const double *fau, const double *cA,
                              const double *cC, int32_t law, double kint, double time) {
  if (!path || n_nodes <= 0 || n_tris < 0 || (n_tris > 0 && !tris) || !x) return WCFE_ERR_ARG;
  if (style != WCFE_VTK_CELL && style != WCFE_VTK_SUBSTRATE) return WCFE_ERR_ARG;
  if (law != 0 && law != 1) return WCFE_ERR_ARG;
  for (long long k = 0; k < 3LL * n_tris; ++k)
    if (tris[k] < 0 || tris[k] >= n_nodes) return WCFE_ERR_ARG;
  VtkInput in;
  in.path = path;
  in.style = style;
  in.n_nodes = n_nodes;
  in.n_tris = n_tris;
  in.tris = tris;
  in.x = x;
  in.x_stride = 3;
  in.ea = ea;
  in.ea_kind = 6;
  in.fau = fau;
  in.fau_stride = 3;
  in.cA = cA;
  in.cC = cC;
  in.law = law;
  in.kint = kint;
  in.time = time;
  std::string err;
  return write_vtk_file(in, err);
}
