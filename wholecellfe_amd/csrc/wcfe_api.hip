// C ABI of the MI355X per-step cell mechanics hot path (include/wcfe.h):
// batch layout in HBM, colouring, LU work lists, the Newton step sequence and
// its hipGraph capture.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "../../include/wcfe.h"
#include "wcfe_internal.hpp"

using namespace wcfe;

namespace {

struct HipError {
  hipError_t e;
  const char *what;
};
#define HIPCHECK(x)                             \
  do {                                          \
    hipError_t err_ = (x);                      \
    if (err_ != hipSuccess) throw HipError{err_, #x}; \
  } while (0)

struct ApiError {
  int code;
  std::string msg;
};
[[noreturn]] void fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  throw ApiError{code, buf};
}

template <class T> T *dalloc(size_t n) {
  if (n == 0) n = 1;
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, n * sizeof(T));
  if (e != hipSuccess) fail(WCFE_ERR_OOM, "hipMalloc of %zu bytes failed: %s", n * sizeof(T), hipGetErrorString(e));
  return static_cast<T *>(p);
}
template <class T> T *dupload(const std::vector<T> &h) {
  T *d = dalloc<T>(h.size());
  if (!h.empty()) HIPCHECK(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

struct HostCell {
  int n_nodes = 0, n_tris = 0, law = 0, node_springs = 0, spring_law = 0, table_id = 0, kinetics = 0, n_spr = 0;
  int solver = 0;
  int step = 0;       // committed load steps (also the Bell stream's step counter)
  double time = NAN;  // load time; NaN = wcfe_options.t0
  bool has_state = false;  // `state` holds the full solver state (Newton iterate, norms)
  CellState state;
  double mu = 1000, E = 1000, nu = 0.3, thick = 1, rho_max = 100, k_lin = 1000;
  // boundary and body forces of the reference cell model
  int ap = 0;
  double k_bend = 0, k_area = 0, area_0 = 0, drag = 0, ap_rate = 0, ap_quantum = 0, ap_density = 0;
  // substrate form: no active stress; constant nodal load (empty = none)
  int passive = 0;
  std::vector<double> f_ext;
  std::vector<double> X, x, ea, cC, fau, spr_bary, spr_anchor, spr_state;
  std::vector<int> tris, spr_tri;
  uint64_t topo_hash = 0;  // of the triangle list (topology_hash)
  // the band numbering of the mesh, computed at the first layout that needs
  // it; valid for the triangle list whose topology_hash it carries
  mutable std::vector<int> rcm_pos;
  mutable int bw_nat = -1, bw_rcm = -1;
  mutable uint64_t rcm_hash = 0;
};

// FNV-1a over the triangle list: a cell's topology as the host's remeshing
// leaves it (wcfe_remesh_cell lays a cell out again only when it changes).
uint64_t topology_hash(int n_nodes, int n_tris, const int *tris) {
  uint64_t h = 1469598103934665603ULL;
  auto mix = [&](uint32_t v) {
    for (int k = 0; k < 4; ++k) {
      h ^= (v >> (8 * k)) & 0xFFu;
      h *= 1099511628211ULL;
    }
  };
  mix((uint32_t)n_nodes);
  mix((uint32_t)n_tris);
  for (long long i = 0; i < 3LL * n_tris; ++i) mix((uint32_t)tris[i]);
  return h;
}

struct Table {
  std::vector<double> px, py, k;
};

// Greedy first-fit element colouring in element-index order (64 colours max).
int color_elements(int n_nodes, int n_tris, const int *tris, std::vector<int> &color) {
  std::vector<unsigned long long> used(n_nodes, 0ull);
  color.assign(n_tris, 0);
  int nc = 0;
  for (int e = 0; e < n_tris; ++e) {
    const unsigned long long forb = used[tris[3 * e]] | used[tris[3 * e + 1]] | used[tris[3 * e + 2]];
    if (forb == ~0ull) return -1;
    const int col = __builtin_ctzll(~forb);
    color[e] = col;
    for (int a = 0; a < 3; ++a) used[tris[3 * e + a]] |= (1ull << col);
    nc = std::max(nc, col + 1);
  }
  return nc;
}

// Boundary vertices by El Topo's rule (an edge with an odd number of incident
// triangles is a boundary edge, eltopo3d/nondestructivetrimesh.cpp:527-543;
// triangles with t0 == t1 are deleted ones, :497) and, per boundary vertex ni,
// the (nj, nint) pairs the reference's boundary-force loops visit
// (UL_growth_explicit.cpp:1497-1528): one-ring triangles by increasing index,
// corners nj in slot order that are boundary vertices other than ni, nint the
// first corner that is neither.  cst: N+1 offsets into cand (an empty range
// marks an interior vertex).
void boundary_candidates(int n_nodes, int n_tris, const int *tris, std::vector<int> &cst, std::vector<int2> &cand) {
  std::vector<unsigned long long> edges;
  edges.reserve(3 * (size_t)n_tris);
  for (int e = 0; e < n_tris; ++e) {
    const int *t = tris + 3 * e;
    if (t[0] == t[1]) continue;
    for (int j = 0; j < 3; ++j) {
      const unsigned a = (unsigned)std::min(t[j], t[(j + 1) % 3]), b = (unsigned)std::max(t[j], t[(j + 1) % 3]);
      edges.push_back(((unsigned long long)a << 32) | b);
    }
  }
  std::sort(edges.begin(), edges.end());
  std::vector<char> bnd(n_nodes, 0);
  for (size_t i = 0; i < edges.size();) {
    size_t j = i;
    while (j < edges.size() && edges[j] == edges[i]) ++j;
    if ((j - i) % 2 == 1) bnd[edges[i] >> 32] = bnd[edges[i] & 0xFFFFFFFFull] = 1;
    i = j;
  }
  std::vector<int> rs(n_nodes + 1, 0), rl;
  for (int e = 0; e < n_tris; ++e)
    for (int a = 0; a < 3; ++a) rs[tris[3 * e + a] + 1]++;
  for (int i = 0; i < n_nodes; ++i) rs[i + 1] += rs[i];
  rl.assign(rs[n_nodes], 0);
  {
    std::vector<int> fill(rs.begin(), rs.end() - 1);
    for (int e = 0; e < n_tris; ++e)
      for (int a = 0; a < 3; ++a) rl[fill[tris[3 * e + a]]++] = e;
  }
  cst.assign(n_nodes + 1, 0);
  cand.clear();
  for (int ni = 0; ni < n_nodes; ++ni) {
    cst[ni] = (int)cand.size();
    if (!bnd[ni]) continue;
    for (int k = rs[ni]; k < rs[ni + 1]; ++k) {
      const int *t = tris + 3 * rl[k];
      if (t[0] == t[1]) continue;
      for (int enj = 0; enj < 3; ++enj) {
        const int nj = t[enj];
        if (!bnd[nj] || nj == ni) continue;
        int nint = -1;
        for (int enk = 0; enk < 3; ++enk)
          if (t[enk] != nj && t[enk] != ni) {
            nint = t[enk];
            break;
          }
        if (nint >= 0) cand.push_back(make_int2(nj, nint));
      }
    }
  }
  cst[n_nodes] = (int)cand.size();
}

// Reverse Cuthill-McKee order of a mesh's nodes (George & Liu, "Computer
// Solution of Large Sparse Positive Definite Systems", 1981): per connected
// component, a pseudo-peripheral start node (repeated breadth-first searches
// from the last level's lowest-degree node while the eccentricity grows),
// then breadth-first by increasing (degree, index); reversed.  Returns pos:
// node -> position.
std::vector<int> rcm_positions(int n, int n_tris, const int *tris) {
  std::vector<std::vector<int>> adj(n);
  for (int e = 0; e < n_tris; ++e)
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        if (a != b) adj[tris[3 * e + a]].push_back(tris[3 * e + b]);
  for (auto &l : adj) {
    std::sort(l.begin(), l.end());
    l.erase(std::unique(l.begin(), l.end()), l.end());
  }
  auto deg = [&](int v) { return (int)adj[v].size(); };
  for (auto &l : adj)
    std::stable_sort(l.begin(), l.end(), [&](int a, int b) { return deg(a) < deg(b); });
  std::vector<int> order, level(n, -1);
  order.reserve(n);
  std::vector<char> done(n, 0);
  auto bfs_levels = [&](int s, std::vector<int> &comp) {  // returns eccentricity
    comp.clear();
    comp.push_back(s);
    level[s] = 0;
    for (size_t h = 0; h < comp.size(); ++h)
      for (int w : adj[comp[h]])
        if (level[w] < 0) {
          level[w] = level[comp[h]] + 1;
          comp.push_back(w);
        }
    return level[comp.back()];
  };
  std::vector<int> comp;
  for (int s0 = 0; s0 < n; ++s0) {
    if (done[s0]) continue;
    // start from the component's lowest-degree node
    int s = s0;
    bfs_levels(s, comp);
    for (int v : comp)
      if (deg(v) < deg(s)) s = v;
    for (int v : comp) level[v] = -1;
    int ecc = bfs_levels(s, comp);
    for (int it = 0; it < 8; ++it) {
      int best = -1;
      for (int v : comp)
        if (level[v] == ecc && (best < 0 || deg(v) < deg(best))) best = v;
      for (int v : comp) level[v] = -1;
      std::vector<int> c2;
      const int e2 = bfs_levels(best, c2);
      if (e2 <= ecc) {
        for (int v : c2) level[v] = -1;
        break;
      }
      s = best;
      ecc = e2;
      comp.swap(c2);
    }
    for (int v : comp) level[v] = -1;
    // Cuthill-McKee from s
    const size_t first = order.size();
    order.push_back(s);
    done[s] = 1;
    for (size_t h = first; h < order.size(); ++h)
      for (int w : adj[order[h]])
        if (!done[w]) {
          done[w] = 1;
          order.push_back(w);
        }
  }
  std::vector<int> pos(n);
  for (int i = 0; i < n; ++i) pos[order[n - 1 - i]] = i;
  return pos;
}

// Block bandwidth (64-row blocks, 2 rows per node) of a mesh's stiffness
// pattern under the node positions pos.
int block_bandwidth(int n_tris, const int *tris, const std::vector<int> &pos) {
  int bw = 0;
  for (int e = 0; e < n_tris; ++e)
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        bw = std::max(bw, std::abs(pos[tris[3 * e + a]] / (NB / 2) - pos[tris[3 * e + b]] / (NB / 2)));
  return bw;
}

// LU schedule: a list of stages (kernel kind, block step k, panel count),
// each with the exclusive prefix sums of its work items over the cells of a
// list (all cells batched, or one cell at a time).
enum StageKind { ST_DIAG, ST_TRSM, ST_STRIP, ST_GEMM1, ST_GEMM, ST_PANEL, ST_SWP, ST_STEP, ST_CELL };
struct Stage {
  int kind, k, np, off, tot;
  long long work;  // 64 x 64 x 64 block products (GEMM and strip stages)
};


struct LuPlan {
  int max_nblk = 0;
  int pivoting = 0;
  int *d_prefix = nullptr;
  std::vector<Stage> batched;
  std::vector<std::vector<Stage>> single;
  // triangular-solve item lists (one item per block row), static pivoting
  struct SolveLists {
    int fwd_off = 0, bwd_off = 0, tot = 0;
  };
  SolveLists solve_batched;
  std::vector<SolveLists> solve_single;
  // the batch split in groups of cells factored on concurrent streams (cells
  // dealt out by decreasing size so that every group holds big and small
  // ones); each group runs the same schedule over its own cells
  std::vector<std::vector<Stage>> grp;
  std::vector<int> grp_cells_off, grp_ncell;  // each group's cell list in d_prefix

  // trailing tiles of an mt x mt block matrix: all, or the lower triangle
  static long long tiles(const CellParams &p, long long mt) { return p.sym ? mt * (mt + 1) / 2 : mt * mt; }
  // Every trailing window is capped at the block bandwidth bw: with no
  // pivoting the factors of a band matrix keep the band, and panel p only
  // reaches blocks up to p + bw.
  static long long items(const CellParams &p, int kind, int k, int np, int pivoting) {
    const int nb = p.nblk;
    const int m = std::min(nb - k - 1, p.bw);
    switch (kind) {
      case ST_DIAG: return (k == 0 && nb > 0) ? 1 : 0;
      case ST_TRSM: return k < nb ? ((pivoting || p.sym) ? m : 2 * m) : 0;
      case ST_STRIP: {  // block column (and row) k + np, after panels k .. k+np-1
        const int ms = std::min(nb - k - np, p.bw);
        return ms >= 1 ? (p.sym ? ms : 2 * ms - 1) : 0;
      }
      case ST_GEMM1: return k < nb ? tiles(p, m) : 0;
      case ST_GEMM: return nb - k - np >= 1 ? tiles(p, std::min(nb - k - np, p.bw)) : 0;
      case ST_PANEL: return k < nb ? 1 : 0;
      case ST_STEP:  // fused block step k (k = -1 .. chain_steps - 2), see k_lu_step; np: the form
        return cell_step_items(p, k, step_flag(np));
      case ST_CELL: return nb > 0 ? 1 : 0;  // the whole factorization, one workgroup
      default: return (k < nb && m > 0) ? (m * NB + 255) / 256 : 0;
    }
  }
  // a block step's form (Stage::np): 1 L items, 2 panel items, 3 split
  // panel items; the kernel's panel argument: 0, 1, 2
  static int step_flag(int np) { return np == 2 ? 1 : np == 3 ? 2 : 0; }
  static long long chain_products(const Chain &v, int k) {
    const int s0 = k + 1;
    if (s0 >= v.nblk || s0 > v.elim) return 0;
    if (s0 == v.elim) return std::max(0, chain_items(v, k, 0) - 1);  // update-only: one product per block
    const int w = std::min(v.nblk - s0 - 1, v.bw);
    const long long npl = k >= 0 ? std::max(0, std::min(v.nblk - 1, k + v.bw) - s0) : 0;
    const long long upd =
        k >= 0 ? 1 + (v.sym ? 1 : 2) * std::min(w, v.bw - 1) + (v.sym ? npl * (npl + 1) / 2 : npl * npl) : 0;
    return upd + w + 1;
  }
  static long long block_products(const CellParams &p, int kind, int k, int np, int pivoting) {
    if (kind == ST_STRIP || kind == ST_GEMM || kind == ST_GEMM1) return (long long)np * items(p, kind, k, np, pivoting);
    if (kind == ST_STEP) {  // useful 64^3 products: updated blocks, L block solves, the inverse
      long long tot = chain_products(chain_of(p, 0), k);
      if (p.tw_m > 0) tot += chain_products(chain_of(p, 1), k);
      return tot;
    }
    if (kind == ST_CELL) {  // every block step: the first in the L-item form, the others in the panel form
      long long tot = 0;
      for (int kk = -1; kk < chain_steps(p) - 1; ++kk) tot += block_products(p, ST_STEP, kk, kk >= 0 ? 2 : 1, pivoting);
      return tot;
    }
    return 0;
  }
  // static pivoting: LU_DEPTH block steps per round -- panel k, then the
  // left-looking strips that bring block column k + j up to date with panels
  // k .. k+j-1 (factoring diagonal block k + j) and its block solve, then one
  // trailing update with all the round's panels; partial pivoting: one block
  // step per round
  // `depth` = panels per round of the static-pivoting schedule (delayed
  // update depth): the trailing matrix is read and written once per `depth`
  // block steps
  int depth = 4;
  bool fused = false;      // static pivoting, depth 1: one fused launch per block step
  bool cell_loop = false;  // ... or one launch for every step, one workgroup per cell
  std::vector<std::array<int, 3>> schedule(int nblk, int pivoting) const {
    std::vector<std::array<int, 3>> s;
    if (!pivoting && cell_loop) {
      if (nblk > 0) s.push_back({ST_CELL, -1, 0});
    } else if (!pivoting && fused) {
      for (int k = -1; k < nblk - 1; ++k) s.push_back({ST_STEP, k, 1});
    } else if (!pivoting) {
      s.push_back({ST_DIAG, 0, 0});
      for (int k = 0; k < nblk; k += depth) {
        const int R = std::min(depth, nblk - k);
        s.push_back({ST_TRSM, k, 0});
        for (int j = 1; j < R; ++j) {
          s.push_back({ST_STRIP, k, j});
          s.push_back({ST_TRSM, k + j, 0});
        }
        s.push_back({ST_GEMM, k, R});
      }
    } else {
      for (int k = 0; k < nblk; ++k) {
        s.push_back({ST_PANEL, k, 0});
        s.push_back({ST_SWP, k, 0});
        s.push_back({ST_TRSM, k, 0});
        s.push_back({ST_GEMM1, k, 1});
      }
    }
    return s;
  }

  // Fused block step lists: k_lu_step hands item i of a launch to XCD
  // chunk xcd_item^-1(i) (eight contiguous chunks, dispatched in order within
  // each); inside every chunk the diagonal and L items (the inverse on their
  // critical path) go first and the trailing tiles fill in after them.
  static void step_heavy_first(std::vector<int> &buf, const std::vector<CellParams> &cp, const Stage &g) {
    if (g.kind != ST_STEP) return;
    const int q = g.tot >> 3, r = g.tot & 7;
    for (int x = 0; x < 8; ++x) {
      const int b0 = g.off + x * q + std::min(x, r), nx = q + (x < r ? 1 : 0);
      std::stable_partition(buf.begin() + b0, buf.begin() + b0 + nx, [&](int v) {
        const CellParams &p = cp[v >> 20];
        int t = v & 0xFFFFF;
        Chain ch = chain_of(p, 0);
        if (p.tw_m > 0) {
          const int na = chain_items(ch, g.k, step_flag(g.np));
          if (t >= na) {
            ch = chain_of(p, 1);
            t -= na;
          }
        }
        if (g.k + 1 >= ch.elim) return false;  // update-only items
        const int wl = std::min(ch.nblk - g.k - 2, ch.bw);
        return t <= (g.np == 2 ? 0 : g.np == 3 ? std::max(wl - 1, 0) : wl);
      });
    }
  }

  // fused block steps (fuse = the fused_step option): 1 per launch, L items
  // while they are at most panel_min, else split panel items while those
  // are, else panel items (one inverse per cell: the least work); 2 L items
  // only; 3 panel items after the first step; 4 one launch for the whole
  // factorization (k_lu_cell: every cell's steps in one workgroup, largest
  // cells first); 5 split panel items (the first L block with the inverse,
  // the others as L items) in every launch with more than panel_min L items,
  // L items in the others
  void build(const std::vector<CellParams> &cp, int piv, int groups, int lu_depth, int fuse = 0, int panel_min = 512) {
    if (lu_depth < 1 || groups < 1) throw ApiError{WCFE_ERR_ARG, "LU schedule needs depth and groups >= 1"};
    pivoting = piv;
    depth = lu_depth;
    fused = fuse && !piv && lu_depth == 1;
    cell_loop = fused && fuse == 4;
    // the step's form: L items (np 1), one panel item per cell (np 2) or
    // split panel items (np 3)
    auto step_form = [&](const std::vector<int> &cells, std::array<int, 3> &st) {
      if (st[0] != ST_STEP || st[1] < 0 || fuse == 2) return;
      if (fuse == 3) {
        st[2] = 2;
        return;
      }
      long long tot = 0, split = 0;
      for (int c : cells) {
        tot += items(cp[c], ST_STEP, st[1], 1, pivoting);
        split += items(cp[c], ST_STEP, st[1], 3, pivoting);
      }
      if (tot > panel_min) st[2] = (fuse == 5 || split <= panel_min) ? 3 : 2;
    };
    std::vector<int> all((size_t)cp.size());
    for (size_t c = 0; c < cp.size(); ++c) all[c] = (int)c;
    const int n = (int)cp.size();
    std::vector<int> buf;
    max_nblk = 0;
    for (auto &p : cp) max_nblk = std::max(max_nblk, chain_steps(p));
    if (n >= (1 << 11)) throw ApiError{WCFE_ERR_ARG, "at most 2047 systems per batch"};
    batched.clear();
    std::vector<int> by_size(all);  // the cell launch takes the longest chains first
    std::stable_sort(by_size.begin(), by_size.end(),
                     [&](int a, int b) { return chain_steps(cp[a]) > chain_steps(cp[b]); });
    for (auto st : schedule(max_nblk, pivoting)) {
      step_form(all, st);
      Stage g{st[0], st[1], st[2], (int)buf.size(), 0, 0};
      for (int c : (st[0] == ST_CELL ? by_size : all)) {
        const long long t = items(cp[c], st[0], st[1], st[2], pivoting);
        if (t >= (1 << 20)) throw ApiError{WCFE_ERR_ARG, "too many LU work items per cell"};
        for (int i = 0; i < t; ++i) buf.push_back((c << 20) | i);
        g.tot += (int)t;
        g.work += block_products(cp[c], st[0], st[1], st[2], pivoting);
      }
      step_heavy_first(buf, cp, g);
      if (g.tot > 0) batched.push_back(g);
    }
    single.assign(n, {});
    for (int c = 0; c < n; ++c)
      for (auto st : schedule(chain_steps(cp[c]), pivoting)) {
        step_form({c}, st);
        const long long t = items(cp[c], st[0], st[1], st[2], pivoting);
        if (t == 0) continue;
        if (t >= (1 << 20)) throw ApiError{WCFE_ERR_ARG, "too many LU work items per cell"};
        single[c].push_back(Stage{st[0], st[1], st[2], (int)buf.size(), (int)t,
                                  block_products(cp[c], st[0], st[1], st[2], pivoting)});
        for (int i = 0; i < t; ++i) buf.push_back((c << 20) | i);
      }
    // solve lists: largest systems first (longest dependency chains), block
    // rows ascending (forward) / descending (backward) within a system
    // one item per pair of block rows (in the order the sweep solves them)
    auto add_lists = [&](const std::vector<int> &cl) {
      SolveLists L;
      L.fwd_off = (int)buf.size();
      for (int c : cl)
        for (int pi = 0; 2 * pi < cp[c].nblk; ++pi) buf.push_back((c << 20) | pi);
      L.tot = (int)buf.size() - L.fwd_off;
      L.bwd_off = (int)buf.size();
      for (int c : cl)
        for (int pi = 0; 2 * pi < cp[c].nblk; ++pi) buf.push_back((c << 20) | pi);
      return L;
    };
    std::vector<int> order(n);
    for (int c = 0; c < n; ++c) order[c] = c;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cp[a].nblk > cp[b].nblk; });
    solve_batched = add_lists(order);
    // concurrent groups
    grp.clear();
    grp_cells_off.clear();
    grp_ncell.clear();
    const int G = std::min(std::max(groups, 1), n);
    if (G > 1) {
      for (int g = 0; g < G; ++g) {
        std::vector<int> mine;
        for (int q = g; q < n; q += G) mine.push_back(order[q]);
        int gmax = 0;
        for (int c : mine) gmax = std::max(gmax, chain_steps(cp[c]));
        std::vector<Stage> st;
        for (auto sk : schedule(gmax, pivoting)) {
          step_form(mine, sk);
          Stage gs{sk[0], sk[1], sk[2], (int)buf.size(), 0, 0};
          for (int c : mine) {
            const long long t = items(cp[c], sk[0], sk[1], sk[2], pivoting);
            for (int i = 0; i < t; ++i) buf.push_back((c << 20) | i);
            gs.tot += (int)t;
            gs.work += block_products(cp[c], sk[0], sk[1], sk[2], pivoting);
          }
          step_heavy_first(buf, cp, gs);
          if (gs.tot > 0) st.push_back(gs);
        }
        grp.push_back(st);
        grp_cells_off.push_back((int)buf.size());
        grp_ncell.push_back((int)mine.size());
        buf.insert(buf.end(), mine.begin(), mine.end());
      }
    }
    solve_single.assign(n, SolveLists{});
    for (int c = 0; c < n; ++c) solve_single[c] = add_lists(std::vector<int>{c});
    d_prefix = dupload(buf);
  }
};

// A batch of cells laid out in HBM.
struct Batch {
  int ncell = 0, N = 0, T = 0, S = 0, ncolor = 0, max_npad = 0, any_ul = 0, any_bell = 0, any_csr = 0;
  long long kpool = 0;
  std::vector<CellParams> cp_h;
  std::vector<int> color_start;                 // global colour-major ranges (ncolor + 1)
  std::vector<std::vector<int>> cell_color;     // per cell: ncolor + 1 global offsets
  std::vector<std::vector<int>> elem_pos;       // per cell: original element -> global id
  std::vector<std::vector<int>> elem_color;     // per cell: colour of each original element
  std::vector<std::vector<int>> spr_pos;        // per cell: original spring -> global id
  std::vector<int> ncolors_cell;
  // device
  double2 *X = nullptr, *x = nullptr, *fau = nullptr;
  double *ea = nullptr, *cC = nullptr;
  int *el_tri = nullptr, *el_cell = nullptr, *el_spr = nullptr;
  double *sp_bary = nullptr, *sp_anchor = nullptr, *sp_state = nullptr;
  int *sp_cell = nullptr, *sp_elem = nullptr, *sp_local = nullptr;
  CellParams *cp = nullptr;
  CellState *cs = nullptr;
  int *node_cell = nullptr, *adj_start = nullptr, *adj = nullptr;
  double *Ge = nullptr;  // per element: residual (6) for the node gather
  // node gather plan: per node, its column nodes (groups) and, per group, the
  // element contributions in element order (offsets of the 2 x 2 block in Ke64)
  int *g_start = nullptr, *g_col = nullptr, *g_bpos = nullptr, *g_cstart = nullptr, *g_ke = nullptr, *g_node = nullptr;
  std::vector<int> g_start_h;  // host copy of g_start
  double *Ke64 = nullptr, *G = nullptr, *r = nullptr, *delta = nullptr, *Rea = nullptr, *MM = nullptr;
  void *K = nullptr;
  void *Dinv = nullptr;
  void *ysol = nullptr;  // triangular-solve intermediate, one per system row
  int *sync = nullptr;   // triangular-solve tickets and per-cell counters
  int nsync = 0;
  int *ipiv = nullptr;
  int *cells = nullptr;
  int *replica = nullptr;
  int *transitions = nullptr;
  LuPlan plan;
  // sparse (block-CSR + PCG) cells
  std::vector<int> csr_cells_h;                 // ids of the sparse cells
  std::vector<int> chunk_off_h, nchunk_h;       // per cell: its chunks (sparse cells)
  std::vector<long long> blk_begin_h, blk_end_h;// per cell: its block range (sparse cells)
  int nchunk = 0;
  long long nnzb = 0;
  int *row_ptr = nullptr, *bcol = nullptr, *diag_pos = nullptr, *el_bpos = nullptr, *csr_cells = nullptr;
  int4 *chunks = nullptr;
  void *Kb = nullptr, *cgvec = nullptr;         // block values; x, r, z, p, q, s (2N each) and dinv (4N)
  CgState *cg = nullptr;
  double *partials = nullptr;
  // per-cell norms of cells above NORM_CHUNK entries (chunk partials, counters)
  double *norm_part = nullptr;
  unsigned *norm_cnt = nullptr;
  int norm_stride = 1;
  NormWork norm_work() const {
    NormWork w;
    if (norm_part) {
      w.part = norm_part;
      w.cnt = norm_cnt;
      w.stride = norm_stride;
    }
    return w;
  }
  // dense numbering: per node its position in the cell's dense system; the
  // permuted right-hand side and correction of the triangular solves
  int *kpos = nullptr;
  // the first solve's forward sweep, run inside the fused block steps: per
  // band row its batch dof (-1: padding) and the forward result
  int *kinv = nullptr;
  float *yfwd = nullptr;
  // batches with two-chain cells: the sweeps' forward result (sys elements)
  void *ysw = nullptr;
  std::vector<std::vector<int>> kpos_h;
  double *rp = nullptr, *dp = nullptr;
  bool solve_wg = false;  // static pivoting: one workgroup per cell walks its band (narrow bands)
  // PCG iterations the batch's slowest sparse cell took in refinement sweep
  // i of the last Newton step (-1: none yet): the first batch of sweep i's
  // next solve
  int pcg_hint[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
  int lu_depth = 4, lu_groups = 4;  // the LU schedule in effect (options, or their automatic choice)
  int max_bw = 0;                    // widest block band of the dense cells
  // boundary and body forces (cells with bterm != 0)
  int any_bterm = 0, any_area = 0, nbnd = 0;
  int *bt_bidx = nullptr, *bt_node = nullptr, *bt_cst = nullptr, *crange = nullptr;
  int2 *bt_cand = nullptr, *bt_frame = nullptr;
  double2 *bt_g = nullptr;
  double *bt_J = nullptr, *bt_area = nullptr, *bt_gd = nullptr, *bt_fext = nullptr;

  void release() {
    void *ptrs[] = {X, x, fau, ea, cC, el_tri, el_cell, el_spr, sp_bary, sp_anchor, sp_state, sp_cell, sp_elem,
                    sp_local, cp, cs, node_cell, adj_start, adj, Ge, g_start, g_col, g_bpos, g_cstart, g_ke, g_node, Ke64, G, r, delta, Rea, MM, K, Dinv, ysol, sync, ipiv, cells,
                    replica, transitions, plan.d_prefix, row_ptr, bcol, diag_pos, el_bpos, csr_cells, chunks, Kb,
                    cgvec, cg, partials, norm_part, norm_cnt, kpos, kinv, yfwd, ysw, rp, dp, bt_bidx, bt_node, bt_cst, crange, bt_cand, bt_frame, bt_g, bt_J, bt_area,
                    bt_gd, bt_fext};
    for (void *p : ptrs)
      if (p) (void)hipFree(p);
    *this = Batch();
  }
};

// Per-phase kernel timing with HIP events on the library's stream (only
// while wcfe_profile_step runs; never inside a captured graph).
struct Profiler {
  struct Rec {
    int phase;
    hipEvent_t a, b;
    long long work;
  };
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  std::vector<Rec> recs;
  hipEvent_t get() {
    if (used == pool.size()) {
      hipEvent_t e;
      HIPCHECK(hipEventCreate(&e));
      pool.push_back(e);
    }
    return pool[used++];
  }
  ~Profiler() {
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
  }
};

}  // namespace

// One captured frame: the batch's committed state copied to pinned host
// memory on the copy stream.  Layout (doubles): X[2N] ea[3N] cC[N] fau[2N],
// then the CellState array.
struct FrameSlot {
  int id = -1, layout = -1;
  size_t bytes = 0;
  double *dev = nullptr, *host = nullptr;
  hipEvent_t done = nullptr;
};

struct wcfe_ctx {
  wcfe_options opt;
  std::string err;
  int device = 0;
  hipStream_t stream = nullptr;
  std::string arch;
  size_t max_lds = 65536;
  int n_cu = 256;  // compute units of the device
  std::map<int, Table> tables;
  std::vector<HostCell> cells;
  bool finalized = false;
  Batch b;
  double *tab_px = nullptr, *tab_py = nullptr, *tab_k = nullptr;
  std::map<int, std::pair<int, int>> table_off;  // id -> (offset, n)
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  Profiler *prof = nullptr;
  int cg_iters_max = 0, cg_breakdowns = 0;  // sparse cells, since the last step began
  bool stepped = false;                      // a Newton step has run since the batch was built
  std::vector<hipStream_t> side;            // concurrent LU groups
  hipEvent_t fork_ev = nullptr;
  std::vector<hipEvent_t> join_ev;
  // frame output: two slots in flight, copied on their own stream
  FrameSlot frames[2];
  int next_frame = 0;
  int layout = 0;  // bumped whenever the batch is laid out (finalize, rebuild)
  hipStream_t copy_stream = nullptr;
  hipEvent_t frame_ready = nullptr;
};

namespace {

int set_err(wcfe_ctx *ctx, int code, const std::string &m) {
  if (ctx) ctx->err = m;
  return code;
}

template <class F> int guarded(wcfe_ctx *ctx, F &&f) {
  try {
    f();
    return WCFE_OK;
  } catch (const ApiError &e) {
    return set_err(ctx, e.code, e.msg);
  } catch (const HipError &e) {
    return set_err(ctx, WCFE_ERR_HIP, std::string(e.what) + ": " + hipGetErrorString(e.e));
  } catch (const std::bad_alloc &) {
    return set_err(ctx, WCFE_ERR_OOM, "host allocation failed");
  }
}

bool mixed(const wcfe_ctx *ctx) { return ctx->opt.precision == WCFE_PREC_MIXED; }

struct PhaseScope {
  wcfe_ctx *ctx;
  int phase;
  long long work;
  hipEvent_t a = nullptr;
  PhaseScope(wcfe_ctx *c, int ph, long long w = 0) : ctx(c), phase(ph), work(w) {
    if (ctx->prof) {
      a = ctx->prof->get();
      HIPCHECK(hipEventRecord(a, ctx->stream));
    }
  }
  ~PhaseScope() {
    if (ctx->prof) {
      hipEvent_t b = ctx->prof->get();
      if (hipEventRecord(b, ctx->stream) == hipSuccess) ctx->prof->recs.push_back({phase, a, b, work});
    }
  }
};

// Timing of one kernel while wcfe_profile_step runs: events handed to the
// launch mark the kernel's own start and end (not the queue around it).
struct KernelTimer {
  wcfe_ctx *ctx;
  int phase;
  long long work;
  hipEvent_t a = nullptr, b = nullptr;
  KernelTimer(wcfe_ctx *c, int ph, long long w) : ctx(c), phase(ph), work(w) {
    if (ctx->prof) {
      a = ctx->prof->get();
      b = ctx->prof->get();
    }
  }
  ~KernelTimer() {
    if (ctx->prof && a) ctx->prof->recs.push_back({phase, a, b, work});
  }
};

View make_view(const wcfe_ctx *ctx, const Batch &b) {
  View v;
  v.X = b.X; v.x = b.x; v.ea = b.ea; v.cC = b.cC; v.fau = b.fau;
  v.el_tri = b.el_tri; v.el_cell = b.el_cell; v.el_spr = b.el_spr; v.el_bpos = b.el_bpos; v.kpos = b.kpos;
  v.sp_bary = b.sp_bary; v.sp_anchor = b.sp_anchor; v.sp_state = b.sp_state;
  v.cp = b.cp; v.cs = b.cs;
  v.tab_px = ctx->tab_px; v.tab_py = ctx->tab_py; v.tab_k = ctx->tab_k;
  v.kappa = ctx->opt.kappa;
  return v;
}

// Lay out cells in HBM.  Elements are stored colour-major (colour k of every
// cell, then colour k+1, ...), so one colour of the whole batch, or of one
// cell, is a contiguous range.
// f(c) for every cell on up to 16 host threads (cells dealt out in turn);
// each f(c) writes only its own cell's parts of the outputs.
template <class F> void parallel_cells(int ncell, F f) {
  const int nth = std::max(1, std::min<int>({16, ncell, (int)std::thread::hardware_concurrency()}));
  if (nth <= 1) {
    for (int c = 0; c < ncell; ++c) f(c);
    return;
  }
  std::atomic<int> next(0);
  std::vector<std::thread> pool;
  for (int t = 0; t < nth; ++t)
    pool.emplace_back([&]() {
      for (int c = next++; c < ncell; c = next++) f(c);
    });
  for (auto &th : pool) th.join();
}

void build_batch(wcfe_ctx *ctx, Batch &b, const std::vector<HostCell> &hc, int elem_size) {
  b.ncell = (int)hc.size();
  std::vector<std::vector<int>> colors(b.ncell);
  b.ncolors_cell.assign(b.ncell, 0);
  b.ncolor = 0;
  for (int c = 0; c < b.ncell; ++c) {
    const int nc = color_elements(hc[c].n_nodes, hc[c].n_tris, hc[c].tris.data(), colors[c]);
    if (nc < 0) fail(WCFE_ERR_ARG, "cell %d needs more than 64 element colours", c);
    b.ncolors_cell[c] = nc;
    b.ncolor = std::max(b.ncolor, nc);
  }
  b.cp_h.assign(b.ncell, CellParams{});
  int N = 0, S = 0, sys = 0;
  long long kp = 0;
  b.max_npad = 0;
  b.any_ul = b.any_bell = 0;
  for (int c = 0; c < b.ncell; ++c) {
    const HostCell &h = hc[c];
    CellParams &p = b.cp_h[c];
    p.law = h.law;
    p.node_springs = h.node_springs;
    p.spring_law = h.spring_law;
    p.kinetics = h.kinetics;
    if (h.spring_law == WCFE_SPRING_TABLE) {
      auto it = ctx->table_off.find(h.table_id);
      if (it == ctx->table_off.end()) fail(WCFE_ERR_ARG, "cell %d uses undefined table %d", c, h.table_id);
      p.table_off = it->second.first;
      p.table_n = it->second.second;
    }
    p.node_off = N;
    p.n_nodes = h.n_nodes;
    p.n_elems = h.n_tris;
    p.spr_off = S;
    p.n_spr = h.n_spr;
    p.ndof = 2 * h.n_nodes;
    // the block-CSR solver is block-Jacobi conjugate gradients on the sparse
    // pattern: the bending tangent is not symmetric and the area term's has a
    // dense rank-one part over the boundary, so those cells take the dense LU
    const bool dense_only = h.k_bend != 0 || h.k_area != 0;
    p.csr = h.solver == WCFE_SOLVER_CSR ||
            (h.solver == WCFE_SOLVER_AUTO && p.ndof > ctx->opt.dense_max_dof && !dense_only);
    if (p.csr && dense_only)
      fail(WCFE_ERR_ARG,
           "cell %d: the bending and area-constraint tangents need the dense solver (block-CSR runs conjugate "
           "gradients on the sparse pattern); use WCFE_SOLVER_DENSE or AUTO for cells with k_bend or k_area > 0",
           c);
    p.npad = p.csr ? 0 : ((p.ndof + NB - 1) / NB) * NB;
    p.nblk = p.npad / NB;
    p.sys_off = sys;
    p.vec_off = 2 * N;
    p.bterm = (h.k_bend != 0 ? BT_BEND : 0) | (h.k_area != 0 ? BT_AREA : 0) | (h.ap ? BT_AP : 0) |
              (h.drag != 0 ? BT_DRAG : 0) | (!h.f_ext.empty() ? BT_EXT : 0);
    p.passive = h.passive;
    p.k_bend = h.k_bend; p.k_area = h.k_area; p.area_0 = h.area_0; p.drag = h.drag;
    p.ap_rate = h.ap_rate; p.ap_quantum = h.ap_quantum; p.ap_density = h.ap_density;
    // the tangent is symmetric unless node-attached springs (area derivative)
    // or the boundary bending / area terms are present
    p.sym = ctx->opt.symmetric && ctx->opt.pivoting == WCFE_PIVOT_STATIC && !h.node_springs &&
            !(p.bterm & (BT_BEND | BT_AREA));
    p.k_off = kp;
    p.mu = h.mu; p.E = h.E; p.nu = h.nu; p.thick = h.thick; p.k_lin = h.k_lin; p.rho_max = h.rho_max;
    N += h.n_nodes;
    S += h.n_spr;
    sys += p.npad;
    kp += (long long)p.npad * p.npad;
    b.max_npad = std::max(b.max_npad, p.npad);
    b.any_ul |= (h.law == WCFE_LAW_UL);
    b.any_bterm |= p.bterm != 0;
    b.any_area |= (p.bterm & BT_AREA) != 0;
    b.any_bell |= (h.kinetics == WCFE_KIN_BELL);
    b.any_csr |= p.csr;
  }
  b.N = N;
  b.S = S;
  b.kpool = kp;
  // dense cells under static pivoting: a band-reducing node numbering (RCM
  // when it narrows the band) and the LU restricted to the band; partial
  // pivoting (row swaps widen the band), the area term (rank-one coupling of
  // the whole boundary) and ctx->opt.band == 0 keep the natural numbering and
  // the full matrix
  b.kpos_h.assign(b.ncell, {});
  parallel_cells(b.ncell, [&](int c) {
    CellParams &p = b.cp_h[c];
    const HostCell &h = hc[c];
    std::vector<int> ident(h.n_nodes);
    for (int i = 0; i < h.n_nodes; ++i) ident[i] = i;
    const int full = std::max(p.nblk - 1, 0);
    p.bw = full;
    b.kpos_h[c] = ident;
    if (!p.csr && ctx->opt.band && ctx->opt.pivoting == WCFE_PIVOT_STATIC && !(p.bterm & BT_AREA) && p.nblk > 1) {
      if (h.bw_rcm < 0 || h.rcm_hash != h.topo_hash || (int)h.rcm_pos.size() != h.n_nodes) {  // once per mesh
        h.bw_nat = block_bandwidth(h.n_tris, h.tris.data(), ident);
        h.rcm_pos = rcm_positions(h.n_nodes, h.n_tris, h.tris.data());
        h.bw_rcm = block_bandwidth(h.n_tris, h.tris.data(), h.rcm_pos);
        h.rcm_hash = h.topo_hash;
      }
      if (h.bw_rcm < h.bw_nat) {
        b.kpos_h[c] = h.rcm_pos;
        p.bw = std::min(full, h.bw_rcm);
      } else {
        p.bw = std::min(full, h.bw_nat);
      }
    }
  });
  // automatic LU schedule: narrow bands are latency bound (one block step per
  // round, one stream); wide or full matrices are throughput bound (delayed
  // updates over 4 block steps, 4 concurrent cell groups)
  bool narrow = true;
  for (const CellParams &p : b.cp_h)
    if (!p.csr && p.bw > SOLVE_WG_MAX_BW) narrow = false;
  b.lu_depth = ctx->opt.lu_depth > 0 ? ctx->opt.lu_depth : (narrow ? 1 : 4);
  b.lu_groups = ctx->opt.lu_groups > 0 ? ctx->opt.lu_groups : (narrow ? 1 : AUTO_GROUPS_MAX);
  // narrow bands: one workgroup per cell runs each triangular sweep (no
  // cross-workgroup hops); otherwise the block rows of all cells are spread
  // over workgroups (k_lu_solve_mw)
  b.solve_wg = ctx->opt.pivoting == WCFE_PIVOT_STATIC && narrow && b.max_npad > 0 &&
               lu_solve_lds_bytes(b.max_npad, elem_size) <= lu_solve_max_dynamic(ctx->max_lds);
  b.max_bw = 0;
  for (CellParams &p : b.cp_h) {
    const int full = std::max(p.nblk - 1, 0);
    // the delayed update reads panel blocks up to depth - 1 past the band; the
    // spread-out solve reads one block past it
    p.env = std::min(full, p.bw + std::max(b.lu_depth - 1, b.solve_wg ? 0 : 1));
    if (!p.csr) b.max_bw = std::max(b.max_bw, p.bw);
  }
  // two-chain band LU: the batch's chain of block steps is its longest cell's;
  // a cell of n blocks and block bandwidth bw factored as two chains takes
  // m + bw steps (m = n - bw - nb top blocks, nb = (n - 2 bw) / 2 bottom
  // blocks: nb <= m - bw, so the chains never update the middle window in the
  // same step).  Only cells longer than the shortest achievable batch chain
  // are split (the others keep their factorization bit for bit).
  if (elem_size == 4 && ctx->opt.twist && ctx->opt.pivoting == WCFE_PIVOT_STATIC && b.lu_depth == 1 &&
      b.solve_wg && ctx->opt.fused_step >= 1 && b.max_bw <= 4) {
    auto split = [](const CellParams &p, int &m, int &nb) {
      if (p.csr || p.bw < 1 || p.bw >= p.nblk - 1 || p.nblk - 2 * p.bw < 2) return false;
      nb = (p.nblk - 2 * p.bw) / 2;
      m = p.nblk - p.bw - nb;
      return m + p.bw < p.nblk;
    };
    int chain = 0;
    double work = 0.0;  // block products of the whole factorization, about nblk (bw+1)(bw+2)/2 per cell
    for (const CellParams &p : b.cp_h) {
      if (p.csr) continue;
      int m, nb;
      chain = std::max(chain, split(p, m, nb) ? m + p.bw : p.nblk);
      work += (double)p.nblk * (p.bw + 1) * (p.sym ? (p.bw + 2) * 0.5 : (p.bw + 1));
    }
    // automatic (twist 1): only when the shorter chain's launches stay
    // latency bound -- at most 8 block products per compute unit and step;
    // a batch whose launches are throughput bound gains nothing from fewer,
    // fuller launches
    if (ctx->opt.twist == 1 && work > 8.0 * ctx->n_cu * chain) chain = INT_MAX;
    for (CellParams &p : b.cp_h) {
      int m, nb;
      if (p.nblk > chain && split(p, m, nb)) {
        p.tw_m = m;
        p.tw_nb = nb;
        // automatic schedule: two concurrent cell groups, so one group's
        // latency-bound launches overlap the other's fuller ones
        if (ctx->opt.lu_groups == 0) b.lu_groups = 2;
      }
    }
  }
  // colour-major element order
  b.color_start.assign(b.ncolor + 1, 0);
  b.cell_color.assign(b.ncell, std::vector<int>(b.ncolor + 1, 0));
  b.elem_pos.assign(b.ncell, {});
  b.elem_color = colors;
  int T = 0;
  for (int c = 0; c < b.ncell; ++c) {
    T += hc[c].n_tris;
    b.elem_pos[c].assign(hc[c].n_tris, -1);
  }
  b.T = T;
  std::vector<int> el_tri(3 * (size_t)T), el_cell(T);
  // elements per (cell, colour), then the colour-major positions: colour k
  // of cell 0, of cell 1, ..., then colour k + 1; in increasing element
  // order inside each (cell, colour)
  std::vector<std::vector<int>> cnt(b.ncell, std::vector<int>(b.ncolor, 0));
  for (int c = 0; c < b.ncell; ++c)
    for (int e = 0; e < hc[c].n_tris; ++e) cnt[c][colors[c][e]]++;
  std::vector<std::vector<int>> ranges(b.ncell, std::vector<int>(2 * b.ncolor, 0));
  int g = 0;
  for (int k = 0; k < b.ncolor; ++k) {
    b.color_start[k] = g;
    for (int c = 0; c < b.ncell; ++c) {
      ranges[c][2 * k] = g;
      g += cnt[c][k];
      ranges[c][2 * k + 1] = g;
    }
  }
  b.color_start[b.ncolor] = g;
  for (int c = 0; c < b.ncell; ++c) {
    const HostCell &h = hc[c];
    std::vector<int> next(b.ncolor);
    for (int k = 0; k < b.ncolor; ++k) next[k] = ranges[c][2 * k];
    for (int e = 0; e < h.n_tris; ++e) {
      const int pos = next[colors[c][e]]++;
      b.elem_pos[c][e] = pos;
      for (int a = 0; a < 3; ++a) el_tri[3 * (size_t)pos + a] = b.cp_h[c].node_off + h.tris[3 * e + a];
      el_cell[pos] = c;
    }
  }
  b.cell_color = ranges;  // 2 entries (begin, end) per colour
  // springs: per cell, ordered by host element global id, then original index
  std::vector<int> el_spr(2 * (size_t)T, 0), sp_cell(S), sp_elem(S), sp_local(S);
  std::vector<double> sp_bary(3 * (size_t)S), sp_anchor(2 * (size_t)S), sp_state(S);
  b.spr_pos.assign(b.ncell, {});
  parallel_cells(b.ncell, [&](int c) {  // disjoint ranges per cell (its springs and elements)
    const HostCell &h = hc[c];
    std::vector<int> order(h.n_spr);
    for (int s = 0; s < h.n_spr; ++s) order[s] = s;
    std::stable_sort(order.begin(), order.end(), [&](int a, int bb) {
      return b.elem_pos[c][h.spr_tri[a]] < b.elem_pos[c][h.spr_tri[bb]];
    });
    b.spr_pos[c].assign(h.n_spr, 0);
    int base = b.cp_h[c].spr_off;
    for (int i = 0; i < h.n_spr; ++i) {
      const int s = order[i], gs = base + i;
      const int ge = b.elem_pos[c][h.spr_tri[s]];
      b.spr_pos[c][s] = gs;
      sp_cell[gs] = c;
      sp_elem[gs] = ge;
      sp_local[gs] = s;
      for (int q = 0; q < 3; ++q) sp_bary[3 * (size_t)gs + q] = h.spr_bary[3 * s + q];
      for (int q = 0; q < 2; ++q) sp_anchor[2 * (size_t)gs + q] = h.spr_anchor[2 * s + q];
      sp_state[gs] = h.spr_state.empty() ? 1.0 : h.spr_state[s];
      if (el_spr[2 * (size_t)ge + 1] == 0 && el_spr[2 * (size_t)ge] == 0) el_spr[2 * (size_t)ge] = gs;
      el_spr[2 * (size_t)ge + 1] = gs + 1;
    }
  });
  // node -> (element, corner) adjacency in increasing element order
  std::vector<int> adj_start(N + 1, 0), adj(3 * (size_t)T), node_cell(N);
  for (int e = 0; e < T; ++e)
    for (int a = 0; a < 3; ++a) adj_start[el_tri[3 * (size_t)e + a] + 1]++;
  for (int i = 0; i < N; ++i) adj_start[i + 1] += adj_start[i];
  {
    std::vector<int> fill(adj_start.begin(), adj_start.end() - 1);
    for (int e = 0; e < T; ++e)
      for (int a = 0; a < 3; ++a) adj[fill[el_tri[3 * (size_t)e + a]]++] = 4 * e + a;
  }
  for (int c = 0; c < b.ncell; ++c)
    for (int i = 0; i < hc[c].n_nodes; ++i) node_cell[b.cp_h[c].node_off + i] = c;
  // node state
  std::vector<double2> X(N), x(N), fau(N);
  std::vector<double> ea(3 * (size_t)N, 0.0), cC(N, 0.0);
  for (int c = 0; c < b.ncell; ++c) {
    const HostCell &h = hc[c];
    const int o = b.cp_h[c].node_off;
    for (int i = 0; i < h.n_nodes; ++i) {
      X[o + i] = make_double2(h.X[2 * i], h.X[2 * i + 1]);
      x[o + i] = h.x.empty() ? X[o + i] : make_double2(h.x[2 * i], h.x[2 * i + 1]);
      fau[o + i] = h.fau.empty() ? make_double2(0, 0) : make_double2(h.fau[2 * i], h.fau[2 * i + 1]);
      if (!h.ea.empty())
        for (int q = 0; q < 3; ++q) ea[3 * (size_t)(o + i) + q] = h.ea[3 * i + q];
      if (!h.cC.empty()) cC[o + i] = h.cC[i];
    }
  }
  std::vector<CellState> cs(b.ncell);
  for (int c = 0; c < b.ncell; ++c) {
    CellState &s = cs[c];
    memset(&s, 0, sizeof s);
    if (hc[c].has_state) s = hc[c].state;
    s.time = std::isnan(hc[c].time) ? ctx->opt.t0 : hc[c].time;
    s.step = hc[c].step;
  }
  std::vector<int> ident(b.ncell), replica(b.ncell);
  for (int c = 0; c < b.ncell; ++c) ident[c] = replica[c] = c;
  // boundary tables of the cells with boundary forces (global node ids)
  std::vector<int> bidx, bnode, bcst, crange;
  std::vector<int2> bcand;
  if (b.any_bterm) {
    bidx.assign(N, -1);
    bcst.push_back(0);
    for (int c = 0; c < b.ncell; ++c) {
      CellParams &p = b.cp_h[c];
      p.bnd_off = (int)bnode.size();
      if (!(p.bterm & (BT_BEND | BT_AREA | BT_AP))) continue;
      std::vector<int> cst;
      std::vector<int2> cand;
      boundary_candidates(hc[c].n_nodes, hc[c].n_tris, hc[c].tris.data(), cst, cand);
      for (int i = 0; i < hc[c].n_nodes; ++i) {
        if (cst[i + 1] == cst[i]) continue;
        bidx[p.node_off + i] = (int)bnode.size();
        bnode.push_back(p.node_off + i);
        for (int k = cst[i]; k < cst[i + 1]; ++k)
          bcand.push_back(make_int2(p.node_off + cand[k].x, p.node_off + cand[k].y));
        bcst.push_back((int)bcand.size());
      }
      p.n_bnd = (int)bnode.size() - p.bnd_off;
    }
    b.nbnd = (int)bnode.size();
    crange.reserve((size_t)b.ncell * 2 * b.ncolor);
    for (int c = 0; c < b.ncell; ++c) crange.insert(crange.end(), b.cell_color[c].begin(), b.cell_color[c].end());
  }
  // upload
  b.X = dupload(X);
  b.x = dupload(x);
  b.fau = dupload(fau);
  b.ea = dupload(ea);
  b.cC = dupload(cC);
  b.el_tri = dupload(el_tri);
  b.el_cell = dupload(el_cell);
  b.el_spr = dupload(el_spr);
  b.sp_bary = dupload(sp_bary);
  b.sp_anchor = dupload(sp_anchor);
  b.sp_state = dupload(sp_state);
  b.sp_cell = dupload(sp_cell);
  b.sp_elem = dupload(sp_elem);
  b.sp_local = dupload(sp_local);
  b.cp = dupload(b.cp_h);
  b.cs = dupload(cs);
  b.node_cell = dupload(node_cell);
  b.adj_start = dupload(adj_start);
  b.adj = dupload(adj);
  b.cells = dupload(ident);
  b.replica = dupload(replica);
  {
    std::vector<int> kpos(N);
    for (int c = 0; c < b.ncell; ++c)
      for (int i = 0; i < hc[c].n_nodes; ++i) kpos[b.cp_h[c].node_off + i] = b.kpos_h[c][i];
    b.kpos = dupload(kpos);
  }
  b.rp = dalloc<double>(2 * (size_t)N);
  b.dp = dalloc<double>(2 * (size_t)N);
  HIPCHECK(hipMemset(b.rp, 0, sizeof(double) * 2 * N));
  HIPCHECK(hipMemset(b.dp, 0, sizeof(double) * 2 * N));
  if (b.any_bterm) {
    b.bt_bidx = dupload(bidx);
    b.bt_node = dupload(bnode);
    b.bt_cst = dupload(bcst);
    b.bt_cand = dupload(bcand);
    b.crange = dupload(crange);
    b.bt_frame = dalloc<int2>(std::max(b.nbnd, 1));
    b.bt_g = dalloc<double2>(std::max(b.nbnd, 1));
    b.bt_J = dalloc<double>(12 * (size_t)std::max(b.nbnd, 1));
    HIPCHECK(hipMemset(b.bt_frame, 0xFF, sizeof(int2) * std::max(b.nbnd, 1)));
    HIPCHECK(hipMemset(b.bt_g, 0, sizeof(double2) * std::max(b.nbnd, 1)));
    HIPCHECK(hipMemset(b.bt_J, 0, sizeof(double) * 12 * std::max(b.nbnd, 1)));
    b.bt_area = dalloc<double>(b.ncell);
    b.bt_gd = dalloc<double>(b.ncell);
    HIPCHECK(hipMemset(b.bt_area, 0, sizeof(double) * b.ncell));
    HIPCHECK(hipMemset(b.bt_gd, 0, sizeof(double) * b.ncell));
    bool any_ext = false;
    for (int c = 0; c < b.ncell; ++c) any_ext |= !hc[c].f_ext.empty();
    if (any_ext) {
      std::vector<double> fext(2 * (size_t)N, 0.0);
      for (int c = 0; c < b.ncell; ++c)
        if (!hc[c].f_ext.empty())
          std::copy(hc[c].f_ext.begin(), hc[c].f_ext.end(), fext.begin() + 2LL * b.cp_h[c].node_off);
      b.bt_fext = dupload(fext);
    }
  }
  b.transitions = dalloc<int>(1);
  HIPCHECK(hipMemset(b.transitions, 0, sizeof(int)));
  b.Ke64 = dalloc<double>(36 * (size_t)T);
  b.Ge = dalloc<double>(6 * (size_t)T);
  b.G = dalloc<double>(2 * (size_t)N);
  b.r = dalloc<double>(2 * (size_t)N);
  b.delta = dalloc<double>(2 * (size_t)N);
  b.Rea = dalloc<double>(3 * (size_t)N);
  b.MM = dalloc<double>(N);
  b.ipiv = dalloc<int>(std::max(sys, 1));
  HIPCHECK(hipMemset(b.ipiv, 0, sizeof(int) * std::max(sys, 1)));
  b.K = (void *)dalloc<unsigned char>((size_t)kp * elem_size);
  b.Dinv = (void *)dalloc<unsigned char>((size_t)std::max(sys, 1) * 128 * elem_size);
  {
    std::vector<int> kinv(std::max(sys, 1), -1);
    for (int c = 0; c < b.ncell; ++c) {
      const CellParams &p = b.cp_h[c];
      if (p.csr) continue;
      for (int i = 0; i < p.n_nodes; ++i)
        for (int d = 0; d < 2; ++d) kinv[p.sys_off + 2 * b.kpos_h[c][i] + d] = 2 * (p.node_off + i) + d;
    }
    b.kinv = dupload(kinv);
    b.yfwd = dalloc<float>(std::max(sys, 1));
    bool twist = false;
    for (const CellParams &p : b.cp_h) twist |= p.tw_m > 0;
    if (twist) b.ysw = (void *)dalloc<unsigned char>((size_t)std::max(sys, 1) * elem_size);
  }
  const size_t yb = lu_solve_exchange_bytes(std::max(sys, 1), elem_size);
  b.ysol = (void *)dalloc<unsigned char>(yb);
  HIPCHECK(hipMemset(b.ysol, 0, yb));
  b.nsync = SOLVE_SYNC_INTS;
  b.sync = dalloc<int>(b.nsync);
  HIPCHECK(hipMemset(b.sync, 0, sizeof(int) * b.nsync));
  {
    int max_ndof = 0;
    for (int c = 0; c < b.ncell; ++c) max_ndof = std::max(max_ndof, b.cp_h[c].ndof);
    if (max_ndof > NORM_CHUNK) {
      b.norm_stride = (max_ndof + NORM_CHUNK - 1) / NORM_CHUNK;
      b.norm_part = dalloc<double>((size_t)b.ncell * b.norm_stride);
      b.norm_cnt = dalloc<unsigned>(b.ncell);
      HIPCHECK(hipMemset(b.norm_cnt, 0, sizeof(unsigned) * b.ncell));
    }
  }
  // sparse cells: block-CSR pattern (each node's neighbours and itself,
  // columns ascending), the element block positions and the PCG chunks
  b.csr_cells_h.clear();
  b.chunk_off_h.assign(b.ncell, 0);
  b.nchunk_h.assign(b.ncell, 0);
  b.blk_begin_h.assign(b.ncell, 0);
  b.blk_end_h.assign(b.ncell, 0);
  std::vector<int> bpos_all;  // block-CSR position of element block (a, b), sparse cells
  if (b.any_csr) {
    std::vector<std::vector<int>> nbr(N);
    for (int e = 0; e < T; ++e) {
      if (!b.cp_h[el_cell[e]].csr) continue;
      for (int a = 0; a < 3; ++a)
        for (int q = 0; q < 3; ++q) nbr[el_tri[3 * (size_t)e + a]].push_back(el_tri[3 * (size_t)e + q]);
    }
    std::vector<int> rp(N + 1, 0), bcol, dpos(N, 0);
    for (int i = 0; i < N; ++i) {
      std::sort(nbr[i].begin(), nbr[i].end());
      nbr[i].erase(std::unique(nbr[i].begin(), nbr[i].end()), nbr[i].end());
      rp[i + 1] = rp[i] + (int)nbr[i].size();
      for (size_t k = 0; k < nbr[i].size(); ++k)
        if (nbr[i][k] == i) dpos[i] = rp[i] + (int)k;
      bcol.insert(bcol.end(), nbr[i].begin(), nbr[i].end());
    }
    std::vector<int> bpos(9 * (size_t)T, 0);
    for (int e = 0; e < T; ++e) {
      if (!b.cp_h[el_cell[e]].csr) continue;
      for (int a = 0; a < 3; ++a) {
        const int na = el_tri[3 * (size_t)e + a];
        for (int q = 0; q < 3; ++q) {
          const int nq = el_tri[3 * (size_t)e + q];
          bpos[9 * (size_t)e + 3 * a + q] =
              rp[na] + (int)(std::lower_bound(nbr[na].begin(), nbr[na].end(), nq) - nbr[na].begin());
        }
      }
    }
    std::vector<int4> ch;
    for (int c = 0; c < b.ncell; ++c) {
      const CellParams &p = b.cp_h[c];
      if (!p.csr) continue;
      b.csr_cells_h.push_back(c);
      b.chunk_off_h[c] = (int)ch.size();
      for (int i = p.node_off; i < p.node_off + p.n_nodes; i += CG_ROWS)
        ch.push_back(make_int4(c, i, std::min(i + CG_ROWS, p.node_off + p.n_nodes), 0));
      b.nchunk_h[c] = (int)ch.size() - b.chunk_off_h[c];
      b.blk_begin_h[c] = rp[p.node_off];
      b.blk_end_h[c] = rp[p.node_off + p.n_nodes];
    }
    b.nchunk = (int)ch.size();
    b.nnzb = rp[N];
    b.row_ptr = dupload(rp);
    b.bcol = dupload(bcol);
    b.diag_pos = dupload(dpos);
    b.el_bpos = dupload(bpos);
    bpos_all = std::move(bpos);
    b.chunks = dupload(ch);
    b.csr_cells = dupload(b.csr_cells_h);
    b.Kb = (void *)dalloc<unsigned char>((size_t)std::max<long long>(b.nnzb, 1) * 4 * elem_size);
    b.cgvec = (void *)dalloc<unsigned char>((size_t)std::max(N, 1) * 16 * elem_size);
    std::vector<CgState> cg(2 * (size_t)b.ncell);  // the scalar state twice (by parity)
    for (int c = 0; c < 2 * b.ncell; ++c) {
      memset(&cg[c], 0, sizeof cg[c]);
      cg[c].chunk_off = b.chunk_off_h[c % b.ncell];
      cg[c].nchunk = b.nchunk_h[c % b.ncell];
      cg[c].done = 1;
    }
    b.cg = dupload(cg);
    b.partials = dalloc<double>(3 * (size_t)std::max(b.nchunk, 1));  // 3 sums per chunk
  }
  {
    // node gather plan: for node n, its distinct column nodes in order of
    // first appearance along its elements (element order) and, for each, the
    // contributing element blocks in element order -- the order in which the
    // colour-by-colour scatter adds them to the same entries
    // Built per cell (its nodes are one range) on host threads and joined
    // in cell order: the same plan as one pass over the nodes.
    struct Part {
      std::vector<int> gcol, gbp, gcnt, gke, gnode, ncnt;
    };
    std::vector<Part> parts(b.ncell);
    auto plan_cell = [&](int c) {
      Part &P = parts[c];
      const int n0 = b.cp_h[c].node_off, n1 = n0 + b.cp_h[c].n_nodes;
      std::vector<int> cols, bps;
      std::vector<std::vector<int>> lists;
      for (int n = n0; n < n1; ++n) {
        cols.clear();
        bps.clear();
        for (auto &l : lists) l.clear();
        for (int k = adj_start[n]; k < adj_start[n + 1]; ++k) {
          const int e = adj[k] >> 2, a = adj[k] & 3;
          for (int q = 0; q < 3; ++q) {
            const int col = el_tri[3 * (size_t)e + q];
            size_t g = 0;
            while (g < cols.size() && cols[g] != col) ++g;
            if (g == cols.size()) {
              cols.push_back(col);
              if (lists.size() < cols.size()) lists.emplace_back();
              bps.push_back(bpos_all.empty() ? -1 : bpos_all[9 * (size_t)e + 3 * a + q]);
            }
            lists[g].push_back(36 * e + 12 * a + 2 * q);
          }
        }
        for (size_t g = 0; g < cols.size(); ++g) {
          P.gcol.push_back(cols[g]);
          P.gnode.push_back(n);
          P.gbp.push_back(bps[g]);
          P.gke.insert(P.gke.end(), lists[g].begin(), lists[g].end());
          P.gcnt.push_back((int)lists[g].size());
        }
        P.ncnt.push_back((int)cols.size());
      }
    };
    parallel_cells(b.ncell, plan_cell);
    std::vector<int> gs(N + 1, 0), gcol, gbp, gcs(1, 0), gke, gnode;
    gcol.reserve(7 * (size_t)N);
    gbp.reserve(7 * (size_t)N);
    gke.reserve(9 * (size_t)T);
    for (int c = 0; c < b.ncell; ++c) {
      const Part &P = parts[c];
      const int n0 = b.cp_h[c].node_off;
      for (size_t i = 0; i < P.ncnt.size(); ++i) gs[n0 + i + 1] = P.ncnt[i];
      gcol.insert(gcol.end(), P.gcol.begin(), P.gcol.end());
      gbp.insert(gbp.end(), P.gbp.begin(), P.gbp.end());
      gnode.insert(gnode.end(), P.gnode.begin(), P.gnode.end());
      gke.insert(gke.end(), P.gke.begin(), P.gke.end());
      for (int k : P.gcnt) gcs.push_back(gcs.back() + k);
    }
    for (int n = 0; n < N; ++n) gs[n + 1] += gs[n];
    if (36LL * T >= (1LL << 31)) fail(WCFE_ERR_ARG, "batch of %d elements exceeds the gather plan's 32-bit offsets", T);
    b.g_start = dupload(gs);
    b.g_col = dupload(gcol);
    b.g_bpos = dupload(gbp);
    b.g_cstart = dupload(gcs);
    b.g_ke = dupload(gke);
    b.g_node = dupload(gnode);
    b.g_start_h = gs;
  }
  b.plan.build(b.cp_h, ctx->opt.pivoting, b.lu_groups, b.lu_depth, elem_size == 4 ? ctx->opt.fused_step : 0,
               2 * ctx->n_cu);
  if (ctx->opt.pivoting == WCFE_PIVOT_PARTIAL &&
      lu_solve_lds_bytes(b.max_npad, elem_size) > lu_solve_max_dynamic(ctx->max_lds))
    fail(WCFE_ERR_ARG, "system of %d unknowns exceeds the LDS-resident triangular solve (%zu bytes max)", b.max_npad,
         ctx->max_lds);
}

struct Scope {
  const int *cells;
  int ncell;
  int node_begin, node_end, spr_begin, spr_end;
  int cell;  // -1 = whole batch
};

Scope scope_all(const Batch &b) { return Scope{b.cells, b.ncell, 0, b.N, 0, b.S, -1}; }
Scope scope_cell(const Batch &b, int c) {
  const CellParams &p = b.cp_h[c];
  return Scope{b.cells + c, 1, p.node_off, p.node_off + p.n_nodes, p.spr_off, p.spr_off + p.n_spr, c};
}

void color_range(const Batch &b, const Scope &sc, int k, int &e0, int &e1) {
  if (sc.cell < 0) {
    e0 = b.color_start[k];
    e1 = b.color_start[k + 1];
  } else {
    e0 = b.cell_color[sc.cell][2 * k];
    e1 = b.cell_color[sc.cell][2 * k + 1];
  }
}

// The sparse cells of a scope: their chunks and their id list (device).
bool scope_sparse(const Batch &b, const Scope &sc, int &cbase, int &nch, const int *&cl, int &ncl) {
  if (!b.any_csr) return false;
  if (sc.cell < 0) {
    cbase = 0;
    nch = b.nchunk;
    cl = b.csr_cells;
    ncl = (int)b.csr_cells_h.size();
  } else {
    if (!b.cp_h[sc.cell].csr) return false;
    cbase = b.chunk_off_h[sc.cell];
    nch = b.nchunk_h[sc.cell];
    cl = sc.cells;
    ncl = 1;
  }
  return nch > 0;
}

template <class T> CgVectors<T> cg_vectors(const Batch &b) {
  T *base = static_cast<T *>(b.cgvec);
  const long long n2 = 2LL * b.N;
  return CgVectors<T>{base,          base + n2,     base + 2 * n2, base + 3 * n2,
                      base + 4 * n2, base + 5 * n2, base + 6 * n2, (long long)std::max(b.nchunk, 1)};
}

BndView bnd_view(const wcfe_ctx *ctx, const Batch &b) {
  BndView bv;
  bv.bidx = b.bt_bidx; bv.node = b.bt_node; bv.cst = b.bt_cst; bv.cand = b.bt_cand;
  bv.frame = b.bt_frame; bv.g = b.bt_g; bv.J = b.bt_J; bv.area = b.bt_area; bv.gd = b.bt_gd;
  bv.replica = b.replica;
  bv.fext = b.bt_fext;
  bv.dt = ctx->opt.dt;
  bv.key_lo = ctx->opt.seed_lo;
  bv.key_hi = ctx->opt.seed_hi;
  return bv;
}

// Element assembly colour by colour; in FULL and RESIDUAL mode followed by
// the node terms (boundary and body forces) of the cells that have them.
template <class T> void enqueue_assemble(wcfe_ctx *ctx, Batch &b, const Scope &sc, int mode) {
  hipStream_t s = ctx->stream;
  View v = make_view(ctx, b);
  T *K = static_cast<T *>(b.K);
  if (sc.cell < 0 && mode != ASM_STRAIN) {
    // whole batch: every element at once, then each node gathers its rows
    const GatherPlan gp{b.g_start, b.g_col, b.g_bpos, b.g_cstart, b.g_ke, b.g_node,
                        b.g_start_h[sc.node_begin], b.g_start_h[sc.node_end]};
    launch_assemble_gather<T>(s, v, 0, b.T, mode, b.node_cell, b.adj_start, b.adj, gp, sc.node_begin, sc.node_end,
                              b.Ge, b.Ke64, b.G, K, static_cast<T *>(b.Kb), b.plan.fused && b.solve_wg ? 1 : 0);
  } else {
    for (int k = 0; k < b.ncolor; ++k) {
      int e0, e1;
      color_range(b, sc, k, e0, e1);
      launch_assemble<T>(s, v, e0, e1, mode, K, static_cast<T *>(b.Kb), b.G, b.Ke64, b.Rea, b.MM);
    }
  }
  if (!b.any_bterm || mode == ASM_STRAIN) return;
  const BndView bv = bnd_view(ctx, b);
  const int b0 = sc.cell < 0 ? 0 : b.cp_h[sc.cell].bnd_off;
  const int b1 = sc.cell < 0 ? b.nbnd : b0 + b.cp_h[sc.cell].n_bnd;
  launch_bnd_frames(s, bv, b.x, b0, b1);
  if (b.any_area) launch_cell_area(s, b.cp, sc.cells, sc.ncell, b.crange, b.ncolor, b.el_tri, b.x, b.bt_area);
  launch_node_terms<T>(s, mode == ASM_FULL, v, bv, b.node_cell, sc.node_begin, sc.node_end, K,
                       static_cast<T *>(b.Kb), b.row_ptr, b.bcol, b.diag_pos, b.G);
}

// Launch the LU stages of a plan for a list of cells.
template <class T>
void enqueue_lu(wcfe_ctx *ctx, hipStream_t s, const LuPlan &plan, const std::vector<Stage> &stages, const int *cells,
                int ncell, const CellParams *cp, CellState *cs, T *K, T *Dinv, int *ipiv, StepFwd fwd = StepFwd{}) {
  const bool piv = plan.pivoting != 0;
  {
    PhaseScope ph(ctx, WCFE_PH_LU_PANEL);
    if (!piv) launch_diag_tau<T>(s, cp, cs, cells, ncell, K);
  }
  for (const Stage &g : stages) {
    const WorkList wl{plan.d_prefix + g.off, g.tot};
    switch (g.kind) {
      case ST_DIAG: {
        PhaseScope ph(ctx, WCFE_PH_LU_PANEL);
        launch_lu_diag_inv<T>(s, cp, cs, wl, g.k, K, Dinv);
        break;
      }
      case ST_PANEL: {
        PhaseScope ph(ctx, WCFE_PH_LU_PANEL);
        launch_lu_panel_pp<T>(s, cp, cs, wl, g.k, K, ipiv, Dinv);
        break;
      }
      case ST_SWP: {
        PhaseScope ph(ctx, WCFE_PH_LU_PANEL);
        launch_lu_laswp<T>(s, cp, wl, g.k, K, ipiv);
        break;
      }
      case ST_TRSM: {
        PhaseScope ph(ctx, WCFE_PH_LU_PANEL);
        launch_lu_trsm<T>(s, cp, cs, wl, g.k, K, Dinv);
        break;
      }
      case ST_STRIP: {
        KernelTimer kt(ctx, WCFE_PH_LU_STRIP, g.work);
        launch_lu_gemm<T>(s, cp, cs, wl, g.k, K, Dinv, 1, g.np, 1, kt.a, kt.b);
        break;
      }
      case ST_GEMM1: {
        KernelTimer kt(ctx, WCFE_PH_LU_GEMM, g.work);
        launch_lu_gemm<T>(s, cp, cs, wl, g.k, K, Dinv, piv ? 0 : 1, 1, 0, kt.a, kt.b);
        break;
      }
      case ST_GEMM: {
        KernelTimer kt(ctx, WCFE_PH_LU_GEMM, g.work);
        launch_lu_gemm<T>(s, cp, cs, wl, g.k, K, Dinv, 1, g.np, 0, kt.a, kt.b);
        break;
      }
      case ST_STEP: {
        if (sizeof(T) != 4) fail(WCFE_ERR_STATE, "the fused block step is built for FP32 factors only");
        KernelTimer kt(ctx, WCFE_PH_LU_GEMM, g.work);
        launch_lu_step<T>(s, cp, cs, wl, g.k, K, Dinv, LuPlan::step_flag(g.np), fwd, kt.a, kt.b);
        break;
      }
      case ST_CELL: {
        if (sizeof(T) != 4) fail(WCFE_ERR_STATE, "the fused block step is built for FP32 factors only");
        KernelTimer kt(ctx, WCFE_PH_LU_GEMM, g.work);
        launch_lu_cell<T>(s, cp, cs, wl, K, Dinv, fwd, kt.a, kt.b);
        break;
      }
    }
  }
}

template <class T> void enqueue_factor(wcfe_ctx *ctx, Batch &b, const Scope &sc) {
  const std::vector<Stage> &stages = sc.cell < 0 ? b.plan.batched : b.plan.single[sc.cell];
  if (sc.cell < 0 && !ctx->prof && b.plan.grp.size() > 1 && ctx->side.size() >= b.plan.grp.size()) {
    // independent cell groups on concurrent streams: one group's latency-bound
    // stages (diagonal blocks, block solves, strips) overlap another's
    // trailing updates
    HIPCHECK(hipEventRecord(ctx->fork_ev, ctx->stream));
    for (size_t g = 0; g < b.plan.grp.size(); ++g) {
      hipStream_t gs = ctx->side[g];
      HIPCHECK(hipStreamWaitEvent(gs, ctx->fork_ev, 0));
      enqueue_lu<T>(ctx, gs, b.plan, b.plan.grp[g], b.plan.d_prefix + b.plan.grp_cells_off[g], b.plan.grp_ncell[g],
                    b.cp, b.cs, static_cast<T *>(b.K), static_cast<T *>(b.Dinv), b.ipiv,
                    StepFwd{b.yfwd, b.kinv, b.G});
      HIPCHECK(hipEventRecord(ctx->join_ev[g], gs));
      HIPCHECK(hipStreamWaitEvent(ctx->stream, ctx->join_ev[g], 0));
    }
  } else {
    enqueue_lu<T>(ctx, ctx->stream, b.plan, stages, sc.cells, sc.ncell, b.cp, b.cs, static_cast<T *>(b.K),
                  static_cast<T *>(b.Dinv), b.ipiv, StepFwd{b.yfwd, b.kinv, b.G});
  }
  const int *cl;
  int cb, nch, ncl;
  if (scope_sparse(b, sc, cb, nch, cl, ncl)) {
    PhaseScope ph(ctx, WCFE_PH_LU_PANEL);
    launch_bsr_diag_inv<T>(ctx->stream, b.chunks + cb, nch, b.diag_pos, static_cast<const T *>(b.Kb),
                           cg_vectors<T>(b).dinv);
  }
}

// Block-Jacobi PCG on the sparse cells of the scope for the refinement
// right-hand side b.r; the correction goes to (or is added to) b.delta.
// Convergence is checked on the host after batches of iterations (each check
// drains the stream): the first after as many as the same refinement sweep
// took in the batch's last Newton step (at least 8; one when that sweep had
// nothing to solve, as the sweeps after every cell's convergence; 16 at
// first and for single-cell scopes), then as many as the slowest cell's residual reduction so far
// predicts it still needs (8 .. 128, a quarter more than predicted);
// iterations after a cell's convergence return at once, so the batch sizes
// change only the number of checks, not the result.
// In a batch the first update uses the scalars of the start or of the last
// check, each later update forms its own from the SpMV before it (the scalar
// state alternates between the two halves of b.cg), and the batch ends with
// the scalars of its last SpMV, which the host reads.
template <class T> void enqueue_pcg(wcfe_ctx *ctx, Batch &b, const Scope &sc, int sweep) {
  const int accumulate = sweep > 0;
  const int *cl;
  int cb, nch, ncl;
  if (!scope_sparse(b, sc, cb, nch, cl, ncl)) return;
  hipStream_t s = ctx->stream;
  const CgVectors<T> v = cg_vectors<T>(b);
  const double rtol = ctx->opt.cg_rtol;
  const int maxit = std::max(1, ctx->opt.cg_max_iter);
  const int lanes = ctx->opt.pcg_lanes > 0 ? ctx->opt.pcg_lanes : 4;
  CgState *half[2] = {b.cg, b.cg + b.ncell};
  launch_pcg_init<T>(s, b.chunks, cb, nch, cl, ncl, b.cs, half[0], b.row_ptr, b.bcol, static_cast<const T *>(b.Kb),
                     b.r, v, b.partials, rtol, ctx->opt.lin_rtol, maxit, lanes);
  int cur = 0;
  std::vector<int> ids = sc.cell < 0 ? b.csr_cells_h : std::vector<int>{sc.cell};
  std::vector<CgState> h(b.ncell);
  int *hint = sc.cell < 0 && sweep < 8 ? &b.pcg_hint[sweep] : nullptr;
  int batch = !hint || *hint < 0 ? 16 : *hint == 0 ? 1 : std::min(std::max(8, *hint), maxit);
  for (int it = 0; it < maxit; it += batch) {
    for (int k = 0; k < batch; ++k) {
      launch_pcg_iteration<T>(s, b.chunks, cb, nch, half[cur], half[1 - cur], k > 0, b.row_ptr, b.bcol,
                              static_cast<const T *>(b.Kb), v, b.partials, rtol, maxit, lanes);
      if (k > 0) cur = 1 - cur;
    }
    launch_pcg_check(s, cl, ncl, b.cs, half[cur], half[1 - cur], b.partials, v.pstride, rtol, maxit);
    cur = 1 - cur;
    HIPCHECK(hipMemcpyAsync(h.data(), half[cur], sizeof(CgState) * b.ncell, hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    bool all = true;
    double need = 0.0;  // iterations the slowest open cell still needs at its rate so far
    for (int c : ids) {
      const CgState &g = h[c];
      all = all && g.done;
      if (g.done || g.iters <= 0 || !(g.rr0 > 0.0) || !(g.rr > 0.0)) continue;
      const double red = g.rr / g.rr0;  // squared residual reduction so far
      const double per = std::log(red) / g.iters;
      const double left = per < 0.0 ? std::log(g.tol * g.tol / red) / per : 1e9;
      need = std::max(need, left);
    }
    if (all) break;
    batch = (int)std::min(128.0, std::max(8.0, std::ceil(1.25 * need)));
  }
  launch_pcg_finish<T>(s, b.chunks + cb, nch, b.cs, v.x, b.delta, accumulate);
  int most = 0;
  for (int c : ids) {
    ctx->cg_iters_max = std::max(ctx->cg_iters_max, h[c].iters);
    ctx->cg_breakdowns += h[c].breakdown;
    most = std::max(most, h[c].iters);
  }
  if (hint) *hint = most;
}

// The refinement residual for a mesh batch: r = -G - K d from the FP64
// element matrices.
template <class T> void enqueue_solve_refine(wcfe_ctx *ctx, Batch &b, const Scope &sc) {
  hipStream_t s = ctx->stream;
  T *K = static_cast<T *>(b.K);
  View v = make_view(ctx, b);
  const int piv = ctx->opt.pivoting == WCFE_PIVOT_PARTIAL;
  const long long v0 = 2LL * sc.node_begin, vn = 2LL * (sc.node_end - sc.node_begin);
  {
    PhaseScope ph(ctx, WCFE_PH_REFINE);
    launch_ir_control(s, b.cs, sc.cells, sc.ncell, ctx->opt.lin_rtol, 1);
    launch_neg_copy(s, b.G + v0, b.r + v0, vn);
  }
  const int npad = sc.cell < 0 ? b.max_npad : b.cp_h[sc.cell].npad;
  for (int it = 0; it < std::max(1, ctx->opt.max_refine); ++it) {
    {
      PhaseScope ph(ctx, WCFE_PH_SOLVE);
      if (piv) {
        launch_lu_solve<T>(s, b.cp, b.cs, sc.cells, sc.ncell, K, static_cast<const T *>(b.Dinv), b.ipiv, piv, b.r,
                           b.delta, npad, it > 0);
      } else if (npad > 0) {
        // in each dense cell's own numbering: rp = P r, solve, delta (+)= P^T dp
        const LuPlan::SolveLists &L = sc.cell < 0 ? b.plan.solve_batched : b.plan.solve_single[sc.cell];
        if (b.solve_wg) {  // the sweep kernel reads r and writes delta through the numbering itself
          // first solve after a fused factorization: its forward sweep ran in the block steps (r = -G)
          const float *yf = (it == 0 && b.plan.fused) ? b.yfwd : nullptr;
          launch_lu_solve_band<T>(s, b.cp, b.cs, sc.cells, sc.ncell, K, static_cast<const T *>(b.Dinv), b.kpos, b.r,
                                  b.delta, it > 0, npad, b.max_bw, yf, static_cast<T *>(b.ysw));
        } else {
          launch_perm_gather(s, b.node_cell, b.cp, b.kpos, b.r, b.rp, sc.node_begin, sc.node_end);
          launch_lu_solve_mw<T>(s, b.cp, b.cs, sc.cells, sc.ncell, WorkList{b.plan.d_prefix + L.fwd_off, L.tot},
                                WorkList{b.plan.d_prefix + L.bwd_off, L.tot}, K, static_cast<const T *>(b.Dinv), b.rp,
                                b.dp, b.ysol, b.sync, 0);
          launch_perm_scatter(s, b.node_cell, b.cp, b.cs, b.kpos, b.dp, b.delta, sc.node_begin, sc.node_end, it > 0);
        }
      }
      enqueue_pcg<T>(ctx, b, sc, it);
    }
    PhaseScope ph(ctx, WCFE_PH_REFINE);
    const BndView bv = bnd_view(ctx, b);
    if (b.any_area) launch_cell_gdot(s, b.cp, sc.cells, sc.ncell, bv, b.delta);
    launch_matvec_ir(s, v, b.node_cell, b.adj_start, b.adj, b.Ke64, b.G, b.delta, b.r, sc.node_begin, sc.node_end,
                     b.cs, bv);
    launch_cell_norm(s, b.cp, b.cs, sc.cells, sc.ncell, b.r, 1, 1, ctx->opt.lin_rtol, b.norm_work());
    // a batch with sparse cells is not captured (its solves check their
    // convergence on the host): from the second sweep on, stop once every
    // cell of the scope has met lin_rtol -- the sweeps left would launch
    // only idle kernels (and a solve's probing iteration)
    if (b.any_csr && it >= 1 && it + 1 < std::max(1, ctx->opt.max_refine)) {
      std::vector<CellState> h(b.ncell);
      HIPCHECK(hipMemcpyAsync(h.data(), b.cs, sizeof(CellState) * b.ncell, hipMemcpyDeviceToHost, s));
      HIPCHECK(hipStreamSynchronize(s));
      bool all = true;
      for (int c = sc.cell < 0 ? 0 : sc.cell; c < (sc.cell < 0 ? b.ncell : sc.cell + 1); ++c) all = all && h[c].ir_done;
      if (all) break;
    }
  }
}

template <class T> void enqueue_newton(wcfe_ctx *ctx, Batch &b, const Scope &sc) {
  hipStream_t s = ctx->stream;
  T *K = static_cast<T *>(b.K);
  const long long v0 = 2LL * sc.node_begin, vn = 2LL * (sc.node_end - sc.node_begin);
  const int npad = sc.cell < 0 ? b.max_npad : b.cp_h[sc.cell].npad;
  // 1. assemble K(x), G(x)
  {
    PhaseScope ph(ctx, WCFE_PH_FILL);
    // fused band LU with the one-workgroup sweeps: the upper band blocks of
    // symmetric cells are written by the factorization before any read
    launch_fill_systems<T>(s, b.cp, sc.cells, sc.ncell, npad, K, b.plan.fused && b.solve_wg ? 1 : 0);
    if (b.any_csr) {
      const long long bb = sc.cell < 0 ? 0 : b.blk_begin_h[sc.cell], be = sc.cell < 0 ? b.nnzb : b.blk_end_h[sc.cell];
      launch_zero<T>(s, static_cast<T *>(b.Kb) + 4 * bb, 4 * (be - bb));
    }
    launch_fill(s, b.G + v0, vn, 0.0);
  }
  {
    PhaseScope ph(ctx, WCFE_PH_ASSEMBLE, sc.cell < 0 ? b.T : b.cp_h[sc.cell].n_elems);
    enqueue_assemble<T>(ctx, b, sc, ASM_FULL);
    launch_cell_norm(s, b.cp, b.cs, sc.cells, sc.ncell, b.G, 0, 0, -1.0, b.norm_work());
  }
  // 2. factor
  enqueue_factor<T>(ctx, b, sc);
  // 3. solve K d = -G with FP64 refinement
  enqueue_solve_refine<T>(ctx, b, sc);
  // 4. x += d, FP64 residual at the new x, convergence test
  {
    PhaseScope ph(ctx, WCFE_PH_RESIDUAL);
    launch_update_x(s, b.x, b.delta, sc.node_begin, sc.node_end);
    launch_fill(s, b.G + v0, vn, 0.0);
    enqueue_assemble<T>(ctx, b, sc, ASM_RESIDUAL);
    launch_cell_norm(s, b.cp, b.cs, sc.cells, sc.ncell, b.G, 2, 0, -1.0, b.norm_work());
    launch_newton_control(s, b.cs, sc.cells, sc.ncell, ctx->opt.newton_rtol, ctx->opt.newton_atol);
  }
  // 5. commit converged cells
  PhaseScope ph(ctx, WCFE_PH_COMMIT);
  const bool any_ul = sc.cell < 0 ? b.any_ul : (b.cp_h[sc.cell].law == LAW_UL);
  if (any_ul) {
    launch_fill(s, b.Rea + 3LL * sc.node_begin, 3LL * (sc.node_end - sc.node_begin), 0.0);
    launch_fill(s, b.MM + sc.node_begin, sc.node_end - sc.node_begin, 0.0);
    enqueue_assemble<T>(ctx, b, sc, ASM_STRAIN);
  }
  View v = make_view(ctx, b);
  launch_commit_nodes(s, v, b.node_cell, b.X, b.x, b.ea, b.cC, b.fau, b.Rea, b.MM, sc.node_begin, sc.node_end,
                      ctx->opt.dt);
  const bool any_bell = sc.cell < 0 ? b.any_bell : (b.cp_h[sc.cell].kinetics == KIN_BELL);
  if (any_bell)
    launch_commit_springs(s, v, b.sp_cell, b.sp_elem, b.sp_local, b.sp_state, b.sp_anchor, sc.spr_begin, sc.spr_end,
                          ctx->opt.dt, ctx->opt.k_on, ctx->opt.k_off0, ctx->opt.f_bell, ctx->opt.seed_lo,
                          ctx->opt.seed_hi, b.replica, b.transitions);
  launch_commit_cells(s, b.cs, sc.cells, sc.ncell, ctx->opt.dt);
}

void enqueue_step(wcfe_ctx *ctx) {
  Batch &b = ctx->b;
  ctx->cg_iters_max = 0;
  ctx->cg_breakdowns = 0;
  if (ctx->opt.batched) {
    if (mixed(ctx)) enqueue_newton<float>(ctx, b, scope_all(b));
    else enqueue_newton<double>(ctx, b, scope_all(b));
  } else {
    for (int c = 0; c < b.ncell; ++c) {
      if (mixed(ctx)) enqueue_newton<float>(ctx, b, scope_cell(b, c));
      else enqueue_newton<double>(ctx, b, scope_cell(b, c));
    }
  }
  HIPCHECK(hipGetLastError());
}

void check_cell(const wcfe_ctx *ctx, int cell) {
  if (!ctx->finalized) fail(WCFE_ERR_STATE, "wcfe_finalize has not been called");
  if (cell < 0 || cell >= ctx->b.ncell) fail(WCFE_ERR_ARG, "cell %d out of range [0, %d)", cell, ctx->b.ncell);
}

void collect_stats(wcfe_ctx *ctx, wcfe_step_stats *st, int *err_code) {
  Batch &b = ctx->b;
  std::vector<CellState> cs(b.ncell);
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  HIPCHECK(hipMemcpy(cs.data(), b.cs, sizeof(CellState) * b.ncell, hipMemcpyDeviceToHost));
  wcfe_step_stats s;
  memset(&s, 0, sizeof s);
  s.n_cells = b.ncell;
  int degenerate = -1, zero = -1, stall = -1, missed = -1;
  for (int c = 0; c < b.ncell; ++c) {
    const CellState &q = cs[c];
    s.n_converged += q.converged;
    s.max_refine_used = std::max(s.max_refine_used, q.ir_iters);
    s.n_static_pivots += q.n_static;
    s.n_failed += q.ir_done ? 0 : 1;
    if (!q.ir_done && missed < 0) missed = c;
    s.max_lin_relres = std::max(s.max_lin_relres, q.lin_relres);
    const double rel = q.g0 > 0 ? q.gnorm / q.g0 : 0.0;
    s.max_newton_rel = std::max(s.max_newton_rel, rel);
    if (q.degenerate && degenerate < 0) degenerate = c;
    if (q.zero_pivot && zero < 0) zero = c;
    if (q.solve_stall && stall < 0) stall = c;
  }
  s.max_cg_iters = ctx->cg_iters_max;
  s.n_cg_breakdown = ctx->cg_breakdowns;
  if (st) *st = s;
  *err_code = WCFE_OK;
  if (degenerate >= 0) {
    *err_code = WCFE_ERR_DEGENERATE;
    ctx->err = "cell " + std::to_string(degenerate) + " has a zero-area or inverted element";
  } else if (zero >= 0) {
    *err_code = WCFE_ERR_ZERO_PIVOT;
    ctx->err = "cell " + std::to_string(zero) + ": exactly singular column in partial-pivot LU";
  } else if (stall >= 0) {
    *err_code = WCFE_ERR_HIP;
    ctx->err = "cell " + std::to_string(stall) + ": triangular-solve wait exceeded its bound";
  } else if (missed >= 0) {
    // the Newton update of such a cell is not the Newton step (the host may
    // go on, e.g. with more refinement sweeps, but is told)
    *err_code = WCFE_ERR_NOT_CONVERGED;
    char buf[160];
    snprintf(buf, sizeof buf, "cell %d: linear solve stopped at relative residual %.3e > lin_rtol %.1e after %d "
             "refinement sweeps (%d cells)", missed, cs[missed].lin_relres, ctx->opt.lin_rtol, cs[missed].ir_iters,
             s.n_failed);
    ctx->err = buf;
  }
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

void wcfe_default_options(wcfe_options *o) {
  memset(o, 0, sizeof *o);
  o->device = 0;
  o->precision = WCFE_PREC_MIXED;
  o->pivoting = WCFE_PIVOT_STATIC;
  o->batched = 1;
  o->use_graph = 1;
  o->max_refine = 8;
  o->lin_rtol = 1e-8;
  o->newton_rtol = 1e-8;
  o->newton_atol = 1e-12;
  o->max_newton = 20;
  o->dt = 0.005;
  o->t0 = 0.5;
  o->seed_lo = 5;
  o->seed_hi = 0;
  o->k_on = 1.0;
  o->k_off0 = 0.1;
  o->f_bell = 4.0;
  o->kappa = 0.5;
  o->symmetric = 1;
  o->dense_max_dof = 8192;
  o->cg_rtol = 1e-5;
  o->cg_max_iter = 20000;
  o->lu_groups = 0;  // auto
  o->lu_depth = 0;   // auto
  o->band = 1;
  o->fused_step = 1;
  o->twist = 1;
  o->pcg_lanes = 0;
}

int wcfe_create(const wcfe_options *opt, wcfe_ctx **out) {
  if (!out) return WCFE_ERR_ARG;
  *out = nullptr;
  wcfe_ctx *ctx = new (std::nothrow) wcfe_ctx();
  if (!ctx) return WCFE_ERR_OOM;
  if (opt) ctx->opt = *opt;
  else wcfe_default_options(&ctx->opt);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= ctx->opt.device) {
    delete ctx;
    return WCFE_ERR_NO_DEVICE;
  }
  int rc = guarded(ctx, [&] {
    HIPCHECK(hipSetDevice(ctx->opt.device));
    hipDeviceProp_t prop;
    HIPCHECK(hipGetDeviceProperties(&prop, ctx->opt.device));
    ctx->arch = prop.gcnArchName;
    if (ctx->arch.find("gfx950") == std::string::npos)
      fail(WCFE_ERR_NO_DEVICE, "device %d is %s; this library is built for gfx950", ctx->opt.device,
           prop.gcnArchName);
    int lds = 0;
    HIPCHECK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->opt.device));
    ctx->max_lds = (size_t)lds;
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->opt.device) == hipSuccess && ncu > 0)
      ctx->n_cu = ncu;
    HIPCHECK(lu_prepare_attributes(ctx->max_lds));
    HIPCHECK(noalloc_init(ctx->opt.device));  // the device's scratch array address (wcfe_force_fa_noalloc)
    HIPCHECK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    if (ctx->opt.lu_groups < 0 || ctx->opt.lu_groups > 16) fail(WCFE_ERR_ARG, "lu_groups must be in [0, 16]");
    if (ctx->opt.lu_depth < 0 || ctx->opt.lu_depth > 8) fail(WCFE_ERR_ARG, "lu_depth must be in [0, 8]");
    if (ctx->opt.fused_step < 0 || ctx->opt.fused_step > 5) fail(WCFE_ERR_ARG, "fused_step must be in [0, 5]");
    if (ctx->opt.twist < 0 || ctx->opt.twist > 2) fail(WCFE_ERR_ARG, "twist must be 0, 1 or 2");
    if (ctx->opt.pcg_lanes != 0 && ctx->opt.pcg_lanes != 4 && ctx->opt.pcg_lanes != 8 && ctx->opt.pcg_lanes != 16)
      fail(WCFE_ERR_ARG, "pcg_lanes must be 0, 4, 8 or 16");
    if (ctx->opt.band != 0 && ctx->opt.band != 1) fail(WCFE_ERR_ARG, "band must be 0 or 1");
    HIPCHECK(hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming));
    for (int g = 0; g < std::max(ctx->opt.lu_groups, AUTO_GROUPS_MAX); ++g) {
      hipStream_t st;
      hipEvent_t ev;
      HIPCHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
      ctx->side.push_back(st);
      HIPCHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
      ctx->join_ev.push_back(ev);
    }
  });
  if (rc != WCFE_OK) {
    // keep the context so the caller can read the message, but report the code
    *out = ctx;
    return rc;
  }
  *out = ctx;
  return WCFE_OK;
}

int wcfe_destroy(wcfe_ctx *ctx) {
  if (!ctx) return WCFE_OK;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
  if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
  ctx->b.release();
  if (ctx->tab_px) (void)hipFree(ctx->tab_px);
  if (ctx->tab_py) (void)hipFree(ctx->tab_py);
  if (ctx->tab_k) (void)hipFree(ctx->tab_k);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  for (hipStream_t st : ctx->side) (void)hipStreamDestroy(st);
  for (hipEvent_t ev : ctx->join_ev) (void)hipEventDestroy(ev);
  if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
  if (ctx->copy_stream) (void)hipStreamSynchronize(ctx->copy_stream);
  for (FrameSlot &f : ctx->frames) {
    if (f.dev) (void)hipFree(f.dev);
    if (f.host) (void)hipHostFree(f.host);
    if (f.done) (void)hipEventDestroy(f.done);
  }
  if (ctx->frame_ready) (void)hipEventDestroy(ctx->frame_ready);
  if (ctx->copy_stream) (void)hipStreamDestroy(ctx->copy_stream);
  delete ctx;
  return WCFE_OK;
}

const char *wcfe_last_error(const wcfe_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int wcfe_device_arch(const wcfe_ctx *ctx, char *buf, int len) {
  if (!ctx || !buf || len <= 0) return WCFE_ERR_ARG;
  snprintf(buf, len, "%s", ctx->arch.c_str());
  return WCFE_OK;
}

int wcfe_set_table(wcfe_ctx *ctx, int32_t id, int32_t n, const double *px, const double *py, const double *k) {
  if (!ctx) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    if (ctx->finalized) fail(WCFE_ERR_STATE, "tables must be set before wcfe_finalize");
    if (n < 1 || n > MAX_TABLE || !px || !py || !k) fail(WCFE_ERR_ARG, "table %d: need 1..%d knots", id, MAX_TABLE);
    for (int j = 1; j < n; ++j)
      if (!(px[j] > px[j - 1])) fail(WCFE_ERR_ARG, "table %d: knots must increase", id);
    Table t;
    t.px.assign(px, px + n);
    t.py.assign(py, py + n);
    t.k.assign(k, k + n);
    ctx->tables[id] = t;
  });
}

}  // extern "C"

namespace {
// Validate a cell description and copy it to the host side of the context.
HostCell host_cell(wcfe_ctx *ctx, const wcfe_cell_desc *d) {
    if (d->n_nodes <= 0 || d->n_tris <= 0 || !d->X || !d->tris) fail(WCFE_ERR_ARG, "empty mesh");
    if (d->law != WCFE_LAW_UL && d->law != WCFE_LAW_CST) fail(WCFE_ERR_ARG, "unknown law %d", d->law);
    if (d->spring_law != WCFE_SPRING_LINEAR && d->spring_law != WCFE_SPRING_TABLE)
      fail(WCFE_ERR_ARG, "unknown spring law %d", d->spring_law);
    if (d->kinetics < WCFE_KIN_NONE || d->kinetics > WCFE_KIN_BELL) fail(WCFE_ERR_ARG, "unknown kinetics %d", d->kinetics);
    if (d->node_springs && d->law != WCFE_LAW_UL)
      fail(WCFE_ERR_ARG, "node-attached springs follow the reference model and need the UL law");
    if (d->node_springs && (!d->cC || !d->fau)) fail(WCFE_ERR_ARG, "node springs need cC and fau");
    if (d->n_spr < 0 || (d->n_spr > 0 && (!d->spr_tri || !d->spr_bary || !d->spr_anchor)))
      fail(WCFE_ERR_ARG, "barycentric springs need spr_tri, spr_bary and spr_anchor");
    if (d->kinetics == WCFE_KIN_MEANFIELD && !d->node_springs)
      fail(WCFE_ERR_ARG, "mean-field kinetics act on node-attached springs");
    if (d->kinetics == WCFE_KIN_BELL && d->n_spr == 0) fail(WCFE_ERR_ARG, "Bell kinetics act on barycentric springs");
    if (d->spring_law == WCFE_SPRING_TABLE && !ctx->tables.count(d->table_id))
      fail(WCFE_ERR_ARG, "table %d has not been set", d->table_id);
    HostCell h;
    h.n_nodes = d->n_nodes;
    h.n_tris = d->n_tris;
    h.law = d->law;
    h.mu = d->mu; h.E = d->E; h.nu = d->nu; h.thick = d->thick;
    h.node_springs = d->node_springs;
    h.rho_max = d->rho_max;
    h.spring_law = d->spring_law;
    h.k_lin = d->k_lin;
    h.table_id = d->table_id;
    h.kinetics = d->kinetics;
    if (d->solver < WCFE_SOLVER_AUTO || d->solver > WCFE_SOLVER_CSR) fail(WCFE_ERR_ARG, "unknown solver %d", d->solver);
    h.solver = d->solver;
    const bool bterm = d->k_bend != 0 || d->k_area != 0 || d->drag != 0 || d->ap;
    if (bterm && d->law != WCFE_LAW_UL)
      fail(WCFE_ERR_ARG, "boundary and body forces follow the reference cell model and need the UL law");
    if (!(d->k_bend >= 0) || !(d->k_area >= 0) || !(d->drag >= 0))
      fail(WCFE_ERR_ARG, "k_bend, k_area and drag must be >= 0");
    if (d->ap && (!(d->ap_rate >= 0) || !(d->ap_quantum >= 0) || !(d->ap_density >= 0)))
      fail(WCFE_ERR_ARG, "actin polymerisation rate, quantum and density must be >= 0");
    if ((d->drag != 0 || d->ap) && !(ctx->opt.dt > 0))
      fail(WCFE_ERR_ARG, "drag and actin polymerisation need wcfe_options.dt > 0");
    h.k_bend = d->k_bend; h.k_area = d->k_area; h.area_0 = d->area_0; h.drag = d->drag;
    h.ap = d->ap ? 1 : 0;
    h.ap_rate = d->ap_rate; h.ap_quantum = d->ap_quantum; h.ap_density = d->ap_density;
    h.passive = d->passive ? 1 : 0;
    if (d->f_ext) {
      h.f_ext.assign(d->f_ext, d->f_ext + 2 * d->n_nodes);
      for (double v : h.f_ext)
        if (!std::isfinite(v)) fail(WCFE_ERR_ARG, "f_ext holds a non-finite value");
    }
    h.n_spr = d->n_spr;
    h.X.assign(d->X, d->X + 2 * d->n_nodes);
    if (d->x) h.x.assign(d->x, d->x + 2 * d->n_nodes);
    h.tris.assign(d->tris, d->tris + 3 * d->n_tris);
    for (int v : h.tris)
      if (v < 0 || v >= d->n_nodes) fail(WCFE_ERR_ARG, "triangle node index %d out of range", v);
    h.topo_hash = topology_hash(d->n_nodes, d->n_tris, d->tris);
    if (d->ea) h.ea.assign(d->ea, d->ea + 3 * d->n_nodes);
    if (d->cC) h.cC.assign(d->cC, d->cC + d->n_nodes);
    if (d->fau) h.fau.assign(d->fau, d->fau + 2 * d->n_nodes);
    if (d->n_spr > 0) {
      h.spr_tri.assign(d->spr_tri, d->spr_tri + d->n_spr);
      for (int t : h.spr_tri)
        if (t < 0 || t >= d->n_tris) fail(WCFE_ERR_ARG, "spring host triangle %d out of range", t);
      h.spr_bary.assign(d->spr_bary, d->spr_bary + 3 * d->n_spr);
      h.spr_anchor.assign(d->spr_anchor, d->spr_anchor + 2 * d->n_spr);
      if (d->spr_state) h.spr_state.assign(d->spr_state, d->spr_state + d->n_spr);
    }
    return h;
}

// Copy every cell's current device state (positions, strains, bond state,
// load time and step) back into the host copies of the cells.
void state_to_host(wcfe_ctx *ctx) {
  Batch &b = ctx->b;
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  // the batch's arrays in one copy each, then split per cell
  std::vector<CellState> cs(b.ncell);
  HIPCHECK(hipMemcpy(cs.data(), b.cs, sizeof(CellState) * b.ncell, hipMemcpyDeviceToHost));
  const size_t N = (size_t)b.N, S = (size_t)b.S;
  std::vector<double> X(2 * N), x(2 * N), ea(3 * N), cC(N), fau(2 * N), st(S), an(2 * S);
  HIPCHECK(hipMemcpy(X.data(), b.X, sizeof(double2) * N, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(x.data(), b.x, sizeof(double2) * N, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(ea.data(), b.ea, sizeof(double) * 3 * N, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(cC.data(), b.cC, sizeof(double) * N, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(fau.data(), b.fau, sizeof(double2) * N, hipMemcpyDeviceToHost));
  if (S > 0) {
    HIPCHECK(hipMemcpy(st.data(), b.sp_state, sizeof(double) * S, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(an.data(), b.sp_anchor, sizeof(double) * 2 * S, hipMemcpyDeviceToHost));
  }
  for (int c = 0; c < b.ncell; ++c) {
    HostCell &h = ctx->cells[c];
    const CellParams &p = b.cp_h[c];
    const size_t o = (size_t)p.node_off, n = (size_t)h.n_nodes;
    h.X.assign(X.begin() + 2 * o, X.begin() + 2 * (o + n));
    h.x.assign(x.begin() + 2 * o, x.begin() + 2 * (o + n));
    h.ea.assign(ea.begin() + 3 * o, ea.begin() + 3 * (o + n));
    h.cC.assign(cC.begin() + o, cC.begin() + o + n);
    h.fau.assign(fau.begin() + 2 * o, fau.begin() + 2 * (o + n));
    if (h.n_spr > 0) {
      h.spr_state.resize(h.n_spr);
      for (int k = 0; k < h.n_spr; ++k) {
        const size_t g = (size_t)b.spr_pos[c][k];
        h.spr_state[k] = st[g];
        h.spr_anchor[2 * k] = an[2 * g];
        h.spr_anchor[2 * k + 1] = an[2 * g + 1];
      }
    }
    h.time = cs[c].time;
    h.step = cs[c].step;
    h.state = cs[c];
    h.has_state = true;
  }
}
}  // namespace

extern "C" {

int wcfe_add_cell(wcfe_ctx *ctx, const wcfe_cell_desc *d, int32_t *cell_id) {
  if (!ctx || !d) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    if (ctx->finalized) fail(WCFE_ERR_STATE, "cells must be added before wcfe_finalize");
    ctx->cells.push_back(host_cell(ctx, d));
    if (cell_id) *cell_id = (int32_t)ctx->cells.size() - 1;
  });
}

int wcfe_finalize(wcfe_ctx *ctx) {
  if (!ctx) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    if (ctx->finalized) fail(WCFE_ERR_STATE, "already finalized");
    if (ctx->cells.empty()) fail(WCFE_ERR_ARG, "no cells");
    HIPCHECK(hipSetDevice(ctx->opt.device));
    // tables
    std::vector<double> px, py, kk;
    for (auto &kv : ctx->tables) {
      ctx->table_off[kv.first] = {(int)px.size(), (int)kv.second.px.size()};
      px.insert(px.end(), kv.second.px.begin(), kv.second.px.end());
      py.insert(py.end(), kv.second.py.begin(), kv.second.py.end());
      kk.insert(kk.end(), kv.second.k.begin(), kv.second.k.end());
    }
    ctx->tab_px = dupload(px);
    ctx->tab_py = dupload(py);
    ctx->tab_k = dupload(kk);
    build_batch(ctx, ctx->b, ctx->cells, mixed(ctx) ? 4 : 8);
    ctx->finalized = true;
    ++ctx->layout;
  });
}

namespace {
// One batched Newton step on the context's stream (graph replay when enabled).
void step_impl(wcfe_ctx *ctx) {
  HIPCHECK(hipSetDevice(ctx->opt.device));
  ctx->stepped = true;
  if (ctx->opt.use_graph && ctx->opt.batched && !ctx->b.any_csr) {
    if (!ctx->gexec) {
      HIPCHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      try {
        enqueue_step(ctx);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(ctx->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      HIPCHECK(hipStreamEndCapture(ctx->stream, &ctx->graph));
      HIPCHECK(hipGraphInstantiate(&ctx->gexec, ctx->graph, nullptr, nullptr, 0));
    }
    HIPCHECK(hipGraphLaunch(ctx->gexec, ctx->stream));
  } else {
    enqueue_step(ctx);
  }
}

// Drop the captured graph and the device batch and lay the batch out again
// from the host copies of the cells (the state the cells were added with).
void rebuild_batch(wcfe_ctx *ctx) {
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
  if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
  ctx->gexec = nullptr;
  ctx->graph = nullptr;
  ctx->b.release();
  build_batch(ctx, ctx->b, ctx->cells, mixed(ctx) ? 4 : 8);
  ctx->stepped = false;
  ++ctx->layout;
}

// Device bytes of the dense systems the cells would need at a threshold.
double dense_bytes(const wcfe_ctx *ctx, int dense_max_dof) {
  double s = 0.0;
  for (const HostCell &h : ctx->cells) {
    const int n = 2 * h.n_nodes;
    const bool csr = h.solver == WCFE_SOLVER_CSR || (h.solver == WCFE_SOLVER_AUTO && n > dense_max_dof);
    if (csr) continue;
    const double np = std::ceil(n / 64.0) * 64.0;
    s += np * np * (mixed(ctx) ? 4.0 : 8.0);
  }
  return s;
}
}  // namespace

int wcfe_newton_step(wcfe_ctx *ctx, wcfe_step_stats *stats) {
  if (!ctx) return WCFE_ERR_ARG;
  int code = WCFE_OK;
  int rc = guarded(ctx, [&] {
    if (!ctx->finalized) fail(WCFE_ERR_STATE, "wcfe_finalize has not been called");
    step_impl(ctx);
    if (stats) collect_stats(ctx, stats, &code);
  });
  return rc != WCFE_OK ? rc : code;
}

int wcfe_autotune(wcfe_ctx *ctx, int32_t flags, int32_t steps, int32_t *chosen, double *ms_best) {
  if (!ctx || steps < 1) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    if (!ctx->finalized) fail(WCFE_ERR_STATE, "wcfe_finalize has not been called");
    if (ctx->stepped) fail(WCFE_ERR_STATE, "autotune runs on the state the cells were added with: call it before "
                                           "any Newton step");
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    size_t free_b = 0, total_b = 0;
    HIPCHECK(hipMemGetInfo(&free_b, &total_b));
    const double current_dense = dense_bytes(ctx, ctx->opt.dense_max_dof);
    auto trial = [&]() {
      try {
        rebuild_batch(ctx);
        step_impl(ctx);  // warm-up (captures the graph)
      } catch (const ApiError &e) {
        if (e.code != WCFE_ERR_OOM) throw;
        // this candidate does not fit (its batch or a workspace of its
        // first step): skip it
        (void)hipStreamSynchronize(ctx->stream);
        if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
        if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
        ctx->gexec = nullptr;
        ctx->graph = nullptr;
        ctx->b.release();
        (void)hipGetLastError();
        return 1e300;
      }
      HIPCHECK(hipEventRecord(e0, ctx->stream));
      for (int i = 0; i < steps; ++i) step_impl(ctx);
      HIPCHECK(hipEventRecord(e1, ctx->stream));
      HIPCHECK(hipEventSynchronize(e1));
      float ms = 0.0f;
      HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
      return (double)ms / steps;
    };
    const wcfe_options orig = ctx->opt;
    wcfe_options best = ctx->opt;
    double best_ms = 1e300;
    std::vector<std::pair<double, wcfe_options>> tried;  // every feasible candidate, for the final rebuild
    // 1. dense / CSR threshold (cells with solver AUTO)
    if (flags & WCFE_TUNE_SOLVER) {
      const int cands[] = {0, 2048, 4096, 8192, 16384, 1 << 30};
      for (int dm : cands) {
        if (dense_bytes(ctx, dm) > std::min(current_dense + 0.8 * (double)free_b, 0.6 * (double)total_b))
          continue;  // would not fit (with room for the workspaces)
        ctx->opt = best;
        ctx->opt.dense_max_dof = dm;
        const double ms = trial();
        if (ms < 1e300) tried.push_back({ms, ctx->opt});
        if (ms < best_ms) {
          best_ms = ms;
          best = ctx->opt;
        }
      }
    }
    // 2. delayed-update depth and concurrent cell groups of the dense LU;
    //    at depth 1 also separate launches instead of the fused block step
    //    (fewer inverses per step: better when many cells share a launch)
    if (flags & WCFE_TUNE_LU) {
      const int depths[] = {1, 2, 4, 6}, groups[] = {1, 2, 4};
      const wcfe_options base = best;
      for (int d : depths)
        for (int g : groups)
          for (int f : {5, 4, 1, 0}) {
            // the per-launch choice (1) covers the L-item, split panel and
            // panel forms; 5 takes split panel items wherever 1 takes panel
            // items; 4 (one workgroup per cell) with one group in FP32
            if (d > 1 && f != 1) continue;
            if ((f == 4 && (g != 1 || !mixed(ctx))) || (f == 5 && !mixed(ctx))) continue;
            ctx->opt = base;
            ctx->opt.lu_depth = d;
            ctx->opt.lu_groups = g;
            ctx->opt.fused_step = f;
            const double ms = trial();
            if (ms < 1e300) tried.push_back({ms, ctx->opt});
            if (ms < best_ms) {
              best_ms = ms;
              best = ctx->opt;
            }
          }
    }
    // 3. with fused block steps in FP32: one chain per cell against two-chain
    //    cells for every cell longer than the batch needs (the automatic
    //    rule splits only when the block steps are latency bound)
    if ((flags & WCFE_TUNE_LU) && mixed(ctx) && best.fused_step >= 1 &&
        (best.lu_depth == 1 || (best.lu_depth == 0 && ctx->b.lu_depth == 1))) {
      const wcfe_options base = best;
      for (int tw : {0, 2}) {
        if (tw == base.twist) continue;
        ctx->opt = base;
        ctx->opt.twist = tw;
        const double ms = trial();
        if (ms < 1e300) tried.push_back({ms, ctx->opt});
        if (ms < best_ms) {
          best_ms = ms;
          best = ctx->opt;
        }
      }
    }
    // 4. block-CSR PCG: lanes per block row in the SpMV (batches with sparse
    //    cells under the best setting so far)
    if (flags & WCFE_TUNE_PCG) {
      const wcfe_options base = best;
      for (int lanes : {4, 8, 16}) {
        ctx->opt = base;
        ctx->opt.pcg_lanes = lanes;
        const double ms = trial();
        if (ms < 1e300) tried.push_back({ms, ctx->opt});
        if (ms < best_ms) {
          best_ms = ms;
          best = ctx->opt;
        }
        if (!ctx->b.any_csr) break;  // no sparse cell: the lane count does not matter
      }
    }
    // the fastest candidate that can be laid out again (a setting near the
    // memory limit may not fit a second time); else the batch's own
    std::stable_sort(tried.begin(), tried.end(),
                     [](const std::pair<double, wcfe_options> &a, const std::pair<double, wcfe_options> &b) {
                       return a.first < b.first;
                     });
    tried.push_back({0.0, orig});
    best_ms = 0.0;
    for (size_t q = 0; q < tried.size(); ++q) {
      ctx->opt = tried[q].second;
      try {
        rebuild_batch(ctx);
        best_ms = q + 1 < tried.size() ? tried[q].first : 0.0;
        break;
      } catch (const ApiError &e) {
        if (e.code != WCFE_ERR_OOM || q + 1 == tried.size()) throw;
        ctx->b.release();
        (void)hipGetLastError();
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (chosen) {
      chosen[0] = ctx->b.lu_depth;
      chosen[1] = ctx->b.lu_groups;
      chosen[2] = ctx->opt.dense_max_dof;
      chosen[3] = ctx->b.plan.cell_loop ? 4 : ctx->b.plan.fused ? ctx->opt.fused_step : 0;
      int ntw = 0;
      for (const CellParams &p : ctx->b.cp_h) ntw += p.tw_m > 0;
      chosen[4] = ntw;
      chosen[5] = ctx->opt.pcg_lanes > 0 ? ctx->opt.pcg_lanes : 4;
    }
    if (ms_best) *ms_best = best_ms < 1e300 ? best_ms : 0.0;
  });
}

int wcfe_load_step(wcfe_ctx *ctx, int32_t *iterations, wcfe_step_stats *stats) {
  if (!ctx) return WCFE_ERR_ARG;
  int code = WCFE_OK;
  int rc = guarded(ctx, [&] {
    if (!ctx->finalized) fail(WCFE_ERR_STATE, "wcfe_finalize has not been called");
    Batch &b = ctx->b;
    std::vector<CellState> cs(b.ncell);
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    HIPCHECK(hipMemcpy(cs.data(), b.cs, sizeof(CellState) * b.ncell, hipMemcpyDeviceToHost));
    std::vector<int> step0(b.ncell);
    for (int c = 0; c < b.ncell; ++c) step0[c] = cs[c].step;
    int it = 0;
    wcfe_step_stats st;
    memset(&st, 0, sizeof st);
    for (;;) {
      int r = wcfe_newton_step(ctx, &st);
      if (r != WCFE_OK) fail(r, "%s", ctx->err.c_str());
      ++it;
      HIPCHECK(hipMemcpy(cs.data(), b.cs, sizeof(CellState) * b.ncell, hipMemcpyDeviceToHost));
      bool all = true;
      for (int c = 0; c < b.ncell; ++c) all &= cs[c].step > step0[c];
      if (all) break;
      if (it >= ctx->opt.max_newton) {
        code = WCFE_ERR_NOT_CONVERGED;
        ctx->err = "load step did not converge in " + std::to_string(it) + " Newton iterations";
        break;
      }
    }
    if (iterations) *iterations = it;
    if (stats) *stats = st;
  });
  return rc != WCFE_OK ? rc : code;
}

int wcfe_profile_step(wcfe_ctx *ctx, double *phase_ms, int32_t *phase_launches, double *phase_work) {
  if (!ctx || !phase_ms) return WCFE_ERR_ARG;
  Profiler prof;
  int rc = guarded(ctx, [&] {
    if (!ctx->finalized) fail(WCFE_ERR_STATE, "wcfe_finalize has not been called");
    HIPCHECK(hipSetDevice(ctx->opt.device));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    ctx->prof = &prof;
    ctx->stepped = true;
    try {
      enqueue_step(ctx);
    } catch (...) {
      ctx->prof = nullptr;
      throw;
    }
    ctx->prof = nullptr;
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < WCFE_NPHASES; ++i) {
      phase_ms[i] = 0.0;
      if (phase_launches) phase_launches[i] = 0;
      if (phase_work) phase_work[i] = 0.0;
    }
    for (const Profiler::Rec &r : prof.recs) {
      float ms = 0.0f;
      HIPCHECK(hipEventElapsedTime(&ms, r.a, r.b));
      phase_ms[r.phase] += ms;
      if (phase_launches) phase_launches[r.phase] += 1;
      if (phase_work) phase_work[r.phase] += (double)r.work;
    }
  });
  ctx->prof = nullptr;
  return rc;
}

int wcfe_copy_probe(wcfe_ctx *ctx, int64_t bytes, int32_t reps, double *ms_per_copy) {
  if (!ctx || bytes < 16 || reps < 1 || !ms_per_copy) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    HIPCHECK(hipSetDevice(ctx->opt.device));
    bytes &= ~15LL;
    void *src = dalloc<unsigned char>((size_t)bytes);
    void *dst = nullptr;
    try {
      dst = dalloc<unsigned char>((size_t)bytes);
    } catch (...) {
      (void)hipFree(src);
      throw;
    }
    hipEvent_t a, b;
    HIPCHECK(hipEventCreate(&a));
    HIPCHECK(hipEventCreate(&b));
    HIPCHECK(hipMemsetAsync(src, 1, (size_t)bytes, ctx->stream));
    launch_copy(ctx->stream, src, dst, bytes);  // warm up
    HIPCHECK(hipEventRecord(a, ctx->stream));
    for (int i = 0; i < reps; ++i) launch_copy(ctx->stream, src, dst, bytes);
    HIPCHECK(hipEventRecord(b, ctx->stream));
    HIPCHECK(hipEventSynchronize(b));
    float ms = 0.0f;
    HIPCHECK(hipEventElapsedTime(&ms, a, b));
    *ms_per_copy = ms / reps;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipFree(src);
    (void)hipFree(dst);
  });
}

int wcfe_synchronize(wcfe_ctx *ctx) {
  if (!ctx) return WCFE_ERR_ARG;
  return guarded(ctx, [&] { HIPCHECK(hipStreamSynchronize(ctx->stream)); });
}

int wcfe_get_positions(wcfe_ctx *ctx, int32_t cell, double *x) {
  if (!ctx || !x) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    check_cell(ctx, cell);
    const CellParams &p = ctx->b.cp_h[cell];
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    HIPCHECK(hipMemcpy(x, ctx->b.x + p.node_off, sizeof(double2) * p.n_nodes, hipMemcpyDeviceToHost));
  });
}

int wcfe_set_positions(wcfe_ctx *ctx, int32_t cell, const double *x) {
  if (!ctx || !x) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    check_cell(ctx, cell);
    const CellParams &p = ctx->b.cp_h[cell];
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    HIPCHECK(hipMemcpy(ctx->b.x + p.node_off, x, sizeof(double2) * p.n_nodes, hipMemcpyHostToDevice));
  });
}

int wcfe_frame_capture(wcfe_ctx *ctx, int32_t *frame) {
  if (!ctx) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    if (!ctx->finalized) fail(WCFE_ERR_STATE, "wcfe_finalize has not been called");
    HIPCHECK(hipSetDevice(ctx->opt.device));
    Batch &b = ctx->b;
    if (!ctx->copy_stream) {
      HIPCHECK(hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking));
      HIPCHECK(hipEventCreateWithFlags(&ctx->frame_ready, hipEventDisableTiming));
    }
    const int id = ctx->next_frame;
    FrameSlot &f = ctx->frames[id & 1];
    if (f.done) HIPCHECK(hipEventSynchronize(f.done));  // the slot's previous copy has landed
    const size_t N = (size_t)b.N;
    const size_t bytes = 8 * N * sizeof(double) + sizeof(CellState) * (size_t)b.ncell;
    if (bytes > f.bytes) {
      if (f.dev) HIPCHECK(hipFree(f.dev));
      if (f.host) HIPCHECK(hipHostFree(f.host));
      f.dev = nullptr;
      f.host = nullptr;
      f.bytes = 0;
      HIPCHECK(hipMalloc(&f.dev, bytes));
      HIPCHECK(hipHostMalloc(&f.host, bytes, hipHostMallocDefault));
      f.bytes = bytes;
    }
    if (!f.done) HIPCHECK(hipEventCreateWithFlags(&f.done, hipEventDisableTiming));
    // snapshot in stream order on the device, then D2H on the copy stream
    // while the following Newton steps run
    hipStream_t s = ctx->stream;
    double *d = f.dev;
    HIPCHECK(hipMemcpyAsync(d, b.X, sizeof(double) * 2 * N, hipMemcpyDeviceToDevice, s));
    HIPCHECK(hipMemcpyAsync(d + 2 * N, b.ea, sizeof(double) * 3 * N, hipMemcpyDeviceToDevice, s));
    HIPCHECK(hipMemcpyAsync(d + 5 * N, b.cC, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
    HIPCHECK(hipMemcpyAsync(d + 6 * N, b.fau, sizeof(double) * 2 * N, hipMemcpyDeviceToDevice, s));
    HIPCHECK(hipMemcpyAsync(d + 8 * N, b.cs, sizeof(CellState) * b.ncell, hipMemcpyDeviceToDevice, s));
    HIPCHECK(hipEventRecord(ctx->frame_ready, s));
    HIPCHECK(hipStreamWaitEvent(ctx->copy_stream, ctx->frame_ready, 0));
    HIPCHECK(hipMemcpyAsync(f.host, f.dev, bytes, hipMemcpyDeviceToHost, ctx->copy_stream));
    HIPCHECK(hipEventRecord(f.done, ctx->copy_stream));
    f.id = id;
    f.layout = ctx->layout;
    ++ctx->next_frame;
    if (frame) *frame = id;
  });
}

namespace {
// The landed host copy of a captured frame.
const FrameSlot &frame_slot(wcfe_ctx *ctx, int frame) {
  if (frame < 0 || frame >= ctx->next_frame) fail(WCFE_ERR_ARG, "frame %d has not been captured", frame);
  const FrameSlot &f = ctx->frames[frame & 1];
  if (f.id != frame) fail(WCFE_ERR_ARG, "frame %d was replaced by a later capture (two frames in flight)", frame);
  if (f.layout != ctx->layout) fail(WCFE_ERR_STATE, "the batch was laid out again after frame %d", frame);
  HIPCHECK(hipEventSynchronize(f.done));
  return f;
}
}  // namespace

int wcfe_frame_get(wcfe_ctx *ctx, int32_t frame, int32_t cell, double *X, double *ea, double *cC, double *fau,
                   double *time) {
  if (!ctx) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    check_cell(ctx, cell);
    const FrameSlot &f = frame_slot(ctx, frame);
    const size_t N = (size_t)ctx->b.N, o = (size_t)ctx->b.cp_h[cell].node_off, n = (size_t)ctx->b.cp_h[cell].n_nodes;
    const double *h = f.host;
    if (X) std::copy(h + 2 * o, h + 2 * (o + n), X);
    if (ea) std::copy(h + 2 * N + 3 * o, h + 2 * N + 3 * (o + n), ea);
    if (cC) std::copy(h + 5 * N + o, h + 5 * N + o + n, cC);
    if (fau) std::copy(h + 6 * N + 2 * o, h + 6 * N + 2 * (o + n), fau);
    if (time) *time = reinterpret_cast<const CellState *>(h + 8 * N)[cell].time;
  });
}

int wcfe_frame_write_vtk(wcfe_ctx *ctx, int32_t frame, int32_t cell, const char *path, int32_t style,
                         const double *cA, int32_t law, double kint) {
  if (!ctx || !path) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    check_cell(ctx, cell);
    if (style != WCFE_VTK_CELL && style != WCFE_VTK_SUBSTRATE) fail(WCFE_ERR_ARG, "unknown VTK style %d", style);
    if (law != 0 && law != 1) fail(WCFE_ERR_ARG, "law must be 0 (linear kint) or 1 (MD table)");
    const FrameSlot &f = frame_slot(ctx, frame);
    const size_t N = (size_t)ctx->b.N, o = (size_t)ctx->b.cp_h[cell].node_off;
    const HostCell &hc = ctx->cells[cell];
    VtkInput in;
    in.path = path;
    in.style = style;
    in.n_nodes = hc.n_nodes;
    in.n_tris = hc.n_tris;
    in.tris = hc.tris.data();
    in.x = f.host + 2 * o;
    in.x_stride = 2;
    in.ea = f.host + 2 * N + 3 * o;
    in.ea_kind = 3;
    in.cC = f.host + 5 * N + o;
    in.fau = f.host + 6 * N + 2 * o;
    in.fau_stride = 2;
    in.cA = cA;
    in.law = law;
    in.kint = kint;
    in.time = reinterpret_cast<const CellState *>(f.host + 8 * N)[cell].time;
    std::string err;
    const int rc = write_vtk_file(in, err);
    if (rc != WCFE_OK) fail(rc, "%s", err.c_str());
  });
}

int wcfe_set_external_forces(wcfe_ctx *ctx, int32_t cell, const double *f_ext) {
  if (!ctx) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    if (cell < 0 || cell >= (int)ctx->cells.size()) fail(WCFE_ERR_ARG, "cell %d out of range", cell);
    HostCell &h = ctx->cells[cell];
    if (h.f_ext.empty()) fail(WCFE_ERR_ARG, "cell %d was added without f_ext", cell);
    const size_t n = 2 * (size_t)h.n_nodes;
    if (f_ext) {
      for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(f_ext[i])) fail(WCFE_ERR_ARG, "f_ext holds a non-finite value");
      h.f_ext.assign(f_ext, f_ext + n);
    } else {
      std::fill(h.f_ext.begin(), h.f_ext.end(), 0.0);
    }
    if (ctx->finalized) {
      HIPCHECK(hipStreamSynchronize(ctx->stream));
      HIPCHECK(hipMemcpy(ctx->b.bt_fext + 2LL * ctx->b.cp_h[cell].node_off, h.f_ext.data(), sizeof(double) * n,
                         hipMemcpyHostToDevice));
    }
  });
}

int wcfe_get_ul_state(wcfe_ctx *ctx, int32_t cell, double *X, double *ea, double *cC, double *fau) {
  if (!ctx) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    check_cell(ctx, cell);
    const CellParams &p = ctx->b.cp_h[cell];
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    if (X) HIPCHECK(hipMemcpy(X, ctx->b.X + p.node_off, sizeof(double2) * p.n_nodes, hipMemcpyDeviceToHost));
    if (ea) HIPCHECK(hipMemcpy(ea, ctx->b.ea + 3LL * p.node_off, sizeof(double) * 3 * p.n_nodes, hipMemcpyDeviceToHost));
    if (cC) HIPCHECK(hipMemcpy(cC, ctx->b.cC + p.node_off, sizeof(double) * p.n_nodes, hipMemcpyDeviceToHost));
    if (fau) HIPCHECK(hipMemcpy(fau, ctx->b.fau + p.node_off, sizeof(double2) * p.n_nodes, hipMemcpyDeviceToHost));
  });
}

int wcfe_set_ul_state(wcfe_ctx *ctx, int32_t cell, const double *X, const double *ea, const double *cC,
                      const double *fau, double time) {
  if (!ctx) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    check_cell(ctx, cell);
    const CellParams &p = ctx->b.cp_h[cell];
    Batch &b = ctx->b;
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    if (X) HIPCHECK(hipMemcpy(b.X + p.node_off, X, sizeof(double2) * p.n_nodes, hipMemcpyHostToDevice));
    if (ea) HIPCHECK(hipMemcpy(b.ea + 3LL * p.node_off, ea, sizeof(double) * 3 * p.n_nodes, hipMemcpyHostToDevice));
    if (cC) HIPCHECK(hipMemcpy(b.cC + p.node_off, cC, sizeof(double) * p.n_nodes, hipMemcpyHostToDevice));
    if (fau) HIPCHECK(hipMemcpy(b.fau + p.node_off, fau, sizeof(double2) * p.n_nodes, hipMemcpyHostToDevice));
    // the next Newton iteration starts a new load step from this state
    CellState st;
    HIPCHECK(hipMemcpy(&st, b.cs + cell, sizeof st, hipMemcpyDeviceToHost));
    if (time == time) st.time = time;
    st.newton_it = 0;
    st.converged = 0;
    HIPCHECK(hipMemcpy(b.cs + cell, &st, sizeof st, hipMemcpyHostToDevice));
  });
}

int wcfe_get_springs(wcfe_ctx *ctx, int32_t cell, double *state, double *anchor) {
  if (!ctx) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    check_cell(ctx, cell);
    const CellParams &p = ctx->b.cp_h[cell];
    if (p.n_spr == 0) return;
    std::vector<double> st(p.n_spr), an(2 * (size_t)p.n_spr);
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    HIPCHECK(hipMemcpy(st.data(), ctx->b.sp_state + p.spr_off, sizeof(double) * p.n_spr, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(an.data(), ctx->b.sp_anchor + 2LL * p.spr_off, sizeof(double) * 2 * p.n_spr,
                       hipMemcpyDeviceToHost));
    for (int s = 0; s < p.n_spr; ++s) {
      const int g = ctx->b.spr_pos[cell][s] - p.spr_off;
      if (state) state[s] = st[g];
      if (anchor) { anchor[2 * s] = an[2 * g]; anchor[2 * s + 1] = an[2 * g + 1]; }
    }
  });
}

int wcfe_get_cell_info(wcfe_ctx *ctx, int32_t cell, double *time, int32_t *step, int32_t *newton_iter,
                       int32_t *n_colors, int32_t *ndof) {
  if (!ctx) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    check_cell(ctx, cell);
    CellState s;
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    HIPCHECK(hipMemcpy(&s, ctx->b.cs + cell, sizeof s, hipMemcpyDeviceToHost));
    if (time) *time = s.time;
    if (step) *step = s.step;
    if (newton_iter) *newton_iter = s.newton_it;
    if (n_colors) *n_colors = ctx->b.ncolors_cell[cell];
    if (ndof) *ndof = ctx->b.cp_h[cell].ndof;
  });
}

int wcfe_get_cell_layout(wcfe_ctx *ctx, int32_t cell, int32_t *solver, int32_t *nblk, int32_t *bw, int32_t *sym,
                         int32_t *chain_top, int32_t *chain_bottom) {
  if (!ctx) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    check_cell(ctx, cell);
    const CellParams &p = ctx->b.cp_h[cell];
    if (solver) *solver = p.csr ? WCFE_SOLVER_CSR : WCFE_SOLVER_DENSE;
    if (nblk) *nblk = p.nblk;
    if (bw) *bw = p.bw;
    if (sym) *sym = p.sym;
    if (chain_top) *chain_top = p.tw_m;
    if (chain_bottom) *chain_bottom = p.tw_nb;
  });
}

int wcfe_num_cells(const wcfe_ctx *ctx, int32_t *n_cells) {
  if (!ctx || !n_cells) return WCFE_ERR_ARG;
  *n_cells = (int32_t)ctx->cells.size();
  return WCFE_OK;
}

int wcfe_get_cell_sizes(wcfe_ctx *ctx, int32_t cell, int32_t *n_nodes, int32_t *n_tris, int32_t *n_spr) {
  if (!ctx) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    if (cell < 0 || cell >= (int)ctx->cells.size())
      fail(WCFE_ERR_ARG, "cell %d out of range [0, %d)", cell, (int)ctx->cells.size());
    const HostCell &h = ctx->cells[cell];
    if (n_nodes) *n_nodes = h.n_nodes;
    if (n_tris) *n_tris = h.n_tris;
    if (n_spr) *n_spr = h.n_spr;
  });
}

int wcfe_residual_fp64(wcfe_ctx *ctx, int32_t cell, double *G) {
  if (!ctx || !G) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    check_cell(ctx, cell);
    Batch &b = ctx->b;
    Scope sc = scope_cell(b, cell);
    const CellParams &p = b.cp_h[cell];
    launch_fill(ctx->stream, b.G + 2LL * p.node_off, 2LL * p.n_nodes, 0.0);
    if (mixed(ctx)) enqueue_assemble<float>(ctx, b, sc, ASM_RESIDUAL);
    else enqueue_assemble<double>(ctx, b, sc, ASM_RESIDUAL);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    HIPCHECK(hipMemcpy(G, b.G + 2LL * p.node_off, sizeof(double) * 2 * p.n_nodes, hipMemcpyDeviceToHost));
  });
}

int wcfe_element_systems(wcfe_ctx *ctx, int32_t cell, double *Ge, double *Ke) {
  if (!ctx || !Ge || !Ke) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    check_cell(ctx, cell);
    Batch &b = ctx->b;
    const CellParams &p = b.cp_h[cell];
    const int T = p.n_elems;
    double *dG = dalloc<double>(6 * (size_t)T), *dK = dalloc<double>(36 * (size_t)T);
    View v = make_view(ctx, b);
    std::vector<double> hG(6 * (size_t)T), hK(36 * (size_t)T);
    int done = 0;
    std::vector<int> order;  // global ids in dump order
    for (int k = 0; k < b.ncolor; ++k) {
      const int e0 = b.cell_color[cell][2 * k], e1 = b.cell_color[cell][2 * k + 1];
      if (e1 <= e0) continue;
      launch_element_dump(ctx->stream, v, e0, e1, dG + 6LL * done, dK + 36LL * done);
      for (int e = e0; e < e1; ++e) order.push_back(e);
      done += e1 - e0;
    }
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    HIPCHECK(hipMemcpy(hG.data(), dG, sizeof(double) * hG.size(), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(hK.data(), dK, sizeof(double) * hK.size(), hipMemcpyDeviceToHost));
    (void)hipFree(dG);
    (void)hipFree(dK);
    std::map<int, int> slot;
    for (int i = 0; i < (int)order.size(); ++i) slot[order[i]] = i;
    for (int e = 0; e < T; ++e) {
      const int i = slot[b.elem_pos[cell][e]];
      memcpy(Ge + 6LL * e, hG.data() + 6LL * i, sizeof(double) * 6);
      memcpy(Ke + 36LL * e, hK.data() + 36LL * i, sizeof(double) * 36);
    }
  });
}

int wcfe_assembled_system(wcfe_ctx *ctx, int32_t cell, double *K, double *G) {
  if (!ctx) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    check_cell(ctx, cell);
    Batch &b = ctx->b;
    Scope sc = scope_cell(b, cell);
    const CellParams &p = b.cp_h[cell];
    launch_fill(ctx->stream, b.G + 2LL * p.node_off, 2LL * p.n_nodes, 0.0);
    const size_t es = mixed(ctx) ? 4 : 8;
    const long long b0 = p.csr ? b.blk_begin_h[cell] : 0, b1 = p.csr ? b.blk_end_h[cell] : 0;
    if (mixed(ctx)) {
      launch_fill_systems<float>(ctx->stream, b.cp, sc.cells, 1, p.npad, static_cast<float *>(b.K));
      if (p.csr) launch_zero<float>(ctx->stream, static_cast<float *>(b.Kb) + 4 * b0, 4 * (b1 - b0));
      enqueue_assemble<float>(ctx, b, sc, ASM_FULL);
    } else {
      launch_fill_systems<double>(ctx->stream, b.cp, sc.cells, 1, p.npad, static_cast<double *>(b.K));
      if (p.csr) launch_zero<double>(ctx->stream, static_cast<double *>(b.Kb) + 4 * b0, 4 * (b1 - b0));
      enqueue_assemble<double>(ctx, b, sc, ASM_FULL);
    }
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    if (G) HIPCHECK(hipMemcpy(G, b.G + 2LL * p.node_off, sizeof(double) * p.ndof, hipMemcpyDeviceToHost));
    if (K && p.csr) {
      // densify the cell's block-CSR rows
      const long long b0 = b.blk_begin_h[cell], nb = b.blk_end_h[cell] - b0;
      std::vector<int> rp(p.n_nodes + 1), col((size_t)nb);
      std::vector<unsigned char> raw((size_t)nb * 4 * es);
      HIPCHECK(hipMemcpy(rp.data(), b.row_ptr + p.node_off, sizeof(int) * rp.size(), hipMemcpyDeviceToHost));
      HIPCHECK(hipMemcpy(col.data(), b.bcol + b0, sizeof(int) * col.size(), hipMemcpyDeviceToHost));
      HIPCHECK(hipMemcpy(raw.data(), static_cast<unsigned char *>(b.Kb) + (size_t)b0 * 4 * es, raw.size(),
                         hipMemcpyDeviceToHost));
      for (size_t i = 0; i < (size_t)p.ndof * p.ndof; ++i) K[i] = 0.0;
      for (int i = 0; i < p.n_nodes; ++i)
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
          const int j = col[k - b0] - p.node_off;
          for (int q = 0; q < 4; ++q) {
            const size_t o = (size_t)(k - b0) * 4 + q;
            const double v = es == 4 ? (double)reinterpret_cast<float *>(raw.data())[o]
                                     : reinterpret_cast<double *>(raw.data())[o];
            K[(size_t)(2 * i + q / 2) * p.ndof + 2 * j + q % 2] = v;
          }
        }
    } else if (K) {
      std::vector<unsigned char> raw((size_t)p.npad * p.npad * es);
      HIPCHECK(hipMemcpy(raw.data(), static_cast<unsigned char *>(b.K) + (size_t)p.k_off * es, raw.size(),
                         hipMemcpyDeviceToHost));
      const std::vector<int> &kp = b.kpos_h[cell];  // back to the natural numbering
      for (int i = 0; i < p.ndof; ++i)
        for (int j = 0; j < p.ndof; ++j) {
          const int pi = 2 * kp[i / 2] + i % 2, pj = 2 * kp[j / 2] + j % 2;
          if (std::abs(pi / NB - pj / NB) > p.env) {  // outside the fill envelope: zero, never stored
            K[(size_t)i * p.ndof + j] = 0.0;
            continue;
          }
          const size_t o = (size_t)pi * p.npad + pj;
          K[(size_t)i * p.ndof + j] =
              es == 4 ? (double)reinterpret_cast<float *>(raw.data())[o] : reinterpret_cast<double *>(raw.data())[o];
        }
    }
  });
}

int wcfe_get_colors(wcfe_ctx *ctx, int32_t cell, int32_t *color) {
  if (!ctx || !color) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    check_cell(ctx, cell);
    const auto &c = ctx->b.elem_color[cell];
    for (size_t e = 0; e < c.size(); ++e) color[e] = c[e];
  });
}

}  // extern "C"

namespace {

// Device buffers owned for the duration of one stateless call.
struct Scratch {
  std::vector<void *> ptrs;
  template <class T> T *up(const T *h, size_t n) {
    T *d = dalloc<T>(n);
    ptrs.push_back(d);
    if (n && h) HIPCHECK(hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice));
    return d;
  }
  template <class T> T *zeros(size_t n) {
    T *d = dalloc<T>(n);
    ptrs.push_back(d);
    HIPCHECK(hipMemset(d, 0, std::max<size_t>(n, 1) * sizeof(T)));
    return d;
  }
  ~Scratch() {
    for (void *p : ptrs) (void)hipFree(p);
  }
};

template <class T>
void dense_solve_impl(wcfe_ctx *ctx, int n_sys, const int32_t *n, const double *A, const double *b, double *x,
                      double *relres, int32_t *iters, int32_t *ipiv_out) {
  hipStream_t s = ctx->stream;
  std::vector<CellParams> cp(n_sys);
  std::vector<long long> a_off(n_sys);
  long long kp = 0, ap = 0;
  int vo = 0, sys = 0, max_npad = 0;
  for (int i = 0; i < n_sys; ++i) {
    if (n[i] <= 0) fail(WCFE_ERR_ARG, "system %d has size %d", i, n[i]);
    CellParams &p = cp[i];
    memset(&p, 0, sizeof p);
    p.ndof = n[i];
    p.npad = ((n[i] + NB - 1) / NB) * NB;
    p.nblk = p.npad / NB;
    p.bw = p.env = std::max(p.nblk - 1, 0);  // caller's matrices: no band assumed
    p.k_off = kp;
    p.sys_off = sys;
    p.vec_off = vo;
    a_off[i] = ap;
    // exactly symmetric systems take the symmetric (lower-triangle) factorisation
    if (ctx->opt.symmetric && ctx->opt.pivoting == WCFE_PIVOT_STATIC) {
      const double *Ai = A + ap;
      bool sym = true;
      for (int r = 0; r < n[i] && sym; ++r)
        for (int c2 = 0; c2 < r; ++c2)
          if (Ai[(size_t)r * n[i] + c2] != Ai[(size_t)c2 * n[i] + r]) { sym = false; break; }
      p.sym = sym;
    }
    kp += (long long)p.npad * p.npad;
    ap += (long long)n[i] * n[i];
    vo += n[i];
    sys += p.npad;
    max_npad = std::max(max_npad, p.npad);
  }
  if (ctx->opt.pivoting == WCFE_PIVOT_PARTIAL &&
      lu_solve_lds_bytes(max_npad, sizeof(T)) > lu_solve_max_dynamic(ctx->max_lds))
    fail(WCFE_ERR_ARG, "system of %d unknowns exceeds the LDS-resident triangular solve", max_npad);
  std::vector<T> Kh((size_t)kp, T(0));
  for (int i = 0; i < n_sys; ++i) {
    const CellParams &p = cp[i];
    for (int r = 0; r < p.npad; ++r)
      for (int c = 0; c < p.npad; ++c)
        Kh[(size_t)p.k_off + (size_t)r * p.npad + c] =
            (r < p.ndof && c < p.ndof) ? (T)A[a_off[i] + (long long)r * p.ndof + c] : (r == c ? T(1) : T(0));
  }
  std::vector<CellState> cs(n_sys);
  for (auto &q : cs) memset(&q, 0, sizeof q);
  std::vector<int> ident(n_sys);
  for (int i = 0; i < n_sys; ++i) ident[i] = i;
  Scratch sc;
  T *dK = sc.up(Kh.data(), Kh.size());
  double *dA = sc.up(A, (size_t)ap);
  long long *dAoff = sc.up(a_off.data(), a_off.size());
  double *db = sc.up(b, (size_t)vo);
  double *dr = sc.zeros<double>((size_t)vo);
  double *dx = sc.zeros<double>((size_t)vo);
  CellParams *dcp = sc.up(cp.data(), cp.size());
  CellState *dcs = sc.up(cs.data(), cs.size());
  int *dcells = sc.up(ident.data(), ident.size());
  int *dipiv = sc.zeros<int>((size_t)std::max(sys, 1));
  T *dDinv = sc.zeros<T>((size_t)std::max(sys, 1) * 128);
  LuPlan plan;
  plan.build(cp, ctx->opt.pivoting, 1, ctx->opt.lu_depth > 0 ? ctx->opt.lu_depth : 4);  // full matrices
  sc.ptrs.push_back(plan.d_prefix);
  const bool piv = ctx->opt.pivoting == WCFE_PIVOT_PARTIAL;
  enqueue_lu<T>(ctx, ctx->stream, plan, plan.batched, dcells, n_sys, dcp, dcs, dK, dDinv, dipiv);
  // refinement: r = b; repeat { d = LU^-1 r; x += d; r = b - A x (FP64) }
  launch_cell_norm(s, dcp, dcs, dcells, n_sys, db, 0, 0);
  launch_ir_control(s, dcs, dcells, n_sys, ctx->opt.lin_rtol, 1);
  HIPCHECK(hipMemcpyAsync(dr, db, sizeof(double) * vo, hipMemcpyDeviceToDevice, s));
  unsigned char *dy = sc.zeros<unsigned char>(lu_solve_exchange_bytes(std::max(sys, 1), sizeof(T)));
  int *dsync = sc.zeros<int>((size_t)SOLVE_SYNC_INTS);
  for (int it = 0; it < std::max(1, ctx->opt.max_refine); ++it) {
    if (piv)
      launch_lu_solve<T>(s, dcp, dcs, dcells, n_sys, dK, dDinv, dipiv, piv, dr, dx, max_npad, it > 0);
    else
      launch_lu_solve_mw<T>(s, dcp, dcs, dcells, n_sys,
                            WorkList{plan.d_prefix + plan.solve_batched.fwd_off, plan.solve_batched.tot},
                            WorkList{plan.d_prefix + plan.solve_batched.bwd_off, plan.solve_batched.tot}, dK, dDinv,
                            dr, dx, dy, dsync, it > 0);
    launch_dense_residual(s, dcp, dcs, dcells, n_sys, dA, dAoff, db, dx, dr);
    launch_cell_norm(s, dcp, dcs, dcells, n_sys, dr, 1, 1, ctx->opt.lin_rtol);
  }
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(s));
  HIPCHECK(hipMemcpy(x, dx, sizeof(double) * vo, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(cs.data(), dcs, sizeof(CellState) * n_sys, hipMemcpyDeviceToHost));
  for (int i = 0; i < n_sys; ++i) {
    if (relres) relres[i] = cs[i].lin_relres;
    if (iters) iters[i] = cs[i].ir_iters;
    if (cs[i].zero_pivot) fail(WCFE_ERR_ZERO_PIVOT, "system %d is exactly singular", i);
    if (cs[i].solve_stall) fail(WCFE_ERR_HIP, "system %d: triangular-solve wait exceeded its bound", i);
  }
  if (ipiv_out) {
    std::vector<int> pv(std::max(sys, 1));
    HIPCHECK(hipMemcpy(pv.data(), dipiv, sizeof(int) * pv.size(), hipMemcpyDeviceToHost));
    int o = 0;
    for (int i = 0; i < n_sys; ++i)
      for (int j = 0; j < n[i]; ++j) ipiv_out[o++] = pv[cp[i].sys_off + j];
  }
}

// one-ring lists (node -> triangles) in increasing triangle order, as
// NonDestructiveTriMesh builds m_vertex_to_triangle_map
// (nondestructivetrimesh.cpp:489-506)
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

// Uniform grid over a substrate's triangles for point location.  Cell size
// ~ sqrt(area / triangles), at most GRID_MAX cells a side.  A triangle is
// listed in every cell its bounding box meets, widened by 1e-9 of the mesh's
// coordinate scale so that points the containment test accepts through
// rounding are still found.  The z-sign test also accepts the reflection of
// a clockwise (or nearly degenerate) triangle through its first vertex, so
// such triangles list that region too.  A substrate off the plane z = 0
// gets one cell holding every triangle.  Lists are in increasing triangle
// index, the reference's visiting order.
struct HostGrid {
  PointGrid g;
  std::vector<int> start, list;
};
constexpr int GRID_MAX = 2048;
HostGrid build_point_grid(int n_snodes, int n_stris, const int32_t *stris, const double *P) {
  HostGrid h;
  h.g = PointGrid{0.0, 0.0, 1.0, 1, 1, nullptr, nullptr};
  bool planar = true;
  double x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
  for (int i = 0; i < n_snodes; ++i) {
    const double *q = P + 3 * (size_t)i;
    if (q[2] != 0.0 || !std::isfinite(q[0]) || !std::isfinite(q[1])) planar = false;
    x0 = std::min(x0, q[0]); x1 = std::max(x1, q[0]);
    y0 = std::min(y0, q[1]); y1 = std::max(y1, q[1]);
  }
  const double W = x1 - x0, H = y1 - y0, ext = std::max(W, H);
  if (!planar || n_stris == 0 || !(ext > 0.0)) {
    h.start = {0, n_stris};
    h.list.resize(n_stris);
    for (int e = 0; e < n_stris; ++e) h.list[e] = e;
    return h;
  }
  const double scale = std::max({ext, fabs(x0), fabs(x1), fabs(y0), fabs(y1)});
  const double pad = 1e-9 * scale;
  double cell = std::sqrt(std::max(W, ext * 1e-6) * std::max(H, ext * 1e-6) / n_stris);
  cell = std::max(cell, ext / GRID_MAX);
  const double ox = x0 - 2 * pad, oy = y0 - 2 * pad, inv_h = 1.0 / cell;
  const int gx = std::min(GRID_MAX, (int)((W + 4 * pad) * inv_h) + 1);
  const int gy = std::min(GRID_MAX, (int)((H + 4 * pad) * inv_h) + 1);
  auto idx = [&](double v, double o, int g) {
    const double f = std::floor((v - o) * inv_h);
    return f < 0.0 ? 0 : (f >= g ? g - 1 : (int)f);
  };
  std::vector<int4> box(n_stris);
  h.start.assign((size_t)gx * gy + 1, 0);
  for (int e = 0; e < n_stris; ++e) {
    const double *q0 = P + 3 * stris[3 * e], *q1 = P + 3 * stris[3 * e + 1], *q2 = P + 3 * stris[3 * e + 2];
    double bx0 = std::min({q0[0], q1[0], q2[0]}), bx1 = std::max({q0[0], q1[0], q2[0]});
    double by0 = std::min({q0[1], q1[1], q2[1]}), by1 = std::max({q0[1], q1[1], q2[1]});
    const double e1x = q1[0] - q0[0], e1y = q1[1] - q0[1], e2x = q2[0] - q0[0], e2y = q2[1] - q0[1];
    const double orient = e1x * e2y - e1y * e2x;
    if (!(orient > 1e-8 * std::hypot(e1x, e1y) * std::hypot(e2x, e2y))) {  // reflected region q0 - (q - q0)
      bx0 = std::min({bx0, q0[0] - e1x, q0[0] - e2x}); bx1 = std::max({bx1, q0[0] - e1x, q0[0] - e2x});
      by0 = std::min({by0, q0[1] - e1y, q0[1] - e2y}); by1 = std::max({by1, q0[1] - e1y, q0[1] - e2y});
    }
    box[e] = int4{idx(bx0 - pad, ox, gx), idx(bx1 + pad, ox, gx), idx(by0 - pad, oy, gy), idx(by1 + pad, oy, gy)};
    for (int iy = box[e].z; iy <= box[e].w; ++iy)
      for (int ix = box[e].x; ix <= box[e].y; ++ix) h.start[(size_t)iy * gx + ix + 1]++;
  }
  for (size_t c = 0; c < (size_t)gx * gy; ++c) h.start[c + 1] += h.start[c];
  h.list.resize(h.start.back());
  std::vector<int> fill(h.start.begin(), h.start.end() - 1);
  for (int e = 0; e < n_stris; ++e)
    for (int iy = box[e].z; iy <= box[e].w; ++iy)
      for (int ix = box[e].x; ix <= box[e].y; ++ix) h.list[fill[(size_t)iy * gx + ix]++] = e;
  h.g = PointGrid{ox, oy, inv_h, gx, gy, nullptr, nullptr};
  return h;
}
PointGrid upload_grid(Scratch &sc, const HostGrid &h) {
  PointGrid g = h.g;
  g.start = sc.up(h.start.data(), h.start.size());
  g.list = sc.up(h.list.data(), std::max<size_t>(h.list.size(), 1));
  return g;
}

void check_mesh(int n_nodes, int n_tris, const int32_t *tris) {
  if (n_nodes <= 0 || n_tris < 0 || (n_tris > 0 && !tris)) fail(WCFE_ERR_ARG, "invalid mesh sizes");
  for (long long i = 0; i < 3LL * n_tris; ++i)
    if (tris[i] < 0 || tris[i] >= n_nodes) fail(WCFE_ERR_ARG, "triangle node index %d out of range", tris[i]);
}

}  // namespace

extern "C" {

int wcfe_eval_exp(wcfe_ctx *ctx, int64_t n, const double *x, double *y) {
  if (!ctx || n < 0 || n > (1LL << 28) || (n > 0 && (!x || !y))) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    HIPCHECK(hipSetDevice(ctx->opt.device));
    if (n == 0) return;
    Scratch sc;
    const double *dx = sc.up(x, (size_t)n);
    double *dy = sc.zeros<double>((size_t)n);
    launch_eval_exp(ctx->stream, n, dx, dy);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    HIPCHECK(hipMemcpy(y, dy, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost));
  });
}

int wcfe_dense_solve(wcfe_ctx *ctx, int32_t n_sys, const int32_t *n, const double *A, const double *b, double *x,
                     double *relres, int32_t *iters, int32_t *ipiv_out) {
  if (!ctx || n_sys <= 0 || !n || !A || !b || !x) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    HIPCHECK(hipSetDevice(ctx->opt.device));
    if (mixed(ctx)) dense_solve_impl<float>(ctx, n_sys, n, A, b, x, relres, iters, ipiv_out);
    else dense_solve_impl<double>(ctx, n_sys, n, A, b, x, relres, iters, ipiv_out);
  });
}

int wcfe_calculate_res(wcfe_ctx *ctx, int32_t n_nodes, int32_t n_tris, const int32_t *tris, const double *X,
                       const double *x, const double *ea, int32_t mat, double time, double *Res, double *Res_ea,
                       double *MM) {
  if (!ctx || !X || !x || !ea || !Res || !Res_ea || !MM) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    check_mesh(n_nodes, n_tris, tris);
    if (mat != 1 && mat != 2) fail(WCFE_ERR_ARG, "mat must be 1 (cell) or 2 (substrate)");
    HIPCHECK(hipSetDevice(ctx->opt.device));
    HostCell h;
    h.n_nodes = n_nodes;
    h.n_tris = n_tris;
    h.law = WCFE_LAW_UL;
    h.mu = mat == 2 ? 1.0e6 : 1000.0;  // UL:724-727
    h.passive = mat == 2;              // t_act = 0 (UL:763)
    h.thick = 1.0;                     // UL:768
    h.X.resize(2 * (size_t)n_nodes);
    h.x.resize(2 * (size_t)n_nodes);
    h.ea.resize(3 * (size_t)n_nodes);
    for (int i = 0; i < n_nodes; ++i) {
      if (X[3 * i + 2] != 0.0 || x[3 * i + 2] != 0.0)
        fail(WCFE_ERR_ARG, "node %d is off the plane z = 0 (the reference's planar cell, iomesh.cpp:519)", i);
      const double *e6 = ea + 6 * (size_t)i;
      if (fabs(e6[2]) > 1e-10 || fabs(e6[4]) > 1e-10 || fabs(e6[5]) > 1e-10)
        fail(WCFE_ERR_ARG, "node %d carries out-of-plane strain", i);
      h.X[2 * i] = X[3 * i]; h.X[2 * i + 1] = X[3 * i + 1];
      h.x[2 * i] = x[3 * i]; h.x[2 * i + 1] = x[3 * i + 1];
      h.ea[3 * i] = e6[0]; h.ea[3 * i + 1] = e6[1]; h.ea[3 * i + 2] = e6[3];
    }
    h.tris.assign(tris, tris + 3 * (size_t)n_tris);
    Batch tb;
    try {
      build_batch(ctx, tb, std::vector<HostCell>{h}, 8);
      CellState cs;
      memset(&cs, 0, sizeof cs);
      cs.time = time;
      cs.converged = 1;
      HIPCHECK(hipMemcpy(tb.cs, &cs, sizeof cs, hipMemcpyHostToDevice));
      Scope sc = scope_all(tb);
      launch_fill(ctx->stream, tb.G, 2LL * n_nodes, 0.0);
      enqueue_assemble<double>(ctx, tb, sc, ASM_RESIDUAL);
      launch_fill(ctx->stream, tb.Rea, 3LL * n_nodes, 0.0);
      launch_fill(ctx->stream, tb.MM, n_nodes, 0.0);
      enqueue_assemble<double>(ctx, tb, sc, ASM_STRAIN);
      HIPCHECK(hipGetLastError());
      HIPCHECK(hipStreamSynchronize(ctx->stream));
      std::vector<double> G(2 * (size_t)n_nodes), R(3 * (size_t)n_nodes);
      HIPCHECK(hipMemcpy(G.data(), tb.G, sizeof(double) * G.size(), hipMemcpyDeviceToHost));
      HIPCHECK(hipMemcpy(R.data(), tb.Rea, sizeof(double) * R.size(), hipMemcpyDeviceToHost));
      HIPCHECK(hipMemcpy(MM, tb.MM, sizeof(double) * n_nodes, hipMemcpyDeviceToHost));
      int degenerate = 0;
      HIPCHECK(hipMemcpy(&cs, tb.cs, sizeof cs, hipMemcpyDeviceToHost));
      degenerate = cs.degenerate;
      for (int i = 0; i < n_nodes; ++i) {
        Res[3 * i] = G[2 * i]; Res[3 * i + 1] = G[2 * i + 1]; Res[3 * i + 2] = 0.0;
        double *o = Res_ea + 6 * (size_t)i;
        o[0] = R[3 * i]; o[1] = R[3 * i + 1]; o[2] = 0.0; o[3] = R[3 * i + 2]; o[4] = 0.0; o[5] = 0.0;
      }
      tb.release();
      if (degenerate) fail(WCFE_ERR_DEGENERATE, "zero-area or inverted element");
    } catch (...) {
      tb.release();
      throw;
    }
  });
}

int wcfe_force_fa(wcfe_ctx *ctx, int32_t n_nodes, int32_t n_tris, const int32_t *tris, const double *x,
                  const double *cC, const double *fau, int32_t law, double kint, double *F) {
  if (!ctx || !x || !cC || !fau || !F) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    check_mesh(n_nodes, n_tris, tris);
    if (law != 0 && law != 1) fail(WCFE_ERR_ARG, "law must be 0 (linear kint) or 1 (MD table)");
    HIPCHECK(hipSetDevice(ctx->opt.device));
    std::vector<int> st, li;
    one_rings(n_nodes, n_tris, tris, st, li);
    Scratch sc;
    int *dst = sc.up(st.data(), st.size()), *dli = sc.up(li.data(), li.size());
    int *dtr = sc.up(tris, 3 * (size_t)n_tris);
    double *dx = sc.up(x, 3 * (size_t)n_nodes), *dc = sc.up(cC, (size_t)n_nodes);
    double *df = sc.up(fau, 3 * (size_t)n_nodes), *dF = sc.zeros<double>(3 * (size_t)n_nodes);
    launch_ref_force_fa(ctx->stream, n_nodes, dst, dli, dtr, dx, dc, df, law, kint, dF);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    HIPCHECK(hipMemcpy(F, dF, sizeof(double) * 3 * n_nodes, hipMemcpyDeviceToHost));
  });
}

int wcfe_force_fa_noalloc(wcfe_ctx *ctx, int32_t n_nodes, int32_t n_tris, const int32_t *d_tris, const double *d_x,
                          const double *d_cC, const double *d_fau, int32_t law, double kint, double *d_F,
                          int32_t *d_n_bad) {
  if (!ctx || !d_tris || !d_x || !d_cC || !d_fau || !d_F) return WCFE_ERR_ARG;
  if (law != 0 && law != 1) return set_err(ctx, WCFE_ERR_ARG, "law must be 0 (linear kint) or 1 (MD table)");
  if (hipSetDevice(ctx->opt.device) != hipSuccess) return set_err(ctx, WCFE_ERR_HIP, "hipSetDevice failed");
  std::string err;
  const int rc = noalloc_force_fa(ctx->opt.device, ctx->stream, n_nodes, n_tris, d_tris, d_x, d_cC, d_fau, law, kint,
                                  d_F, d_n_bad, err);
  if (rc != WCFE_OK) return set_err(ctx, rc, err);
  return WCFE_OK;
}

int wcfe_update_fa(wcfe_ctx *ctx, int32_t n_nodes, const double *X, const double *x, double *cA, double *cC,
                   double *fau, double dt, int32_t n_snodes, int32_t n_stris, const int32_t *stris, const double *SX,
                   const double *Sx, int32_t law, double kint, int32_t *n_unlocated) {
  if (!ctx || !X || !x || !cA || !cC || !fau || !SX || !Sx) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    if (n_nodes <= 0) fail(WCFE_ERR_ARG, "no cell nodes");
    check_mesh(n_snodes, n_stris, stris);
    if (law != 0 && law != 1) fail(WCFE_ERR_ARG, "law must be 0 (linear kint) or 1 (MD table)");
    HIPCHECK(hipSetDevice(ctx->opt.device));
    Scratch sc;
    double *dX = sc.up(X, 3 * (size_t)n_nodes), *dx = sc.up(x, 3 * (size_t)n_nodes);
    double *dA = sc.up(cA, (size_t)n_nodes), *dC = sc.up(cC, (size_t)n_nodes), *dF = sc.up(fau, 3 * (size_t)n_nodes);
    int *dst = sc.up(stris, 3 * (size_t)n_stris);
    double *dSX = sc.up(SX, 3 * (size_t)n_snodes), *dSx = sc.up(Sx, 3 * (size_t)n_snodes);
    int *dun = sc.zeros<int>(1);
    const PointGrid grid = upload_grid(sc, build_point_grid(n_snodes, n_stris, stris, SX));
    launch_ref_update_fa(ctx->stream, n_nodes, dX, dx, dA, dC, dF, dt, grid, dst, dSX, dSx, law, kint, dun);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    HIPCHECK(hipMemcpy(cA, dA, sizeof(double) * n_nodes, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(cC, dC, sizeof(double) * n_nodes, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(fau, dF, sizeof(double) * 3 * n_nodes, hipMemcpyDeviceToHost));
    int un = 0;
    HIPCHECK(hipMemcpy(&un, dun, sizeof(int), hipMemcpyDeviceToHost));
    if (n_unlocated) *n_unlocated = un;
  });
}

int wcfe_force_fa_subs(wcfe_ctx *ctx, int32_t n_nodes, int32_t n_tris, const int32_t *tris, const double *x_cell,
                       const double *cC, const double *fau, int32_t n_snodes, int32_t n_stris, const int32_t *stris,
                       const double *Sx, int32_t law, double kint, double *F) {
  if (!ctx || !x_cell || !cC || !fau || !Sx || !F) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    check_mesh(n_nodes, n_tris, tris);
    check_mesh(n_snodes, n_stris, stris);
    if (law != 0 && law != 1) fail(WCFE_ERR_ARG, "law must be 0 (linear kint) or 1 (MD table)");
    HIPCHECK(hipSetDevice(ctx->opt.device));
    std::vector<int> st, li;
    one_rings(n_nodes, n_tris, tris, st, li);
    Scratch sc;
    int *drs = sc.up(st.data(), st.size()), *dri = sc.up(li.data(), std::max<size_t>(li.size(), 1));
    int *dtr = sc.up(tris, 3 * (size_t)n_tris), *dst = sc.up(stris, 3 * (size_t)n_stris);
    double *dx = sc.up(x_cell, 3 * (size_t)n_nodes), *dc = sc.up(cC, (size_t)n_nodes);
    double *df = sc.up(fau, 3 * (size_t)n_nodes), *dSx = sc.up(Sx, 3 * (size_t)n_snodes);
    const PointGrid grid = upload_grid(sc, build_point_grid(n_snodes, n_stris, stris, Sx));
    int *dht = sc.zeros<int>((size_t)n_nodes * SUBS_MAX_HITS), *dnh = sc.zeros<int>((size_t)n_nodes);
    double *dhv = sc.zeros<double>((size_t)n_nodes * SUBS_MAX_HITS * 9);
    int *dov = sc.zeros<int>(1);
    launch_ref_force_fa_subs(ctx->stream, n_nodes, drs, dri, dtr, dx, dc, df, grid, dst, dSx, law, kint, dht, dhv,
                             dnh, dov);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    int ov = 0;
    HIPCHECK(hipMemcpy(&ov, dov, sizeof(int), hipMemcpyDeviceToHost));
    if (ov) fail(WCFE_ERR_ARG, "%d cell nodes lie in more than %d substrate triangles", ov, SUBS_MAX_HITS);
    std::vector<int> ht((size_t)n_nodes * SUBS_MAX_HITS), nh(n_nodes);
    std::vector<double> hv(ht.size() * 9);
    HIPCHECK(hipMemcpy(ht.data(), dht, sizeof(int) * ht.size(), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(nh.data(), dnh, sizeof(int) * nh.size(), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(hv.data(), dhv, sizeof(double) * hv.size(), hipMemcpyDeviceToHost));
    // the reference's accumulation order: cell node, then triangle, ascending
    std::fill(F, F + 3 * (size_t)n_snodes, 0.0);
    for (int i = 0; i < n_nodes; ++i)
      for (int k = 0; k < nh[i]; ++k) {
        const size_t hk = (size_t)i * SUBS_MAX_HITS + k;
        const int32_t *t = stris + 3 * (size_t)ht[hk];
        for (int a = 0; a < 3; ++a)
          for (int c = 0; c < 3; ++c) F[3 * (size_t)t[a] + c] += hv[9 * hk + 3 * a + c];
      }
  });
}

int wcfe_get_schedule(wcfe_ctx *ctx, int32_t *lu_depth, int32_t *lu_groups, int32_t *max_band, int32_t *band_solve,
                      int32_t *fused_step) {
  if (!ctx) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    if (!ctx->finalized) fail(WCFE_ERR_STATE, "wcfe_finalize has not been called");
    if (lu_depth) *lu_depth = ctx->b.lu_depth;
    if (lu_groups) *lu_groups = ctx->b.lu_groups;
    if (max_band) *max_band = ctx->b.max_bw;
    if (band_solve) *band_solve = ctx->b.solve_wg ? 1 : 0;
    if (fused_step) *fused_step = ctx->b.plan.cell_loop ? 2 : ctx->b.plan.fused ? 1 : 0;
  });
}

int wcfe_area_tot(wcfe_ctx *ctx, int32_t n_nodes, int32_t n_tris, const int32_t *tris, const double *x,
                  double *out) {
  if (!ctx || !x || !out) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    check_mesh(n_nodes, n_tris, tris);
    if (n_tris <= 0) fail(WCFE_ERR_ARG, "no triangles");
    HIPCHECK(hipSetDevice(ctx->opt.device));
    Scratch sc;
    int *dtr = sc.up(tris, 3 * (size_t)n_tris);
    double *dx = sc.up(x, 3 * (size_t)n_nodes), *dout = sc.zeros<double>(4);
    launch_ref_area_tot(ctx->stream, n_tris, dtr, dx, dout);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    HIPCHECK(hipMemcpy(out, dout, sizeof(double) * 4, hipMemcpyDeviceToHost));
  });
}

}  // extern "C"

namespace {
// One reference-shaped boundary-force call: candidate tables on the host, one
// GPU thread per vertex.
int ref_boundary_call(wcfe_ctx *ctx, int which, int n_nodes, int n_tris, const int32_t *tris, const double *x,
                      const RefBoundaryArgs &a, const double *u, double *F, int32_t *n_open) {
  if (!ctx || !x || !F) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    check_mesh(n_nodes, n_tris, tris);
    HIPCHECK(hipSetDevice(ctx->opt.device));
    std::vector<int> cst;
    std::vector<int2> cand;
    boundary_candidates(n_nodes, n_tris, tris, cst, cand);
    Scratch sc;
    int *dcs = sc.up(cst.data(), cst.size());
    int2 *dca = sc.up(cand.data(), cand.size());
    double *dx = sc.up(x, 3 * (size_t)n_nodes), *dF = sc.zeros<double>(3 * (size_t)n_nodes);
    double *du = u ? sc.up(u, (size_t)n_nodes) : nullptr;
    int *dopen = sc.zeros<int>(1);
    launch_ref_boundary(ctx->stream, which, n_nodes, dcs, dca, dx, a, du, dF, dopen);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    HIPCHECK(hipMemcpy(F, dF, sizeof(double) * 3 * n_nodes, hipMemcpyDeviceToHost));
    int op = 0;
    HIPCHECK(hipMemcpy(&op, dopen, sizeof(int), hipMemcpyDeviceToHost));
    if (n_open) *n_open = op;
  });
}
}  // namespace

extern "C" {

int wcfe_bending_force_2d(wcfe_ctx *ctx, int32_t n_nodes, int32_t n_tris, const int32_t *tris, const double *x,
                          double k_bend, double *F, int32_t *n_open) {
  RefBoundaryArgs a{};
  a.k_bend = k_bend;
  return ref_boundary_call(ctx, REF_BEND, n_nodes, n_tris, tris, x, a, nullptr, F, n_open);
}

int wcfe_area_constraint_force(wcfe_ctx *ctx, int32_t n_nodes, int32_t n_tris, const int32_t *tris, const double *x,
                               double tot_area, double k_area, double area_0, double *F, int32_t *n_open) {
  RefBoundaryArgs a{};
  a.tot_area = tot_area;
  a.k_area = k_area;
  a.area_0 = area_0;
  return ref_boundary_call(ctx, REF_AREA, n_nodes, n_tris, tris, x, a, nullptr, F, n_open);
}

int wcfe_force_ap(wcfe_ctx *ctx, int32_t n_nodes, int32_t n_tris, const int32_t *tris, const double *x, double dt,
                  double rate, double quantum, double density, const double *u, uint32_t draw, double *F,
                  int32_t *n_open) {
  RefBoundaryArgs a{};
  a.dt = dt;
  a.rate = rate;
  a.quantum = quantum;
  a.density = density;
  if (ctx) {
    a.key_lo = ctx->opt.seed_lo;
    a.key_hi = ctx->opt.seed_hi;
  }
  a.draw = draw;
  return ref_boundary_call(ctx, REF_AP, n_nodes, n_tris, tris, x, a, u, F, n_open);
}

}  // extern "C"

namespace {
const char CKP_MAGIC[8] = {'W', 'C', 'F', 'E', 'C', 'K', 'P', '1'};
// version 2 added the boundary and body force parameters after the scalars,
// version 3 the substrate form (passive flag, constant nodal load), version 4
// the size of the per-cell solver state record (stored as raw bytes): a file
// whose record size differs from this build's is refused.  Versions 1-3 carry
// no record size and are refused.
const int CKP_VERSION = 4;
struct CkpWriter {
  FILE *f;
  void raw(const void *p, size_t n) {
    if (n && fwrite(p, 1, n, f) != n) fail(WCFE_ERR_ARG, "checkpoint write failed");
  }
  void i32(int v) { raw(&v, 4); }
  void f64(double v) { raw(&v, 8); }
  template <class T> void vec(const std::vector<T> &v, size_t n) {
    if (v.size() == n) raw(v.data(), n * sizeof(T));
    else {
      std::vector<T> z(n, T(0));
      raw(z.data(), n * sizeof(T));
    }
  }
};
struct CkpReader {
  FILE *f;
  void raw(void *p, size_t n) {
    if (n && fread(p, 1, n, f) != n) fail(WCFE_ERR_ARG, "checkpoint file is truncated");
  }
  int i32() { int v; raw(&v, 4); return v; }
  double f64() { double v; raw(&v, 8); return v; }
  template <class T> void vec(std::vector<T> &v, long long n) {
    if (n < 0 || n > (1LL << 34)) fail(WCFE_ERR_ARG, "checkpoint array size %lld out of range", n);
    v.resize((size_t)n);
    raw(v.data(), (size_t)n * sizeof(T));
  }
};
}  // namespace

extern "C" {

int wcfe_replace_cell(wcfe_ctx *ctx, int32_t cell, const wcfe_cell_desc *d) {
  if (!ctx || !d) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    check_cell(ctx, cell);
    HostCell h = host_cell(ctx, d);
    state_to_host(ctx);  // the other cells keep their current state
    h.time = ctx->cells[cell].time;
    h.step = ctx->cells[cell].step;
    ctx->cells[cell] = std::move(h);
    rebuild_batch(ctx);
    ctx->stepped = true;
  });
}



int wcfe_remesh_cell(wcfe_ctx *ctx, int32_t cell, int32_t n_nodes, int32_t n_tris, const int32_t *tris,
                     const int32_t *src0, const int32_t *src1, const double *w, const double *x_new,
                     const int32_t *spr_tri, const double *spr_bary, int32_t *relaid) {
  if (!ctx || !tris || n_nodes <= 0 || n_tris <= 0) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    check_cell(ctx, cell);
    HostCell &old = ctx->cells[cell];
    Batch &b = ctx->b;
    const CellParams &p = b.cp_h[cell];
    if ((src0 == nullptr) != (src1 == nullptr) || (src0 == nullptr) != (w == nullptr))
      fail(WCFE_ERR_ARG, "src0, src1 and w come together (or all NULL: the nodes are kept as they are)");
    if (!src0 && n_nodes != old.n_nodes) fail(WCFE_ERR_ARG, "a changed node count needs the node maps");
    for (long long i = 0; i < 3LL * n_tris; ++i)
      if (tris[i] < 0 || tris[i] >= n_nodes) fail(WCFE_ERR_ARG, "triangle node index %d out of range", tris[i]);
    bool identity = true;
    if (src0) {
      for (int i = 0; i < n_nodes; ++i) {
        if (src0[i] < 0 || src0[i] >= old.n_nodes || src1[i] < 0 || src1[i] >= old.n_nodes)
          fail(WCFE_ERR_ARG, "node map of new node %d outside the old cell's %d nodes", i, old.n_nodes);
        if (!(w[i] >= 0.0 && w[i] <= 1.0)) fail(WCFE_ERR_ARG, "interpolation weight of node %d outside [0, 1]", i);
        identity &= src0[i] == i && (w[i] == 1.0 || src1[i] == i);
      }
      identity &= n_nodes == old.n_nodes;
    }
    if ((spr_tri == nullptr) != (spr_bary == nullptr)) fail(WCFE_ERR_ARG, "spr_tri and spr_bary come together");
    for (int k = 0; spr_tri && k < old.n_spr; ++k)
      if (spr_tri[k] < 0 || spr_tri[k] >= n_tris) fail(WCFE_ERR_ARG, "spring host triangle %d out of range", spr_tri[k]);
    const uint64_t hash = topology_hash(n_nodes, n_tris, tris);
    const bool same_mesh = hash == old.topo_hash && n_tris == old.n_tris && n_nodes == old.n_nodes &&
                           std::equal(tris, tris + 3LL * n_tris, old.tris.begin());
    const bool same_springs =
        !spr_tri || (std::equal(spr_tri, spr_tri + old.n_spr, old.spr_tri.begin()) &&
                     std::equal(spr_bary, spr_bary + 3LL * old.n_spr, old.spr_bary.begin()));
    if (same_mesh && identity && same_springs) {
      // the topology did not change: positions (if given) in place, no layout
      HIPCHECK(hipStreamSynchronize(ctx->stream));
      if (x_new)
        HIPCHECK(hipMemcpy(b.x + p.node_off, x_new, sizeof(double2) * p.n_nodes, hipMemcpyHostToDevice));
      if (relaid) *relaid = 0;
      return;
    }
    // the cell's new nodes from its old ones, on the device
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    Scratch sc;
    std::vector<int> s0(n_nodes), s1(n_nodes);
    std::vector<double> ww(n_nodes, 1.0);
    for (int i = 0; i < n_nodes; ++i) {
      s0[i] = src0 ? src0[i] : i;
      s1[i] = src1 ? src1[i] : i;
      if (w) ww[i] = w[i];
    }
    const int *ds0 = sc.up(s0.data(), s0.size()), *ds1 = sc.up(s1.data(), s1.size());
    const double *dw = sc.up(ww.data(), ww.size());
    double2 *Xn = sc.zeros<double2>(n_nodes), *xn = sc.zeros<double2>(n_nodes), *fn = sc.zeros<double2>(n_nodes);
    double *en = sc.zeros<double>(3 * (size_t)n_nodes), *cn = sc.zeros<double>(n_nodes);
    launch_remap_nodes(ctx->stream, n_nodes, ds0, ds1, dw, p.node_off, b.X, b.x, b.ea, b.cC, b.fau, Xn, xn, en, cn, fn);
    HIPCHECK(hipGetLastError());
    state_to_host(ctx);  // every cell's state (this cell's springs too)
    HostCell h = old;
    h.n_nodes = n_nodes;
    h.n_tris = n_tris;
    h.tris.assign(tris, tris + 3LL * n_tris);
    h.topo_hash = hash;
    h.X.resize(2 * (size_t)n_nodes);
    h.x.resize(2 * (size_t)n_nodes);
    h.ea.resize(3 * (size_t)n_nodes);
    h.cC.resize(n_nodes);
    h.fau.resize(2 * (size_t)n_nodes);
    HIPCHECK(hipMemcpy(h.X.data(), Xn, sizeof(double2) * n_nodes, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(h.x.data(), xn, sizeof(double2) * n_nodes, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(h.ea.data(), en, sizeof(double) * 3 * n_nodes, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(h.cC.data(), cn, sizeof(double) * n_nodes, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(h.fau.data(), fn, sizeof(double2) * n_nodes, hipMemcpyDeviceToHost));
    if (x_new) h.x.assign(x_new, x_new + 2 * (size_t)n_nodes);
    if (!h.f_ext.empty()) {  // a constant load per node does not interpolate: the host sets it again
      h.f_ext.assign(2 * (size_t)n_nodes, 0.0);
    }
    if (spr_tri) {  // bond states and anchors stay with their springs
      h.spr_tri.assign(spr_tri, spr_tri + old.n_spr);
      h.spr_bary.assign(spr_bary, spr_bary + 3LL * old.n_spr);
    }
    h.has_state = false;  // the cell starts its load step again on the new mesh
    ctx->cells[cell] = std::move(h);
    rebuild_batch(ctx);
    ctx->stepped = true;
    if (relaid) *relaid = 1;
  });
}

int wcfe_checkpoint_save(wcfe_ctx *ctx, const char *path) {
  if (!ctx || !path) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    if (!ctx->finalized) fail(WCFE_ERR_STATE, "wcfe_finalize has not been called");
    state_to_host(ctx);
    FILE *f = fopen(path, "wb");
    if (!f) fail(WCFE_ERR_ARG, "cannot open %s for writing", path);
    CkpWriter w{f};
    try {
      w.raw(CKP_MAGIC, 8);
      w.i32(CKP_VERSION);
      w.i32((int)sizeof(CellState));
      w.i32((int)ctx->cells.size());
      w.i32((int)ctx->tables.size());
      for (auto &kv : ctx->tables) {
        const int n = (int)kv.second.px.size();
        w.i32(kv.first);
        w.i32(n);
        w.vec(kv.second.px, n);
        w.vec(kv.second.py, n);
        w.vec(kv.second.k, n);
      }
      for (const HostCell &h : ctx->cells) {
        const size_t N = h.n_nodes, T = h.n_tris, S = h.n_spr;
        for (int v : {h.n_nodes, h.n_tris, h.law, h.node_springs, h.spring_law, h.table_id, h.kinetics, h.n_spr,
                      h.solver, h.step})
          w.i32(v);
        for (double v : {h.mu, h.E, h.nu, h.thick, h.rho_max, h.k_lin, h.time}) w.f64(v);
        w.i32(h.ap);
        for (double v : {h.k_bend, h.k_area, h.area_0, h.drag, h.ap_rate, h.ap_quantum, h.ap_density}) w.f64(v);
        w.i32(h.passive);
        w.i32(h.f_ext.empty() ? 0 : 1);
        if (!h.f_ext.empty()) w.vec(h.f_ext, 2 * N);
        w.vec(h.X, 2 * N);
        w.vec(h.x, 2 * N);
        w.vec(h.ea, 3 * N);
        w.vec(h.cC, N);
        w.vec(h.fau, 2 * N);
        w.vec(h.tris, 3 * T);
        w.vec(h.spr_tri, S);
        w.vec(h.spr_bary, 3 * S);
        w.vec(h.spr_anchor, 2 * S);
        w.vec(h.spr_state, S);
        w.raw(&h.state, sizeof(CellState));  // Newton iterate state (mid-load-step cells resume exactly)
      }
    } catch (...) {
      fclose(f);
      throw;
    }
    if (fclose(f) != 0) fail(WCFE_ERR_ARG, "closing %s failed", path);
  });
}

int wcfe_checkpoint_load(wcfe_ctx *ctx, const char *path) {
  if (!ctx || !path) return WCFE_ERR_ARG;
  return guarded(ctx, [&] {
    if (ctx->finalized || !ctx->cells.empty()) fail(WCFE_ERR_STATE, "load a checkpoint into an empty context");
    FILE *f = fopen(path, "rb");
    if (!f) fail(WCFE_ERR_ARG, "cannot open %s", path);
    CkpReader r{f};
    std::map<int, Table> tables;
    std::vector<HostCell> cells;
    try {
      char magic[8];
      r.raw(magic, 8);
      if (memcmp(magic, CKP_MAGIC, 8) != 0) fail(WCFE_ERR_ARG, "%s is not a wcfe checkpoint", path);
      const int version = r.i32();
      if (version != CKP_VERSION)
        fail(WCFE_ERR_ARG, "unsupported checkpoint version %d (this build reads version %d)", version, CKP_VERSION);
      const int state_bytes = r.i32();
      if (state_bytes != (int)sizeof(CellState))
        fail(WCFE_ERR_ARG, "checkpoint solver-state record is %d bytes, this build's is %d", state_bytes,
             (int)sizeof(CellState));
      const int ncell = r.i32(), ntab = r.i32();
      if (ncell < 1 || ncell >= (1 << 11) || ntab < 0 || ntab > 1024) fail(WCFE_ERR_ARG, "bad checkpoint header");
      for (int t = 0; t < ntab; ++t) {
        const int id = r.i32(), n = r.i32();
        if (n < 1 || n > MAX_TABLE) fail(WCFE_ERR_ARG, "bad table size in checkpoint");
        Table tb;
        r.vec(tb.px, n);
        r.vec(tb.py, n);
        r.vec(tb.k, n);
        tables[id] = tb;
      }
      for (int c = 0; c < ncell; ++c) {
        HostCell h;
        int *iv[] = {&h.n_nodes, &h.n_tris, &h.law, &h.node_springs, &h.spring_law, &h.table_id, &h.kinetics,
                     &h.n_spr, &h.solver, &h.step};
        for (int *v : iv) *v = r.i32();
        double *dv[] = {&h.mu, &h.E, &h.nu, &h.thick, &h.rho_max, &h.k_lin, &h.time};
        for (double *v : dv) *v = r.f64();
        h.ap = r.i32();
        double *bv[] = {&h.k_bend, &h.k_area, &h.area_0, &h.drag, &h.ap_rate, &h.ap_quantum, &h.ap_density};
        for (double *v : bv) *v = r.f64();
        h.passive = r.i32() ? 1 : 0;
        const int has_ext = r.i32();
        if (h.n_nodes <= 0 || h.n_tris <= 0 || h.n_spr < 0) fail(WCFE_ERR_ARG, "bad cell sizes in checkpoint");
        const long long N = h.n_nodes, T = h.n_tris, S = h.n_spr;
        if (has_ext) r.vec(h.f_ext, 2 * N);
        r.vec(h.X, 2 * N);
        r.vec(h.x, 2 * N);
        r.vec(h.ea, 3 * N);
        r.vec(h.cC, N);
        r.vec(h.fau, 2 * N);
        r.vec(h.tris, 3 * T);
        r.vec(h.spr_tri, S);
        r.vec(h.spr_bary, 3 * S);
        r.vec(h.spr_anchor, 2 * S);
        r.vec(h.spr_state, S);
        r.raw(&h.state, sizeof(CellState));
        cells.push_back(std::move(h));
      }
    } catch (...) {
      fclose(f);
      throw;
    }
    fclose(f);
    // every cell passes the checks wcfe_add_cell applies (enumerations, index
    // ranges, table ids, parameter signs) before the context takes the file
    std::map<int, Table> keep_tables;
    std::swap(keep_tables, ctx->tables);
    ctx->tables = tables;
    std::vector<HostCell> valid;
    try {
      for (HostCell &h : cells) {
        wcfe_cell_desc d;
        memset(&d, 0, sizeof d);
        d.n_nodes = h.n_nodes; d.n_tris = h.n_tris;
        d.X = h.X.data(); d.x = h.x.data(); d.tris = h.tris.data();
        d.law = h.law; d.mu = h.mu; d.E = h.E; d.nu = h.nu; d.thick = h.thick; d.ea = h.ea.data();
        d.node_springs = h.node_springs; d.cC = h.cC.data(); d.fau = h.fau.data(); d.rho_max = h.rho_max;
        d.n_spr = h.n_spr; d.spr_tri = h.spr_tri.data(); d.spr_bary = h.spr_bary.data();
        d.spr_anchor = h.spr_anchor.data(); d.spr_state = h.spr_state.data();
        d.spring_law = h.spring_law; d.k_lin = h.k_lin; d.table_id = h.table_id; d.kinetics = h.kinetics;
        d.solver = h.solver; d.k_bend = h.k_bend; d.k_area = h.k_area; d.area_0 = h.area_0; d.drag = h.drag;
        d.ap = h.ap; d.ap_rate = h.ap_rate; d.ap_quantum = h.ap_quantum; d.ap_density = h.ap_density;
        d.passive = h.passive; d.f_ext = h.f_ext.empty() ? nullptr : h.f_ext.data();
        HostCell v = host_cell(ctx, &d);
        v.time = h.time;
        v.step = h.step;
        v.state = h.state;
        v.has_state = true;
        valid.push_back(std::move(v));
      }
    } catch (...) {
      ctx->tables = std::move(keep_tables);
      throw;
    }
    ctx->cells = std::move(valid);
  });
}

}  // extern "C"
