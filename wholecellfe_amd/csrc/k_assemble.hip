// Assembly, FP64 residual/refinement vectors, Newton control and commit
// kernels for gfx950.
//
// Assembly is colour ordered: one launch per colour; within a colour no two
// elements share a node, so every += into the dense system and the residual
// has a single writer and the summation order per entry is fixed by the
// colour sequence (run-to-run bitwise deterministic, no atomics).
#include "wcfe_internal.hpp"

namespace wcfe {

// ---------------------------------------------------------------------------
// dense system fill: zero, identity on the padded diagonal (rows >= ndof)
// ---------------------------------------------------------------------------
template <class T> struct Vec16;
template <> struct Vec16<float> { typedef float4 type; static constexpr int n = 4; };
template <> struct Vec16<double> { typedef double2 type; static constexpr int n = 2; };

template <class T>
// lower_only: symmetric cells get the blocks up to the diagonal only (the
// fused band LU writes every upper band block, U = L^T, before anything
// reads it); a symmetric two-chain cell's bottom chain factors the block
// reversal, whose lower blocks are the real upper ones: block rows of the
// bottom part get the blocks from the diagonal on, the middle window's rows
// the whole envelope
__global__ __launch_bounds__(256) void k_fill_systems(const CellParams *cp, const int *cells, int ncell, T *K,
                                                      int lower_only) {
  const int ci = blockIdx.y;
  if (ci >= ncell) return;
  const CellParams p = cp[cells[ci]];
  typedef typename Vec16<T>::type V;
  constexpr int VN = Vec16<T>::n;
  // blocks within the fill envelope |i - j| <= env only; one wave per row
  // (a row holds at most (2 env + 1) * 64 / VN vectors)
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int r = 4 * blockIdx.x + wave; r < p.npad; r += 4 * gridDim.x) {
    const int rb = r / NB;
    const bool lo = lower_only && p.sym;
    const bool bottom = lo && p.tw_m > 0 && rb >= p.tw_m + p.bw, top = lo && (p.tw_m == 0 || rb < p.tw_m);
    const int c0 = (bottom ? rb : max(0, rb - p.env)) * (NB / VN),
              c1 = min(p.nblk, rb + 1 + (top ? 0 : p.env)) * (NB / VN);
    V *row = reinterpret_cast<V *>(K + p.k_off + (long long)r * p.npad);
    for (int c = c0 + lane; c < c1; c += 64) {
      T vals[VN];
#pragma unroll
      for (int k = 0; k < VN; ++k) vals[k] = (r >= p.ndof && r == c * VN + k) ? T(1) : T(0);
      V w;
      memcpy(&w, vals, sizeof(V));
      row[c] = w;
    }
  }
}

template <class T>
void launch_fill_systems(hipStream_t s, const CellParams *cp, const int *cells, int ncell, int max_npad, T *K,
                         int lower_only) {
  if (ncell <= 0) return;
  // 4 rows per block and pass; at most 128 blocks per cell (the largest
  // cells loop) so that small cells launch few empty blocks
  int gx = (max_npad + 3) / 4 < 128 ? (max_npad + 3) / 4 : 128;
  if (gx < 1) gx = 1;
  hipLaunchKernelGGL(k_fill_systems<T>, dim3(gx, ncell), dim3(256), 0, s, cp, cells, ncell, K, lower_only);
}
template void launch_fill_systems<float>(hipStream_t, const CellParams *, const int *, int, int, float *, int);
template void launch_fill_systems<double>(hipStream_t, const CellParams *, const int *, int, int, double *, int);

__global__ __launch_bounds__(256) void k_fill(double *p, long long n, double value) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = value;
}
void launch_fill(hipStream_t s, double *p, long long n, double value) {
  if (n <= 0) return;
  long long b = (n + 255) / 256;
  int g = (int)(b < 4096 ? b : 4096);
  hipLaunchKernelGGL(k_fill, dim3(g), dim3(256), 0, s, p, n, value);
}

__global__ __launch_bounds__(256) void k_neg_copy(const double *src, double *dst, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = -src[i];
}
void launch_neg_copy(hipStream_t s, const double *src, double *dst, long long n) {
  if (n <= 0) return;
  long long b = (n + 255) / 256;
  int g = (int)(b < 4096 ? b : 4096);
  hipLaunchKernelGGL(k_neg_copy, dim3(g), dim3(256), 0, s, src, dst, n);
}

// ---------------------------------------------------------------------------
// element assembly (one colour per launch)
// ---------------------------------------------------------------------------
template <class T, int MODE>
__global__ __launch_bounds__(128) void k_assemble(View v, int e_begin, int e_end, T *K, T *Kb, double *G,
                                                  double *Ke64, double *Rea, double *MM, CellState *cs) {
  const int e = e_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= e_end) return;
  const int c = v.el_cell[e];
  const int n[3] = {v.el_tri[3 * e], v.el_tri[3 * e + 1], v.el_tri[3 * e + 2]};
  if (MODE == ASM_STRAIN) {
    if (!v.cs[c].converged || v.cp[c].law != LAW_UL) return;
    double re[9], mm[3];
    element_strain(v, e, re, mm);
    for (int a = 0; a < 3; ++a) {
      for (int q = 0; q < 3; ++q) Rea[3 * n[a] + q] += re[3 * a + q];
      MM[n[a]] += mm[a];
    }
    return;
  }
  double Ge[6], Ke[36];
  bool ok = element_system<MODE == ASM_FULL>(v, e, Ge, Ke);
  if (!ok) cs[c].degenerate = 1;
  for (int a = 0; a < 3; ++a) {
    G[2 * n[a]] += Ge[2 * a];
    G[2 * n[a] + 1] += Ge[2 * a + 1];
  }
  if (MODE == ASM_FULL && v.cp[c].csr) {
    // block-CSR: the 3 x 3 node blocks of the element
    const int *bp = v.el_bpos + 9LL * e;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        T *B = Kb + 4LL * bp[3 * a + b];
        B[0] += T(Ke[(2 * a) * 6 + 2 * b]);
        B[1] += T(Ke[(2 * a) * 6 + 2 * b + 1]);
        B[2] += T(Ke[(2 * a + 1) * 6 + 2 * b]);
        B[3] += T(Ke[(2 * a + 1) * 6 + 2 * b + 1]);
      }
  } else if (MODE == ASM_FULL) {
    const CellParams &p = v.cp[c];
    T *A = K + p.k_off;
    const int kp[3] = {v.kpos[n[0]], v.kpos[n[1]], v.kpos[n[2]]};
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int pp = 0; pp < 2; ++pp) {
        T *Ar = A + (long long)(2 * kp[a] + pp) * p.npad;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          const int col = 2 * kp[b];
          Ar[col] += T(Ke[(2 * a + pp) * 6 + 2 * b]);
          Ar[col + 1] += T(Ke[(2 * a + pp) * 6 + 2 * b + 1]);
        }
      }
  }
  if (MODE == ASM_FULL && Ke64) {
    double2 *dst = reinterpret_cast<double2 *>(Ke64 + 36 * (long long)e);
#pragma unroll
    for (int k = 0; k < 18; ++k) dst[k] = make_double2(Ke[2 * k], Ke[2 * k + 1]);
  }
}

template <class T>
void launch_assemble(hipStream_t s, const View &v, int e_begin, int e_end, int mode, T *K, T *Kb, double *G,
                     double *Ke64, double *Rea, double *MM) {
  const int n = e_end - e_begin;
  if (n <= 0) return;
  dim3 grid((n + 127) / 128), block(128);
  CellState *cs = const_cast<CellState *>(v.cs);
  if (mode == ASM_FULL)
    hipLaunchKernelGGL((k_assemble<T, ASM_FULL>), grid, block, 0, s, v, e_begin, e_end, K, Kb, G, Ke64, Rea, MM,
                       cs);
  else if (mode == ASM_RESIDUAL)
    hipLaunchKernelGGL((k_assemble<T, ASM_RESIDUAL>), grid, block, 0, s, v, e_begin, e_end, K, Kb, G, Ke64, Rea, MM,
                       cs);
  else
    hipLaunchKernelGGL((k_assemble<T, ASM_STRAIN>), grid, block, 0, s, v, e_begin, e_end, K, Kb, G, Ke64, Rea, MM,
                       cs);
}
template void launch_assemble<float>(hipStream_t, const View &, int, int, int, float *, float *, double *, double *,
                                     double *, double *);
template void launch_assemble<double>(hipStream_t, const View &, int, int, int, double *, double *, double *,
                                      double *, double *, double *);

// ---------------------------------------------------------------------------
// whole-batch assembly in two launches: every element of every colour at once,
// then one thread per node gathering its rows.  A node's rows are written by
// that node's thread only, and it adds its elements' contributions in element
// order -- colour-major, the order the colour-by-colour scatter adds them -- so
// G and K are bit for bit the scatter's.
// ---------------------------------------------------------------------------
template <int MODE>
__global__ __launch_bounds__(128) void k_element_pass(View v, int e_begin, int e_end, double *Ge_out, double *Ke64,
                                                      CellState *cs) {
  const int e = e_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= e_end) return;
  double Ge[6], Ke[36];
  const bool ok = element_system<MODE == ASM_FULL>(v, e, Ge, Ke);
  if (!ok) cs[v.el_cell[e]].degenerate = 1;
  double2 *g = reinterpret_cast<double2 *>(Ge_out + 6LL * e);
#pragma unroll
  for (int k = 0; k < 3; ++k) g[k] = make_double2(Ge[2 * k], Ge[2 * k + 1]);
  if (MODE == ASM_FULL) {
    double2 *dst = reinterpret_cast<double2 *>(Ke64 + 36LL * e);
#pragma unroll
    for (int k = 0; k < 18; ++k) dst[k] = make_double2(Ke[2 * k], Ke[2 * k + 1]);
  }
}

// RESIDUAL mode: each node gathers its two residual rows, element order
__global__ __launch_bounds__(256) void k_node_gather(const int *adj_start, const int *adj, int n0, int n1,
                                                     const double *Ge, double *G) {
  const int n = n0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= n1) return;
  double g0 = G[2 * n], g1 = G[2 * n + 1];
  for (int k = adj_start[n]; k < adj_start[n + 1]; ++k) {
    const int pk = adj[k];
    const int e = pk >> 2, a = pk & 3;
    g0 += Ge[6LL * e + 2 * a];
    g1 += Ge[6LL * e + 2 * a + 1];
  }
  G[2 * n] = g0;
  G[2 * n + 1] = g1;
}

// FULL mode, one thread per group (row node n, column node): the 2 x 2
// block's contributions in element order, the order in which the
// colour-by-colour scatter adds them; the thread of the node's first group
// also gathers the node's residual rows.
template <class T>
__global__ __launch_bounds__(256) void k_group_gather(View v, const int *node_cell, const int *adj_start,
                                                      const int *adj, GatherPlan gp, const double *Ge,
                                                      const double *Ke64, double *G, T *K, T *Kb, int lower_only) {
  const int g = gp.g0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= gp.g1) return;
  const int n = gp.node[g];
  if (g == gp.start[n]) {
    double g0 = G[2 * n], g1 = G[2 * n + 1];
    for (int k = adj_start[n]; k < adj_start[n + 1]; ++k) {
      const int pk = adj[k];
      const int e = pk >> 2, a = pk & 3;
      g0 += Ge[6LL * e + 2 * a];
      g1 += Ge[6LL * e + 2 * a + 1];
    }
    G[2 * n] = g0;
    G[2 * n + 1] = g1;
  }
  const CellParams &p = v.cp[node_cell[n]];
  // lower_only: a symmetric cell's blocks above the diagonal block are
  // written by the fused band LU (U = L^T) before any read; a two-chain
  // cell's blocks that involve its bottom part: those below the diagonal
  if (lower_only && p.sym && !p.csr) {
    const int bi = v.kpos[n] >> 5, bj = v.kpos[gp.col[g]] >> 5;
    if (p.tw_m > 0 && max(bi, bj) >= p.tw_m + p.bw ? bi > bj : bi < bj) return;
  }
  T *b0, *b1;
  if (p.csr) {
    b0 = Kb + 4LL * gp.bpos[g];
    b1 = b0 + 2;
  } else {
    b0 = K + p.k_off + 2LL * v.kpos[n] * p.npad + 2 * v.kpos[gp.col[g]];
    b1 = b0 + p.npad;
  }
  // the block's entries are zero (the fill or the sparse zero) before its one
  // group adds to them: no read.  The contributions four at a time, their
  // offsets and then their values loaded together, added in order.
  T a00 = T(0), a01 = T(0), a10 = T(0), a11 = T(0);
  const int q1 = gp.cstart[g + 1];
  for (int qb = gp.cstart[g]; qb < q1; qb += 4) {
    int o[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) o[u] = qb + u < q1 ? gp.ke[qb + u] : -1;
    double2 x[4], y[4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (o[u] >= 0) {
        x[u] = *reinterpret_cast<const double2 *>(Ke64 + o[u]);
        y[u] = *reinterpret_cast<const double2 *>(Ke64 + o[u] + 6);
      }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (o[u] >= 0) {
        a00 += T(x[u].x);
        a01 += T(x[u].y);
        a10 += T(y[u].x);
        a11 += T(y[u].y);
      }
  }
  b0[0] = a00;
  b0[1] = a01;
  b1[0] = a10;
  b1[1] = a11;
}

template <class T>
void launch_assemble_gather(hipStream_t s, const View &v, int e_begin, int e_end, int mode, const int *node_cell,
                            const int *adj_start, const int *adj, const GatherPlan &gp, int n0, int n1, double *Ge,
                            double *Ke64, double *G, T *K, T *Kb, int lower_only) {
  CellState *cs = const_cast<CellState *>(v.cs);
  const int ne = e_end - e_begin, nn = n1 - n0;
  if (ne > 0) {
    if (mode == ASM_FULL)
      hipLaunchKernelGGL(k_element_pass<ASM_FULL>, dim3((ne + 127) / 128), dim3(128), 0, s, v, e_begin, e_end, Ge,
                         Ke64, cs);
    else
      hipLaunchKernelGGL(k_element_pass<ASM_RESIDUAL>, dim3((ne + 127) / 128), dim3(128), 0, s, v, e_begin, e_end,
                         Ge, Ke64, cs);
  }
  if (nn <= 0) return;
  const int ng = gp.g1 - gp.g0;
  if (mode == ASM_FULL) {
    if (ng > 0)
      hipLaunchKernelGGL((k_group_gather<T>), dim3((ng + 255) / 256), dim3(256), 0, s, v, node_cell, adj_start, adj,
                         gp, Ge, Ke64, G, K, Kb, lower_only);
  } else
    hipLaunchKernelGGL(k_node_gather, dim3((nn + 255) / 256), dim3(256), 0, s, adj_start, adj, n0, n1, Ge, G);
}
template void launch_assemble_gather<float>(hipStream_t, const View &, int, int, int, const int *, const int *,
                                            const int *, const GatherPlan &, int, int, double *, double *, double *,
                                            float *, float *, int);
template void launch_assemble_gather<double>(hipStream_t, const View &, int, int, int, const int *, const int *,
                                             const int *, const GatherPlan &, int, int, double *, double *, double *,
                                             double *, double *, int);

__global__ __launch_bounds__(128) void k_element_dump(View v, int e_begin, int e_end, double *Ge, double *Ke) {
  const int e = e_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= e_end) return;
  double g[6], k[36];
  element_system<true>(v, e, g, k);
  const long long o = e - e_begin;
  for (int i = 0; i < 6; ++i) Ge[6 * o + i] = g[i];
  for (int i = 0; i < 36; ++i) Ke[36 * o + i] = k[i];
}
void launch_element_dump(hipStream_t s, const View &v, int e_begin, int e_end, double *Ge, double *Ke) {
  const int n = e_end - e_begin;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_element_dump, dim3((n + 127) / 128), dim3(128), 0, s, v, e_begin, e_end, Ge, Ke);
}

// ---------------------------------------------------------------------------
// per-cell FP64 norms (fixed-order tree reduction, deterministic)
// which: 0 = residual after assembly (sets gnorm, bnorm and g0 at the first
//        iterate), 1 = refinement residual (rnorm), 2 = residual after the
//        update (gnorm)
// ---------------------------------------------------------------------------
__device__ void norm_finish(CellState &s, int which, double nrm, double lin_rtol) {
  if (which == 0) {
    s.gnorm = nrm;
    s.bnorm = nrm;
    if (s.newton_it == 0) s.g0 = nrm;
  } else if (which == 1) {
    s.rnorm = nrm;
    if (lin_rtol >= 0.0 && !s.ir_done) {  // refinement control (k_ir_control) for this cell
      s.ir_iters += 1;
      s.lin_relres = s.bnorm > 0.0 ? s.rnorm / s.bnorm : 0.0;
      if (s.rnorm <= lin_rtol * s.bnorm) s.ir_done = 1;
    }
  } else {
    s.gnorm = nrm;
  }
}

// One workgroup per cell (blockIdx.x) and chunk of NORM_CHUNK entries
// (blockIdx.y).  A cell of one chunk is summed by its one workgroup; a larger
// cell's chunks write partial sums and the last chunk to arrive adds them in
// chunk order (fixed order: bitwise reproducible).
__global__ __launch_bounds__(256) void k_cell_norm(const CellParams *cp, CellState *cs, const int *cells,
                                                   const double *vec, int which, int only_active, double lin_rtol,
                                                   NormWork nw) {
  const int c = cells[blockIdx.x];
  if (only_active && cs[c].ir_done) return;
  const CellParams &p = cp[c];
  const int nch = nw.part ? (p.ndof + NORM_CHUNK - 1) / NORM_CHUNK : 1;
  const int chunk = blockIdx.y;
  if (chunk >= nch) return;
  const double *v = vec + (long long)p.vec_off;
  const int lo = nch > 1 ? chunk * NORM_CHUNK : 0;
  const int hi = nch > 1 ? min(p.ndof, lo + NORM_CHUNK) : p.ndof;
  // four partial sums per thread: four loads in flight
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int i0 = lo + threadIdx.x; i0 < hi; i0 += 4 * 256) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = i0 + 256 * u;
      const double x = i < hi ? v[i] : 0.0;
      acc[u] += x * x;
    }
  }
  __shared__ double red[256];
  red[threadIdx.x] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  if (nch == 1) {
    norm_finish(cs[c], which, sqrt(red[0]), lin_rtol);
    return;
  }
  double *part = nw.part + (long long)c * nw.stride;
  part[chunk] = red[0];
  const unsigned prev = __hip_atomic_fetch_add(nw.cnt + c, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
  if (prev != (unsigned)(nch - 1)) return;  // not the last chunk of the cell to arrive
  __hip_atomic_store(nw.cnt + c, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  double sum = 0.0;
  for (int k = 0; k < nch; ++k) sum += __hip_atomic_load(part + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  norm_finish(cs[c], which, sqrt(sum), lin_rtol);
}
void launch_cell_norm(hipStream_t s, const CellParams *cp, CellState *cs, const int *cells, int ncell,
                      const double *vec, int which, int only_active, double lin_rtol, const NormWork &nw) {
  if (ncell <= 0) return;
  const int nx = nw.part ? std::max(1, nw.stride) : 1;
  hipLaunchKernelGGL(k_cell_norm, dim3(ncell, nx), dim3(256), 0, s, cp, cs, cells, vec, which, only_active, lin_rtol,
                     nw);
}

// ---------------------------------------------------------------------------
// FP64 refinement residual r = -G - K d, element by element from the stored
// FP64 element matrices, gathered per node over its incident elements in
// element order (deterministic, no atomics)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_matvec_ir(const int *el_tri, const int *node_cell, const int *adj_start,
                                                   const int *adj, const double *Ke64, const double *G,
                                                   const double *delta, double *r, int node_begin, int node_end,
                                                   const CellState *cs, const CellParams *cp, BndView bv) {
  const int i = node_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= node_end) return;
  const int c = node_cell[i];
  if (cs[c].ir_done) return;
  const double2 *d2 = reinterpret_cast<const double2 *>(delta);
  double y0 = 0.0, y1 = 0.0;
  for (int k = adj_start[i]; k < adj_start[i + 1]; ++k) {
    const int pk = adj[k];
    const int e = pk >> 2, a = pk & 3;
    const double *Ke = Ke64 + 36 * (long long)e + 12 * a;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const double2 d = d2[el_tri[3 * e + b]];
      y0 += Ke[2 * b] * d.x + Ke[2 * b + 1] * d.y;
      y1 += Ke[6 + 2 * b] * d.x + Ke[6 + 2 * b + 1] * d.y;
    }
  }
  if (bv.bidx && cp[c].bterm) {
    // node terms: drag diagonal, the boundary vertex's 2 x 6 block and the
    // area term's rank-one part k g_i (sum_m g_m . d_m)
    const CellParams &p = cp[c];
    const double2 di = d2[i];
    if (p.bterm & BT_DRAG) {
      const double cd = p.drag / bv.dt;
      y0 += cd * di.x;
      y1 += cd * di.y;
    }
    const int b = bv.bidx[i];
    if (b >= 0 && (p.bterm & (BT_BEND | BT_AREA))) {
      const int2 fr = bv.frame[b];
      if (fr.x >= 0) {
        const double *J = bv.J + 12LL * b;
        const double2 d1 = d2[fr.x], dd2 = d2[fr.y];
        y0 += J[0] * di.x + J[1] * di.y + J[2] * d1.x + J[3] * d1.y + J[4] * dd2.x + J[5] * dd2.y;
        y1 += J[6] * di.x + J[7] * di.y + J[8] * d1.x + J[9] * d1.y + J[10] * dd2.x + J[11] * dd2.y;
        if (p.bterm & BT_AREA) {
          const double2 g = bv.g[b];
          y0 += p.k_area * g.x * bv.gd[c];
          y1 += p.k_area * g.y * bv.gd[c];
        }
      }
    }
  }
  r[2 * i] = -G[2 * i] - y0;
  r[2 * i + 1] = -G[2 * i + 1] - y1;
}
void launch_matvec_ir(hipStream_t s, const View &v, const int *node_cell, const int *adj_start, const int *adj,
                      const double *Ke64, const double *G, const double *delta, double *r, int node_begin,
                      int node_end, const CellState *cs, const BndView &bv) {
  const int n = node_end - node_begin;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_matvec_ir, dim3((n + 255) / 256), dim3(256), 0, s, v.el_tri, node_cell, adj_start, adj, Ke64,
                     G, delta, r, node_begin, node_end, cs, v.cp, bv);
}

// Dense cells solve in their band-reducing numbering: rp = P r before the
// triangular solves, delta (+)= P^T dp after (cells still refining only).
__global__ __launch_bounds__(256) void k_perm_gather(const int *node_cell, const CellParams *cp, const int *kpos,
                                                     const double *r, double *rp, int n0, int n1) {
  const int n = n0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= n1) return;
  const CellParams &p = cp[node_cell[n]];
  if (p.csr) return;
  const long long o = 2LL * (p.node_off + kpos[n]);
  rp[o] = r[2LL * n];
  rp[o + 1] = r[2LL * n + 1];
}
__global__ __launch_bounds__(256) void k_perm_scatter(const int *node_cell, const CellParams *cp, const CellState *cs,
                                                      const int *kpos, const double *dp, double *delta, int n0, int n1,
                                                      int accumulate) {
  const int n = n0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= n1) return;
  const int c = node_cell[n];
  const CellParams &p = cp[c];
  if (p.csr || cs[c].ir_done) return;
  const long long o = 2LL * (p.node_off + kpos[n]);
  double *d = delta + 2LL * n;
  d[0] = accumulate ? d[0] + dp[o] : dp[o];
  d[1] = accumulate ? d[1] + dp[o + 1] : dp[o + 1];
}
void launch_perm_gather(hipStream_t s, const int *node_cell, const CellParams *cp, const int *kpos, const double *r,
                        double *rp, int n0, int n1) {
  if (n1 <= n0) return;
  hipLaunchKernelGGL(k_perm_gather, dim3((n1 - n0 + 255) / 256), dim3(256), 0, s, node_cell, cp, kpos, r, rp, n0, n1);
}
void launch_perm_scatter(hipStream_t s, const int *node_cell, const CellParams *cp, const CellState *cs,
                         const int *kpos, const double *dp, double *delta, int n0, int n1, int accumulate) {
  if (n1 <= n0) return;
  hipLaunchKernelGGL(k_perm_scatter, dim3((n1 - n0 + 255) / 256), dim3(256), 0, s, node_cell, cp, cs, kpos, dp, delta,
                     n0, n1, accumulate);
}

__global__ void k_ir_control(CellState *cs, const int *cells, int ncell, double lin_rtol, int reset) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncell) return;
  CellState &s = cs[cells[i]];
  if (reset) {
    s.ir_done = 0;
    s.ir_iters = 0;
    s.lin_relres = 0.0;
    return;
  }
  if (s.ir_done) return;
  s.ir_iters += 1;
  s.lin_relres = s.bnorm > 0.0 ? s.rnorm / s.bnorm : 0.0;
  if (s.rnorm <= lin_rtol * s.bnorm) s.ir_done = 1;
}
void launch_ir_control(hipStream_t s, CellState *cs, const int *cells, int ncell, double lin_rtol, int reset) {
  if (ncell <= 0) return;
  hipLaunchKernelGGL(k_ir_control, dim3((ncell + 63) / 64), dim3(64), 0, s, cs, cells, ncell, lin_rtol, reset);
}

__global__ __launch_bounds__(256) void k_update_x(double2 *x, const double *delta, int node_begin, int node_end) {
  const int i = node_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= node_end) return;
  double2 p = x[i];
  p.x += delta[2 * i];
  p.y += delta[2 * i + 1];
  x[i] = p;
}
void launch_update_x(hipStream_t s, double2 *x, const double *delta, int node_begin, int node_end) {
  const int n = node_end - node_begin;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_update_x, dim3((n + 255) / 256), dim3(256), 0, s, x, delta, node_begin, node_end);
}

__global__ void k_newton_control(CellState *cs, const int *cells, int ncell, double rtol, double atol) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncell) return;
  CellState &s = cs[cells[i]];
  s.newton_it += 1;
  s.converged = (s.gnorm <= rtol * s.g0 || s.gnorm <= atol) ? 1 : 0;
  if (!s.ir_done) s.lin_failed = 1;
}
void launch_newton_control(hipStream_t s, CellState *cs, const int *cells, int ncell, double rtol, double atol) {
  if (ncell <= 0) return;
  hipLaunchKernelGGL(k_newton_control, dim3((ncell + 63) / 64), dim3(64), 0, s, cs, cells, ncell, rtol, atol);
}

// ---------------------------------------------------------------------------
// commit of a converged load step
// ---------------------------------------------------------------------------
// UL:607-612 strain update; UL:2046-2171 mean-field bond kinetics with a rigid
// substrate (subs_u = 0): fa_u += x - X, catch-slip k_off, forward Euler c_C,
// dissipative rescale; the new reference configuration is the converged x.
__global__ __launch_bounds__(256) void k_commit_nodes(View v, const int *node_cell, double2 *X, double2 *x, double *ea,
                                                      double *cC, double2 *fau, const double *Rea, const double *MM,
                                                      int node_begin, int node_end, double dt) {
  const int i = node_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= node_end) return;
  const int c = node_cell[i];
  if (!v.cs[c].converged) return;
  const CellParams &p = v.cp[c];
  if (p.law == LAW_UL) {
    const double m = MM[i];
    for (int q = 0; q < 3; ++q) ea[3 * i + q] += Rea[3 * i + q] / m;
  }
  const double2 xi = x[i], Xi = X[i];
  if (p.node_springs) {
    double2 u = fau[i];
    u.x = u.x + (xi.x - Xi.x);
    u.y = u.y + (xi.y - Xi.y);
    if (p.kinetics == KIN_MEANFIELD) {
#pragma clang fp contract(off)
      double force;
      if (p.spring_law == SPRING_LINEAR) {
        force = p.k_lin * sqrt(u.x * u.x + u.y * u.y);  // 2D norm, UL1:2077
      } else {
        double f[2], J[4];
        spring_force(v, p, u.x, u.y, f, J);
        force = sqrt(f[0] * f[0] + f[1] * f[1]);  // UL:2137
      }
      const double k_on = 1.0, ka = 0.004, kb = 10.0, Fa = 15.0, Fb = 15.0;  // UL:2139-2143
      const double k_off = ka * shared_exp(force / Fa) + kb * shared_exp(-force / Fb);
      const double cold = cC[i];
      double cnew = cold + dt * k_on * (1.0 - cold) - dt * k_off * cold;
      if (cnew < 0.0) cnew = 0.0;
      cC[i] = cnew;
      const double keep = cold - dt * k_off * cold;
      if (keep > 0.0 && cnew > 0.0) {
        u.x = u.x * keep / cnew;
        u.y = u.y * keep / cnew;
      } else {
        u.x = 0.0 * u.x;
        u.y = 0.0 * u.y;
      }
    }
    fau[i] = u;
  }
  if (p.law == LAW_UL) X[i] = xi;
}
void launch_commit_nodes(hipStream_t s, const View &v, const int *node_cell, double2 *X, double2 *x, double *ea,
                         double *cC, double2 *fau, const double *Rea, const double *MM, int node_begin, int node_end,
                         double dt) {
  const int n = node_end - node_begin;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_commit_nodes, dim3((n + 255) / 256), dim3(256), 0, s, v, node_cell, X, x, ea, cC, fau, Rea, MM,
                     node_begin, node_end, dt);
}

// Bell kinetics per barycentric spring after a converged load step:
// bound -> unbound with p = 1 - exp(-k_off0 e^{|F(D)|/f_b} dt), D = |u|,
// unbound -> bound
// with p = 1 - exp(-k_on dt), re-anchored at the current attachment point.
// Uniform u = (w + 0.5) 2^-32, w = Philox(ctr = (replica, spring, step, 0)).
__global__ __launch_bounds__(256) void k_commit_springs(View v, const int *sp_cell, const int *sp_elem,
                                                        const int *sp_local, double *state, double *anchor,
                                                        int s_begin, int s_end, double dt, double k_on,
                                                        double k_off0, double f_b, unsigned key_lo, unsigned key_hi,
                                                        const int *replica_of_cell, int *transitions) {
  const int s = s_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= s_end) return;
  const int c = sp_cell[s];
  const CellState &cst = v.cs[c];
  if (!cst.converged) return;
  const CellParams &p = v.cp[c];
  if (p.kinetics != KIN_BELL) return;
  const int e = sp_elem[s];
  const double *bb = v.sp_bary + 3 * s;
  // attachment point (b0 x0 + b1 x1) + b2 x2, rounded term by term without
  // contraction so a re-anchored bond lands on the same bits as a plain FP64
  // evaluation of the same expression
  const double2 x0 = v.x[v.el_tri[3 * e]], x1 = v.x[v.el_tri[3 * e + 1]], x2 = v.x[v.el_tri[3 * e + 2]];
  double px, py;
  {
#pragma clang fp contract(off)
    px = bb[0] * x0.x + bb[1] * x1.x + bb[2] * x2.x;
    py = bb[0] * x0.y + bb[1] * x1.y + bb[2] * x2.y;
  }
  unsigned ctr[4] = {(unsigned)replica_of_cell[c], (unsigned)sp_local[s], (unsigned)cst.step, 0u};
  philox4x32_10(ctr, key_lo, key_hi);
  const double u = ((double)ctr[0] + 0.5) * (1.0 / 4294967296.0);
  // the force and both probabilities rounded as written, with the shared exp
  if (state[s] != 0.0) {
    double ux, uy;
    {
#pragma clang fp contract(off)
      ux = px - anchor[2 * s];
      uy = py - anchor[2 * s + 1];
    }
    const double F = spring_magnitude(v, p, ux, uy);
    const double koff = k_off0 * shared_exp(F / f_b);
    const double prob = 1.0 - shared_exp(-koff * dt);
    if (u < prob) {
      state[s] = 0.0;
      atomicAdd(transitions, 1);
    }
  } else {
    const double prob = 1.0 - shared_exp(-k_on * dt);
    if (u < prob) {
      state[s] = 1.0;
      anchor[2 * s] = px;
      anchor[2 * s + 1] = py;
      atomicAdd(transitions, 1);
    }
  }
}
void launch_commit_springs(hipStream_t s, const View &v, const int *sp_cell, const int *sp_elem, const int *sp_local,
                           double *state, double *anchor, int s_begin, int s_end, double dt, double k_on,
                           double k_off0, double f_b, unsigned key_lo, unsigned key_hi, const int *replica_of_cell,
                           int *transitions) {
  const int n = s_end - s_begin;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_commit_springs, dim3((n + 255) / 256), dim3(256), 0, s, v, sp_cell, sp_elem, sp_local, state,
                     anchor, s_begin, s_end, dt, k_on, k_off0, f_b, key_lo, key_hi, replica_of_cell, transitions);
}

// A remeshed cell's nodes from its old nodes (the host's remeshing maps): new
// node i takes w_i old[src0_i] + (1 - w_i) old[src1_i] of the start-of-step
// and current positions, the strain, the bound fraction and the spring
// extension -- El Topo's midpoint interpolation of a split vertex
// (edgesplitter.cpp:563-577) is w = 1/2, a kept or renumbered vertex
// (defragmentation, surftrack.cpp:487-600) w = 1.  Rounded as written.
__global__ __launch_bounds__(256) void k_remap_nodes(int n_new, const int *src0, const int *src1, const double *w,
                                                     int old_off, const double2 *X, const double2 *x, const double *ea,
                                                     const double *cC, const double2 *fau, double2 *Xn, double2 *xn,
                                                     double *ean, double *cCn, double2 *faun) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_new) return;
  const int a = old_off + src0[i], b = old_off + src1[i];
  const double wa = w[i], wb = 1.0 - w[i];
  auto mix2 = [&](const double2 *v) { return make_double2(wa * v[a].x + wb * v[b].x, wa * v[a].y + wb * v[b].y); };
  Xn[i] = mix2(X);
  xn[i] = mix2(x);
  faun[i] = mix2(fau);
  cCn[i] = wa * cC[a] + wb * cC[b];
  for (int q = 0; q < 3; ++q) ean[3 * i + q] = wa * ea[3 * a + q] + wb * ea[3 * b + q];
}
void launch_remap_nodes(hipStream_t s, int n_new, const int *src0, const int *src1, const double *w, int old_off,
                        const double2 *X, const double2 *x, const double *ea, const double *cC, const double2 *fau,
                        double2 *Xn, double2 *xn, double *ean, double *cCn, double2 *faun) {
  if (n_new <= 0) return;
  hipLaunchKernelGGL(k_remap_nodes, dim3((n_new + 255) / 256), dim3(256), 0, s, n_new, src0, src1, w, old_off, X, x,
                     ea, cC, fau, Xn, xn, ean, cCn, faun);
}

__global__ void k_commit_cells(CellState *cs, const int *cells, int ncell, double dt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncell) return;
  CellState &s = cs[cells[i]];
  if (!s.converged) return;
  s.time += dt;
  s.step += 1;
  s.newton_it = 0;
}
void launch_commit_cells(hipStream_t s, CellState *cs, const int *cells, int ncell, double dt) {
  if (ncell <= 0) return;
  hipLaunchKernelGGL(k_commit_cells, dim3((ncell + 63) / 64), dim3(64), 0, s, cs, cells, ncell, dt);
}

// ---------------------------------------------------------------------------
// static-pivot threshold: tau = 64 eps max_i |K_ii| per cell (the stiffness
// matrices are SPD in the configurations, so only a nearly singular system
// has pivots that small)
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_diag_tau(const CellParams *cp, CellState *cs, const int *cells, const T *K) {
  const int c = cells[blockIdx.x];
  const CellParams &p = cp[c];
  if (p.npad == 0) return;  // sparse cell: no dense system
  const T *A = K + p.k_off;
  double m = 0.0;
  for (int i = threadIdx.x; i < p.ndof; i += 256) m = fmax(m, fabs((double)A[(long long)i * p.npad + i]));
  __shared__ double red[256];
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double eps = sizeof(T) == 4 ? 5.9604644775390625e-08 : 1.1102230246251565e-16;
    cs[c].tau = 64.0 * eps * (red[0] > 0.0 ? red[0] : 1.0);
    cs[c].n_static = 0;
    cs[c].zero_pivot = 0;
  }
}
template <class T>
void launch_diag_tau(hipStream_t s, const CellParams *cp, CellState *cs, const int *cells, int ncell, const T *K) {
  if (ncell <= 0) return;
  hipLaunchKernelGGL(k_diag_tau<T>, dim3(ncell), dim3(256), 0, s, cp, cs, cells, K);
}
template void launch_diag_tau<float>(hipStream_t, const CellParams *, CellState *, const int *, int, const float *);
template void launch_diag_tau<double>(hipStream_t, const CellParams *, CellState *, const int *, int, const double *);

// plain device copy, 16 bytes per thread per iteration (bandwidth probe and
// the known-size reference for the FETCH_SIZE / WRITE_SIZE counters)
__global__ __launch_bounds__(256) void k_copy(const float4 *src, float4 *dst, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
void launch_copy(hipStream_t s, const void *src, void *dst, long long bytes) {
  const long long n = bytes / 16;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, s, static_cast<const float4 *>(src), static_cast<float4 *>(dst),
                     n);
}

}  // namespace wcfe
