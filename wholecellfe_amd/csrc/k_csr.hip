// Sparse path for large cells: block-CSR (2 x 2 blocks, one block row per
// node) assembly target and a batched block-Jacobi preconditioned conjugate
// gradient in the working precision (FP32 for mixed precision), used as the
// inner solver of the FP64 iterative refinement.
//
// Layout: row_ptr / col over the batch's global node numbering (rows of
// dense-solver cells are empty); values T[4] per block, row-major 2 x 2;
// diag_pos[i] = block index of (i, i).  PCG vectors are T[2 N] over the
// batch.  Work is split in chunks of at most CG_ROWS consecutive nodes of one
// cell (element-wise kernels: one thread per node; the SpMV: one workgroup of
// 1024 threads per chunk); every dot product is summed per chunk by a
// fixed-order tree and per cell over its chunks by another, so the iteration
// is bitwise reproducible run to run.
#include "wcfe_internal.hpp"

namespace wcfe {

template <class T> struct V2;
template <> struct V2<float> { typedef float2 type; };
template <> struct V2<double> { typedef double2 type; };
template <class T> struct V4;
template <> struct V4<float> { typedef float4 type; };
template <> struct V4<double> { typedef double4 type; };

template <class T>
__global__ __launch_bounds__(256) void k_zero_t(T *p, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = T(0);
}
template <class T> void launch_zero(hipStream_t s, T *p, long long n) {
  if (n <= 0) return;
  const long long b = (n + 255) / 256;
  hipLaunchKernelGGL(k_zero_t<T>, dim3((int)(b < 4096 ? b : 4096)), dim3(256), 0, s, p, n);
}

// inverse of each 2 x 2 diagonal block (block-Jacobi preconditioner)
template <class T>
__global__ __launch_bounds__(CG_ROWS) void k_bsr_diag_inv(const int4 *chunks, const int *diag_pos, const T *val,
                                                          T *dinv) {
  const int4 ch = chunks[blockIdx.x];
  const int i = ch.y + threadIdx.x;
  if (i >= ch.z) return;
  const T *B = val + 4LL * diag_pos[i];
  const T a = B[0], b = B[1], c = B[2], d = B[3];
  const T det = a * d - b * c;
  const T id = det != T(0) ? T(1) / det : T(0);
  T *o = dinv + 4LL * i;
  o[0] = d * id;
  o[1] = -b * id;
  o[2] = -c * id;
  o[3] = a * id;
}

template <class T>
__device__ __forceinline__ void apply2(const T *M, T x0, T x1, T &y0, T &y1) {
  y0 = M[0] * x0 + M[1] * x1;
  y1 = M[2] * x0 + M[3] * x1;
}

// Preconditioned conjugate gradients in the single-reduction form of
// Chronopoulos and Gear ("s-step iterative methods for symmetric linear
// systems", J. Comput. Appl. Math. 25, 1989): with u = M^-1 r and w = K u,
// gamma = r.u and delta = w.u come from ONE reduction per iteration, and
//   beta = gamma_new / gamma,  alpha = gamma_new / (delta - beta gamma_new / alpha)
//   p = u + beta p,  s = w + beta s  (s = K p),  x += alpha p,  r -= alpha s
// -- the same iterates as the two-reduction form in exact arithmetic.  Per
// iteration two launches: the update kernel, whose workgroups each add up
// the cell's partial dot products of the last SpMV (one triple per chunk) and
// form the scalars themselves (all the same bits: same sums in the same
// order) before the element-wise update of p, s, x, r, u; then one SpMV with
// the partial dot products (r.u, w.u, r.r).  The scalar state is kept twice
// (two CgState arrays, by parity): an update kernel reads one and its first
// chunk's workgroup writes the other, so no workgroup reads what another
// writes.  A one-workgroup scalar kernel forms the scalars at the start of a
// solve and at each host convergence check.

// r = rhs (FP64 -> T), x = 0, u = M^-1 r, p = s = 0 (so the first update
// gives p = u, s = w)
template <class T>
__global__ __launch_bounds__(CG_ROWS) void k_pcg_init(const int4 *chunks, const CellState *cs, const double *rhs,
                                                      const T *dinv, T *x, T *r, T *u, T *p, T *sv) {
  const int4 ch = chunks[blockIdx.x];
  if (cs[ch.x].ir_done) return;
  const int i = ch.y + threadIdx.x;
  if (i >= ch.z) return;
  const long long k = 2LL * i;
  const T r0 = (T)rhs[k], r1 = (T)rhs[k + 1];
  T u0, u1;
  apply2(dinv + 4LL * i, r0, r1, u0, u1);
  r[k] = r0; r[k + 1] = r1;
  u[k] = u0; u[k + 1] = u1;
  x[k] = T(0); x[k + 1] = T(0);
  p[k] = T(0); p[k + 1] = T(0);
  sv[k] = T(0); sv[k + 1] = T(0);
}

// w = K u and the partial sums (r.u, w.u, r.r) of one chunk (CG_ROWS rows),
// one workgroup of SPMV_THREADS.  LANES lanes per block row: lane (row, slot)
// takes the row's blocks slot, slot + LANES, ..., so a wave reads its rows'
// blocks as one contiguous range; SPMV_THREADS / LANES rows per pass,
// LANES / 4 passes.  The lane sums: xor butterfly (every lane gets the same
// bits); the partial sums: per thread over its passes in order, xor butterfly
// over a wave's rows, then the waves by a fixed tree (bitwise reproducible).
constexpr int SPMV_THREADS = 1024;
template <class T, int LANES>
__global__ __launch_bounds__(SPMV_THREADS) void k_pcg_spmv_dot(const int4 *chunks, int cbase, const CgState *cg,
                                                               const int *row_ptr, const int *col, const T *val,
                                                               const T *r, const T *u, T *w, double *partials,
                                                               long long pstride) {
  constexpr int NW = SPMV_THREADS / 64, RPP = SPMV_THREADS / LANES;
  static_assert(CG_ROWS % RPP == 0, "a chunk is a whole number of passes");
  __shared__ double red[3][NW];
  const int4 ch = chunks[blockIdx.x];
  if (cg[ch.x].done) return;
  typedef typename V2<T>::type v2;
  typedef typename V4<T>::type v4;
  const v2 *u2 = reinterpret_cast<const v2 *>(u);
  const v2 *r2 = reinterpret_cast<const v2 *>(r);
  const v4 *B4 = reinterpret_cast<const v4 *>(val);
  v2 *w2 = reinterpret_cast<v2 *>(w);
  const int lane = threadIdx.x & 63, slot = threadIdx.x & (LANES - 1), wave = threadIdx.x >> 6;
  double ru = 0.0, wu = 0.0, rr = 0.0;
#pragma unroll
  for (int pass = 0; pass < CG_ROWS / RPP; ++pass) {
    const int i = ch.y + pass * RPP + threadIdx.x / LANES;
    T y0 = T(0), y1 = T(0);
    if (i < ch.z) {
      const int ke = row_ptr[i + 1];
      for (int k = row_ptr[i] + slot; k < ke; k += LANES) {
        const v2 uj = u2[col[k]];
        const v4 B = B4[k];
        y0 += B.x * uj.x + B.y * uj.y;
        y1 += B.z * uj.x + B.w * uj.y;
      }
    }
#pragma unroll
    for (int o = 1; o < LANES; o <<= 1) {
      y0 += __shfl_xor(y0, o);
      y1 += __shfl_xor(y1, o);
    }
    if (slot == 0 && i < ch.z) {
      v2 wi;
      wi.x = y0;
      wi.y = y1;
      w2[i] = wi;
      const v2 ui = u2[i], ri = r2[i];
      ru += (double)(ri.x * ui.x + ri.y * ui.y);
      wu += (double)(y0 * ui.x + y1 * ui.y);
      rr += (double)(ri.x * ri.x + ri.y * ri.y);
    }
  }
#pragma unroll
  for (int o = LANES; o < 64; o <<= 1) {
    ru += __shfl_xor(ru, o);
    wu += __shfl_xor(wu, o);
    rr += __shfl_xor(rr, o);
  }
  if (lane == 0) {
    red[0][wave] = ru;
    red[1][wave] = wu;
    red[2][wave] = rr;
  }
  __syncthreads();
  if (threadIdx.x < 3) {  // three arrays of pstride partial sums, one per chunk
    double *v = red[threadIdx.x];
    double s8[NW / 2];
#pragma unroll
    for (int q = 0; q < NW / 2; ++q) s8[q] = v[2 * q] + v[2 * q + 1];
#pragma unroll
    for (int h = NW / 4; h >= 1; h >>= 1)
#pragma unroll
      for (int q = 0; q < h; ++q) s8[q] = s8[2 * q] + s8[2 * q + 1];
    partials[threadIdx.x * pstride + cbase + blockIdx.x] = s8[0];
  }
}

// The scalars of one step from the sums of the last SpMV: convergence
// ||r|| <= tol ||r0|| (r of the update just made; tol: g.tol) or the iteration cap;
// alpha and beta of the next update.  A cell already done carries its state.
__device__ __forceinline__ void pcg_step_scalars(CgState &g, double gamma, double delta, double rr, double rtol,
                                                 int max_iter) {
  if (g.done) return;
  g.iters += 1;
  g.rr = rr;
  (void)rtol;
  if (rr <= g.tol * g.tol * g.rr0 || g.iters >= max_iter || !(rr == rr)) {
    g.done = 1;
    return;
  }
  const double beta = g.rz > 0.0 ? gamma / g.rz : 0.0;
  const double den = delta - beta * gamma / g.alpha;
  g.pq = den;
  if (!(den > 0.0)) {  // the step's curvature p.Kp <= 0
    g.breakdown = 1;
    g.done = 1;
    return;
  }
  g.beta = beta;
  g.alpha = gamma / den;
  g.rz = gamma;
}

// the start of a solve: gamma, delta of the initial u; alpha = gamma / delta,
// beta = 0; done if refinement is finished or r = 0
__device__ __forceinline__ void pcg_init_scalars(CgState &g, double gamma, double delta, double rr, double tol) {
  g.tol = tol;
  g.iters = 0;
  g.breakdown = 0;
  g.rr = g.rr0 = rr;
  g.rz = gamma;
  g.beta = 0.0;
  g.done = rr > 0.0 ? 0 : 1;
  if (!g.done) {
    if (delta > 0.0) {
      g.alpha = gamma / delta;
    } else {  // K u . u <= 0: the system is not positive definite along u
      g.alpha = 0.0;
      g.breakdown = 1;
      g.done = 1;
    }
  }
}

// a cell's three sums over its chunks' partials by NT threads: strided
// per-thread sums, then a fixed tree (the same bits in every workgroup)
template <int NT>
__device__ __forceinline__ void sum_partials(const double *o, long long pstride, int n, double (*red)[NT],
                                             double &s0, double &s1, double &s2) {
  s0 = s1 = s2 = 0.0;
  for (int k = threadIdx.x; k < n; k += NT) {
    s0 += o[k];
    s1 += o[pstride + k];
    s2 += o[2 * pstride + k];
  }
  red[0][threadIdx.x] = s0;
  red[1][threadIdx.x] = s1;
  red[2][threadIdx.x] = s2;
  __syncthreads();
  for (int h = NT / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      red[0][threadIdx.x] += red[0][threadIdx.x + h];
      red[1][threadIdx.x] += red[1][threadIdx.x + h];
      red[2][threadIdx.x] += red[2][threadIdx.x + h];
    }
    __syncthreads();
  }
  s0 = red[0][0];
  s1 = red[1][0];
  s2 = red[2][0];
}

// STEP: the scalars from the last SpMV's partials (every workgroup of the
// cell; the cell's first chunk writes them to gout), then the update
//   p = u + beta p, s = w + beta s, x += alpha p, r -= alpha s, u = M^-1 r;
// otherwise the update with gin's scalars (those of the start of the solve
// or of the last check)
template <class T, bool STEP>
__global__ __launch_bounds__(CG_ROWS) void k_pcg_update(const int4 *chunks, int cbase, const CgState *gin,
                                                        CgState *gout, const double *partials, long long pstride,
                                                        double rtol, int max_iter, const T *dinv, T *p, T *sv,
                                                        const T *w, T *x, T *r, T *u) {
  __shared__ double red[3][CG_ROWS];
  typedef typename V2<T>::type v2;
  typedef typename V4<T>::type v4;
  const int4 ch = chunks[blockIdx.x];
  CgState g = gin[ch.x];
  if (g.done && !STEP) return;
  // the node's vectors first: their loads do not wait for the scalars
  const int i = ch.y + threadIdx.x;
  const bool own = i < ch.z && !g.done;
  v2 pv{}, sv2{}, uv{}, wv{}, xv{}, rv{};
  v4 dv{};
  if (own) {
    pv = reinterpret_cast<const v2 *>(p)[i];
    sv2 = reinterpret_cast<const v2 *>(sv)[i];
    uv = reinterpret_cast<const v2 *>(u)[i];
    wv = reinterpret_cast<const v2 *>(w)[i];
    xv = reinterpret_cast<const v2 *>(x)[i];
    rv = reinterpret_cast<const v2 *>(r)[i];
    dv = reinterpret_cast<const v4 *>(dinv)[i];
  }
  if (STEP) {
    if (!g.done) {
      double s0, s1, s2;
      sum_partials<CG_ROWS>(partials + g.chunk_off, pstride, g.nchunk, red, s0, s1, s2);
      pcg_step_scalars(g, s0, s1, s2, rtol, max_iter);
    }
    if (cbase + (int)blockIdx.x == g.chunk_off && threadIdx.x == 0) gout[ch.x] = g;
  }
  if (g.done || i >= ch.z) return;
  const T alpha = (T)g.alpha, beta = (T)g.beta;
  const T p0 = uv.x + beta * pv.x, p1 = uv.y + beta * pv.y;
  const T s0 = wv.x + beta * sv2.x, s1 = wv.y + beta * sv2.y;
  reinterpret_cast<v2 *>(p)[i] = v2{p0, p1};
  reinterpret_cast<v2 *>(sv)[i] = v2{s0, s1};
  reinterpret_cast<v2 *>(x)[i] = v2{xv.x + alpha * p0, xv.y + alpha * p1};
  const T r0 = rv.x - alpha * s0, r1 = rv.y - alpha * s1;
  reinterpret_cast<v2 *>(r)[i] = v2{r0, r1};
  const T d[4] = {dv.x, dv.y, dv.z, dv.w};
  T u0, u1;
  apply2(d, r0, r1, u0, u1);
  reinterpret_cast<v2 *>(u)[i] = v2{u0, u1};
}

// a new solve: cells still refining run it, the others stay done
__global__ void k_pcg_reset_done(const int *cells, int ncell, const CellState *cs, CgState *cg) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ncell) cg[cells[i]].done = cs[cells[i]].ir_done ? 1 : 0;
}

// Per-cell scalars, one workgroup of 1024 threads per sparse cell of the
// list, from the partial sums of the cell's chunks (three arrays, read
// coalesced) by a fixed-order tree; gin -> gout (the same array for CG_INIT).
constexpr int SCALAR_THREADS = 1024;
__global__ __launch_bounds__(SCALAR_THREADS) void k_pcg_scalars(const int *cells, const CellState *cs,
                                                                const CgState *gin, CgState *gout,
                                                                const double *partials, long long pstride, int mode,
                                                                double rtol, double lin_rtol, int max_iter) {
  __shared__ double red[3][SCALAR_THREADS];
  const int c = cells[blockIdx.x];
  CgState g = gin[c];
  const bool skip = mode == CG_INIT ? cs[c].ir_done != 0 : g.done != 0;
  if (skip) {
    if (threadIdx.x == 0) {
      if (mode == CG_INIT) {  // no solve for this cell: no iterations either
        g.done = 1;
        g.iters = 0;
      }
      gout[c] = g;
    }
    return;
  }
  double s0, s1, s2;
  sum_partials<SCALAR_THREADS>(partials + g.chunk_off, pstride, g.nchunk, red, s0, s1, s2);
  if (threadIdx.x != 0) return;
  if (mode == CG_INIT) {
    // a later refinement sweep needs only the rest of the way to lin_rtol
    // (a tenth of it, for the FP32 system's own error): the solve's
    // tolerance from the cell's relative residual so far (0 before the
    // first sweep)
    const double lr = cs[c].lin_relres;
    double tol = rtol;
    if (lr > 0.0 && lin_rtol > 0.0) tol = fmin(0.1, fmax(rtol, 0.1 * lin_rtol / lr));
    pcg_init_scalars(g, s0, s1, s2, tol);
  } else {
    pcg_step_scalars(g, s0, s1, s2, rtol, max_iter);
  }
  gout[c] = g;
}

// delta (FP64) = x, or += x for the later refinement sweeps
template <class T>
__global__ __launch_bounds__(CG_ROWS) void k_pcg_finish(const int4 *chunks, const CellState *cs, const T *x,
                                                        double *delta, int accumulate) {
  const int4 ch = chunks[blockIdx.x];
  if (cs[ch.x].ir_done) return;
  const int i = ch.y + threadIdx.x;
  if (i >= ch.z) return;
  for (int k = 0; k < 2; ++k) {
    const long long j = 2LL * i + k;
    delta[j] = accumulate ? delta[j] + (double)x[j] : (double)x[j];
  }
}

template <class T>
void launch_bsr_diag_inv(hipStream_t s, const int4 *chunks, int nchunk, const int *diag_pos, const T *val, T *dinv) {
  if (nchunk <= 0) return;
  hipLaunchKernelGGL(k_bsr_diag_inv<T>, dim3(nchunk), dim3(CG_ROWS), 0, s, chunks, diag_pos, val, dinv);
}

template <class T>
static void launch_spmv_dot(hipStream_t s, const int4 *chunks, int cbase, int nchunk, const CgState *cg,
                            const int *row_ptr, const int *col, const T *val, const CgVectors<T> &v,
                            double *partials, int lanes) {
  const dim3 g(nchunk), b(SPMV_THREADS);
  if (lanes == 4)
    hipLaunchKernelGGL((k_pcg_spmv_dot<T, 4>), g, b, 0, s, chunks, cbase, cg, row_ptr, col, val, v.r, v.z, v.q,
                       partials, v.pstride);
  else if (lanes == 16)
    hipLaunchKernelGGL((k_pcg_spmv_dot<T, 16>), g, b, 0, s, chunks, cbase, cg, row_ptr, col, val, v.r, v.z, v.q,
                       partials, v.pstride);
  else
    hipLaunchKernelGGL((k_pcg_spmv_dot<T, 8>), g, b, 0, s, chunks, cbase, cg, row_ptr, col, val, v.r, v.z, v.q,
                       partials, v.pstride);
}

template <class T>
void launch_pcg_init(hipStream_t s, const int4 *chunks, int cbase, int nchunk, const int *cells, int ncell,
                     const CellState *cs, CgState *cg, const int *row_ptr, const int *col, const T *val,
                     const double *rhs, const CgVectors<T> &v, double *partials, double rtol, double lin_rtol,
                     int max_iter, int lanes) {
  if (nchunk <= 0) return;
  hipLaunchKernelGGL(k_pcg_reset_done, dim3((ncell + 63) / 64), dim3(64), 0, s, cells, ncell, cs, cg);
  hipLaunchKernelGGL(k_pcg_init<T>, dim3(nchunk), dim3(CG_ROWS), 0, s, chunks + cbase, cs, rhs, v.dinv, v.x, v.r,
                     v.z, v.p, v.s);
  // w = K u and the partials of the first step
  launch_spmv_dot<T>(s, chunks + cbase, cbase, nchunk, cg, row_ptr, col, val, v, partials, lanes);
  hipLaunchKernelGGL(k_pcg_scalars, dim3(ncell), dim3(SCALAR_THREADS), 0, s, cells, cs, cg, cg, partials, v.pstride,
                     (int)CG_INIT, rtol, lin_rtol, max_iter);
}

template <class T>
void launch_pcg_iteration(hipStream_t s, const int4 *chunks, int cbase, int nchunk, const CgState *gin,
                          CgState *gout, int step, const int *row_ptr, const int *col, const T *val,
                          const CgVectors<T> &v, double *partials, double rtol, int max_iter, int lanes) {
  if (nchunk <= 0) return;
  chunks += cbase;
  if (step)
    hipLaunchKernelGGL((k_pcg_update<T, true>), dim3(nchunk), dim3(CG_ROWS), 0, s, chunks, cbase, gin, gout,
                       partials, v.pstride, rtol, max_iter, v.dinv, v.p, v.s, v.q, v.x, v.r, v.z);
  else
    hipLaunchKernelGGL((k_pcg_update<T, false>), dim3(nchunk), dim3(CG_ROWS), 0, s, chunks, cbase, gin, gout,
                       partials, v.pstride, rtol, max_iter, v.dinv, v.p, v.s, v.q, v.x, v.r, v.z);
  launch_spmv_dot<T>(s, chunks, cbase, nchunk, step ? gout : gin, row_ptr, col, val, v, partials, lanes);
}

void launch_pcg_check(hipStream_t s, const int *cells, int ncell, const CellState *cs, const CgState *gin,
                      CgState *gout, const double *partials, long long pstride, double rtol, int max_iter) {
  if (ncell <= 0) return;
  hipLaunchKernelGGL(k_pcg_scalars, dim3(ncell), dim3(SCALAR_THREADS), 0, s, cells, cs, gin, gout, partials, pstride,
                     (int)CG_STEP, rtol, 0.0, max_iter);
}

template <class T>
void launch_pcg_finish(hipStream_t s, const int4 *chunks, int nchunk, const CellState *cs, const T *x, double *delta,
                       int accumulate) {
  if (nchunk <= 0) return;
  hipLaunchKernelGGL(k_pcg_finish<T>, dim3(nchunk), dim3(CG_ROWS), 0, s, chunks, cs, x, delta, accumulate);
}

#define WCFE_CSR_INST(T)                                                                                          \
  template void launch_zero<T>(hipStream_t, T *, long long);                                                      \
  template void launch_bsr_diag_inv<T>(hipStream_t, const int4 *, int, const int *, const T *, T *);              \
  template void launch_pcg_init<T>(hipStream_t, const int4 *, int, int, const int *, int, const CellState *,       \
                                   CgState *, const int *, const int *, const T *, const double *,                \
                                   const CgVectors<T> &, double *, double, double, int, int);                     \
  template void launch_pcg_iteration<T>(hipStream_t, const int4 *, int, int, const CgState *, CgState *, int,     \
                                        const int *, const int *, const T *, const CgVectors<T> &, double *,      \
                                        double, int, int);                                                        \
  template void launch_pcg_finish<T>(hipStream_t, const int4 *, int, const CellState *, const T *, double *, int);
WCFE_CSR_INST(float)
WCFE_CSR_INST(double)

}  // namespace wcfe
