// Sparse path for large cells: block-CSR (2 x 2 blocks, one block row per
// node) assembly target and a batched block-Jacobi preconditioned conjugate
// gradient in the working precision (FP32 for mixed precision), used as the
// inner solver of the FP64 iterative refinement.
//
// Layout: row_ptr / col over the batch's global node numbering (rows of
// dense-solver cells are empty); values T[4] per block, row-major 2 x 2;
// diag_pos[i] = block index of (i, i).  PCG vectors are T[2 N] over the
// batch.  Work is split in chunks of at most CG_ROWS consecutive nodes of one
// cell (one thread per node); every dot product is summed per chunk by a
// fixed-order tree and per cell over its chunks in order, so the iteration is
// bitwise reproducible run to run.
#include "wcfe_internal.hpp"

namespace wcfe {

template <class T> struct V2;
template <> struct V2<float> { typedef float2 type; };
template <> struct V2<double> { typedef double2 type; };

// fixed-order tree sum over the 256 threads of the workgroup
__device__ __forceinline__ double block_sum256(double v, double *red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

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

// r = rhs (FP64 -> T), x = 0, z = M^-1 r, p = z; partials (r.z, r.r)
template <class T>
__global__ __launch_bounds__(CG_ROWS) void k_pcg_init(const int4 *chunks, int cbase, const CellState *cs,
                                                      const double *rhs, const T *dinv, T *x, T *r, T *z, T *p,
                                                      double *partials) {
  __shared__ double red[CG_ROWS];
  const int4 ch = chunks[blockIdx.x];
  if (cs[ch.x].ir_done) return;
  const int i = ch.y + threadIdx.x;
  double rz = 0.0, rr = 0.0;
  if (i < ch.z) {
    const T r0 = (T)rhs[2LL * i], r1 = (T)rhs[2LL * i + 1];
    T z0, z1;
    apply2(dinv + 4LL * i, r0, r1, z0, z1);
    r[2LL * i] = r0; r[2LL * i + 1] = r1;
    z[2LL * i] = z0; z[2LL * i + 1] = z1;
    p[2LL * i] = z0; p[2LL * i + 1] = z1;
    x[2LL * i] = T(0); x[2LL * i + 1] = T(0);
    rz = (double)(r0 * z0 + r1 * z1);
    rr = (double)(r0 * r0 + r1 * r1);
  }
  rz = block_sum256(rz, red);
  rr = block_sum256(rr, red);
  if (threadIdx.x == 0) {
    partials[2LL * (cbase + blockIdx.x)] = rz;
    partials[2LL * (cbase + blockIdx.x) + 1] = rr;
  }
}

// q = K p; partial p.q
template <class T>
__global__ __launch_bounds__(CG_ROWS) void k_pcg_spmv_dot(const int4 *chunks, int cbase, const CgState *cg,
                                                          const int *row_ptr, const int *col, const T *val,
                                                          const T *p, T *q, double *partials) {
  __shared__ double red[CG_ROWS];
  const int4 ch = chunks[blockIdx.x];
  if (cg[ch.x].done) return;
  const int i = ch.y + threadIdx.x;
  double pq = 0.0;
  if (i < ch.z) {
    typedef typename V2<T>::type v2;
    const v2 *p2 = reinterpret_cast<const v2 *>(p);
    T y0 = T(0), y1 = T(0);
    for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
      const v2 pj = p2[col[k]];
      const T *B = val + 4LL * k;
      y0 += B[0] * pj.x + B[1] * pj.y;
      y1 += B[2] * pj.x + B[3] * pj.y;
    }
    q[2LL * i] = y0;
    q[2LL * i + 1] = y1;
    const v2 pi = p2[i];
    pq = (double)(pi.x * y0 + pi.y * y1);
  }
  pq = block_sum256(pq, red);
  if (threadIdx.x == 0) partials[2LL * (cbase + blockIdx.x)] = pq;
}

// x += alpha p, r -= alpha q, z = M^-1 r; partials (r.z, r.r)
template <class T>
__global__ __launch_bounds__(CG_ROWS) void k_pcg_update(const int4 *chunks, int cbase, const CgState *cg,
                                                        const T *dinv, const T *p, const T *q, T *x, T *r, T *z,
                                                        double *partials) {
  __shared__ double red[CG_ROWS];
  const int4 ch = chunks[blockIdx.x];
  if (cg[ch.x].done) return;
  const T alpha = (T)cg[ch.x].alpha;
  const int i = ch.y + threadIdx.x;
  double rz = 0.0, rr = 0.0;
  if (i < ch.z) {
    const long long k = 2LL * i;
    x[k] += alpha * p[k];
    x[k + 1] += alpha * p[k + 1];
    const T r0 = r[k] - alpha * q[k], r1 = r[k + 1] - alpha * q[k + 1];
    r[k] = r0;
    r[k + 1] = r1;
    T z0, z1;
    apply2(dinv + 4LL * i, r0, r1, z0, z1);
    z[k] = z0;
    z[k + 1] = z1;
    rz = (double)(r0 * z0 + r1 * z1);
    rr = (double)(r0 * r0 + r1 * r1);
  }
  rz = block_sum256(rz, red);
  rr = block_sum256(rr, red);
  if (threadIdx.x == 0) {
    partials[2LL * (cbase + blockIdx.x)] = rz;
    partials[2LL * (cbase + blockIdx.x) + 1] = rr;
  }
}

// p = z + beta p
template <class T>
__global__ __launch_bounds__(CG_ROWS) void k_pcg_paxpy(const int4 *chunks, const CgState *cg, const T *z, T *p) {
  const int4 ch = chunks[blockIdx.x];
  if (cg[ch.x].done) return;
  const T beta = (T)cg[ch.x].beta;
  const int i = ch.y + threadIdx.x;
  if (i >= ch.z) return;
  p[2LL * i] = z[2LL * i] + beta * p[2LL * i];
  p[2LL * i + 1] = z[2LL * i + 1] + beta * p[2LL * i + 1];
}

// Per-cell scalars, one workgroup per sparse cell of the list; partial sums
// of the cell's chunks in chunk order.
//   CG_INIT:  rz, rr0 from (r.z, r.r); done if refinement is finished or r = 0
//   CG_ALPHA: alpha = rz / p.q
//   CG_BETA:  beta = rz_new / rz; convergence ||r|| <= rtol ||r0|| or iteration cap
__global__ __launch_bounds__(256) void k_pcg_scalars(const int *cells, const CellState *cs, CgState *cg,
                                                     const double *partials, int mode, double rtol, int max_iter) {
  __shared__ double red[256];
  CgState &g = cg[cells[blockIdx.x]];
  const bool skip = mode == CG_INIT ? cs[cells[blockIdx.x]].ir_done != 0 : g.done != 0;
  if (skip) {
    if (mode == CG_INIT && threadIdx.x == 0) g.done = 1;
    return;
  }
  double s0 = 0.0, s1 = 0.0;
  for (int k = threadIdx.x; k < g.nchunk; k += 256) {
    s0 += partials[2LL * (g.chunk_off + k)];
    s1 += partials[2LL * (g.chunk_off + k) + 1];
  }
  s0 = block_sum256(s0, red);
  if (mode != CG_ALPHA) s1 = block_sum256(s1, red);
  if (threadIdx.x != 0) return;
  if (mode == CG_INIT) {
    g.rz = s0;
    g.rr = g.rr0 = s1;
    g.iters = 0;
    g.breakdown = 0;
    g.done = s1 > 0.0 ? 0 : 1;
  } else if (mode == CG_ALPHA) {
    g.pq = s0;
    if (s0 > 0.0) {
      g.alpha = g.rz / s0;
    } else {  // K p . p <= 0: the system is not positive definite along p
      g.alpha = 0.0;
      g.breakdown = 1;
      g.done = 1;
    }
  } else {
    g.iters += 1;
    g.beta = g.rz > 0.0 ? s0 / g.rz : 0.0;
    g.rz = s0;
    g.rr = s1;
    if (s1 <= rtol * rtol * g.rr0 || g.iters >= max_iter || !(s1 == s1)) g.done = 1;
  }
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
void launch_pcg_init(hipStream_t s, const int4 *chunks, int cbase, int nchunk, const CellState *cs,
                     const double *rhs, const CgVectors<T> &v, double *partials) {
  if (nchunk <= 0) return;
  hipLaunchKernelGGL(k_pcg_init<T>, dim3(nchunk), dim3(CG_ROWS), 0, s, chunks + cbase, cbase, cs, rhs, v.dinv, v.x,
                     v.r, v.z, v.p, partials);
}

template <class T>
void launch_pcg_iteration(hipStream_t s, const int4 *chunks, int cbase, int nchunk, const int *cells, int ncell,
                          const CellState *cs, CgState *cg, const int *row_ptr, const int *col, const T *val,
                          const CgVectors<T> &v, double *partials, double rtol, int max_iter) {
  if (nchunk <= 0) return;
  chunks += cbase;
  hipLaunchKernelGGL(k_pcg_spmv_dot<T>, dim3(nchunk), dim3(CG_ROWS), 0, s, chunks, cbase, cg, row_ptr, col, val,
                     v.p, v.q, partials);
  hipLaunchKernelGGL(k_pcg_scalars, dim3(ncell), dim3(256), 0, s, cells, cs, cg, partials, (int)CG_ALPHA, rtol,
                     max_iter);
  hipLaunchKernelGGL(k_pcg_update<T>, dim3(nchunk), dim3(CG_ROWS), 0, s, chunks, cbase, cg, v.dinv, v.p, v.q, v.x,
                     v.r, v.z, partials);
  hipLaunchKernelGGL(k_pcg_scalars, dim3(ncell), dim3(256), 0, s, cells, cs, cg, partials, (int)CG_BETA, rtol,
                     max_iter);
  hipLaunchKernelGGL(k_pcg_paxpy<T>, dim3(nchunk), dim3(CG_ROWS), 0, s, chunks, cg, v.z, v.p);
}

void launch_pcg_scalars(hipStream_t s, const int *cells, int ncell, const CellState *cs, CgState *cg,
                        const double *partials, int mode, double rtol, int max_iter) {
  if (ncell <= 0) return;
  hipLaunchKernelGGL(k_pcg_scalars, dim3(ncell), dim3(256), 0, s, cells, cs, cg, partials, mode, rtol, max_iter);
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
  template void launch_pcg_init<T>(hipStream_t, const int4 *, int, int, const CellState *, const double *,        \
                                   const CgVectors<T> &, double *);                                               \
  template void launch_pcg_iteration<T>(hipStream_t, const int4 *, int, int, const int *, int, const CellState *, \
                                        CgState *, const int *, const int *, const T *, const CgVectors<T> &,     \
                                        double *, double, int);                                                   \
  template void launch_pcg_finish<T>(hipStream_t, const int4 *, int, const CellState *, const T *, double *, int);
WCFE_CSR_INST(float)
WCFE_CSR_INST(double)

}  // namespace wcfe
