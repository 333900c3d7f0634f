// Batched dense blocked LU for gfx950: right-looking, 64 x 64 blocks,
// row-major matrices padded to a multiple of 64 (identity on the padding).
//
// Per block step k, over every cell of the work list at once:
//   static pivoting : k_lu_diag (one workgroup factors the diagonal block in
//                     LDS; pivots below tau = sqrt(eps) max|K_ii| are
//                     replaced by +-tau, Li & Demmel's static pivoting, and
//                     the error is removed by FP64 iterative refinement)
//                     -> k_lu_trsm (U row panel and L column panel)
//   partial pivoting: k_lu_panel_pp (one workgroup factors the 64-wide
//                     column panel with row interchanges, LAPACK getf2
//                     semantics inside the panel) -> k_lu_laswp (interchanges
//                     applied to the trailing columns only, LINPACK style)
//                     -> k_lu_trsm (U row panel)
//   then            : k_lu_gemm (trailing update A_ij -= L_ik U_kj on MFMA:
//                     v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64)
// k_lu_solve applies the factors to one right-hand side per cell with the
// whole vector resident in LDS.
#include "wcfe_internal.hpp"

namespace wcfe {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int find_item_cell(const int *prefix, int ncell, int item) {
  int lo = 0, hi = ncell;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prefix[mid] <= item) lo = mid; else hi = mid;
  }
  return lo;
}

template <class T> struct Pitch { static constexpr int v = 68; };
template <> struct Pitch<double> { static constexpr int v = 66; };

// 64 x 64 block (row-major, leading dimension ld) -> LDS tile with pitch P.
template <class T, int P>
__device__ __forceinline__ void load_block(T *s, const T *g, long long ld) {
  if (sizeof(T) == 4) {
    for (int idx = threadIdx.x; idx < 64 * 16; idx += blockDim.x) {
      const int r = idx >> 4, c = (idx & 15) * 4;
      const float4 v = *reinterpret_cast<const float4 *>(g + r * ld + c);
      *reinterpret_cast<float4 *>(s + r * P + c) = v;
    }
  } else {
    for (int idx = threadIdx.x; idx < 64 * 32; idx += blockDim.x) {
      const int r = idx >> 5, c = (idx & 31) * 2;
      const double2 v = *reinterpret_cast<const double2 *>(g + r * ld + c);
      *reinterpret_cast<double2 *>(s + r * P + c) = v;
    }
  }
}
template <class T, int P>
__device__ __forceinline__ void store_block(T *g, const T *s, long long ld) {
  if (sizeof(T) == 4) {
    for (int idx = threadIdx.x; idx < 64 * 16; idx += blockDim.x) {
      const int r = idx >> 4, c = (idx & 15) * 4;
      *reinterpret_cast<float4 *>(g + r * ld + c) = *reinterpret_cast<const float4 *>(s + r * P + c);
    }
  } else {
    for (int idx = threadIdx.x; idx < 64 * 32; idx += blockDim.x) {
      const int r = idx >> 5, c = (idx & 31) * 2;
      *reinterpret_cast<double2 *>(g + r * ld + c) = *reinterpret_cast<const double2 *>(s + r * P + c);
    }
  }
}

// ---------------------------------------------------------------------------
// static pivoting: factor the diagonal block in LDS
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_lu_diag(const CellParams *cp, CellState *cs, const int *cells,
                                                 const int *prefix, int ncell, int k, T *K) {
  constexpr int P = Pitch<T>::v;
  __shared__ __attribute__((aligned(16))) T a[64 * P];
  const int ci = find_item_cell(prefix, ncell, blockIdx.x);
  const int c = cells[ci];
  const CellParams p = cp[c];
  T *blk = K + p.k_off + (long long)(k * NB) * p.npad + k * NB;
  load_block<T, P>(a, blk, p.npad);
  const T tau = (T)cs[c].tau;
  int nstatic = 0;
  __syncthreads();
  for (int j = 0; j < 64; ++j) {
    if (threadIdx.x == 0) {
      T piv = a[j * P + j];
      if (!(fabs((double)piv) >= (double)tau)) {
        a[j * P + j] = piv >= T(0) ? tau : -tau;
        ++nstatic;
      }
    }
    __syncthreads();
    const T piv = a[j * P + j];
    if (threadIdx.x > j && threadIdx.x < 64) a[threadIdx.x * P + j] /= piv;
    __syncthreads();
    const int m = 63 - j;
    for (int idx = threadIdx.x; idx < m * m; idx += 256) {
      const int r = j + 1 + idx / m, cc = j + 1 + idx % m;
      a[r * P + cc] -= a[r * P + j] * a[j * P + cc];
    }
    __syncthreads();
  }
  store_block<T, P>(blk, a, p.npad);
  if (threadIdx.x == 0 && nstatic) atomicAdd(&cs[c].n_static, nstatic);
}

template <class T>
void launch_lu_diag(hipStream_t s, const CellParams *cp, CellState *cs, const WorkList &wl, int k, T *K) {
  if (wl.total <= 0) return;
  hipLaunchKernelGGL(k_lu_diag<T>, dim3(wl.total), dim3(256), 0, s, cp, cs, wl.cells, wl.prefix, wl.ncell, k, K);
}

// ---------------------------------------------------------------------------
// partial pivoting: panel factorisation (one workgroup per cell)
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(512) void k_lu_panel_pp(const CellParams *cp, CellState *cs, const int *cells,
                                                     const int *prefix, int ncell, int k, T *K, int *ipiv) {
  __shared__ T prow[64];
  __shared__ double rv[16];
  __shared__ int ri[16];
  __shared__ int s_piv;
  const int ci = find_item_cell(prefix, ncell, blockIdx.x);
  const int c = cells[ci];
  const CellParams p = cp[c];
  const long long ld = p.npad;
  T *A = K + p.k_off;
  int *piv_out = ipiv + p.sys_off;
  const int k0 = k * NB;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int j = 0; j < 64; ++j) {
    const int col = k0 + j;
    double best = -1.0;
    int bi = 0x7fffffff;
    for (int r = col + threadIdx.x; r < p.npad; r += blockDim.x) {
      const double v = fabs((double)A[r * ld + col]);
      if (v > best) { best = v; bi = r; }
    }
    // wave argmax, ties to the lower row index
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_down(best, off, 64);
      const int oi = __shfl_down(bi, off, 64);
      if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
    }
    if (lane == 0) { rv[wave] = best; ri[wave] = bi; }
    __syncthreads();
    if (threadIdx.x == 0) {
      double b = rv[0];
      int i = ri[0];
      for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
        if (rv[w] > b || (rv[w] == b && ri[w] < i)) { b = rv[w]; i = ri[w]; }
      // a column of NaNs yields no candidate: keep the diagonal row
      if (i < col || i >= p.npad) i = col;
      s_piv = i;
      piv_out[col] = i;
      if (b == 0.0) cs[c].zero_pivot = 1;
    }
    __syncthreads();
    const int pr = s_piv;
    if (pr != col && threadIdx.x < 64) {
      const T t = A[col * ld + k0 + threadIdx.x];
      A[col * ld + k0 + threadIdx.x] = A[pr * ld + k0 + threadIdx.x];
      A[pr * ld + k0 + threadIdx.x] = t;
    }
    __syncthreads();
    if (threadIdx.x < 64) prow[threadIdx.x] = A[col * ld + k0 + threadIdx.x];
    __syncthreads();
    const T pv = prow[j];
    if (pv != T(0)) {
      for (int r = col + 1 + threadIdx.x; r < p.npad; r += blockDim.x) {
        T *row = A + r * ld + k0;
        const T l = row[j] / pv;
        row[j] = l;
        if (l != T(0))
          for (int cc = j + 1; cc < 64; ++cc) row[cc] -= l * prow[cc];
      }
    }
    __syncthreads();
  }
}
template <class T>
void launch_lu_panel_pp(hipStream_t s, const CellParams *cp, CellState *cs, const WorkList &wl, int k, T *K,
                        int *ipiv) {
  if (wl.total <= 0) return;
  hipLaunchKernelGGL(k_lu_panel_pp<T>, dim3(wl.total), dim3(512), 0, s, cp, cs, wl.cells, wl.prefix, wl.ncell, k, K,
                     ipiv);
}

// row interchanges of block k applied to the trailing columns [k0+64, npad)
template <class T>
__global__ __launch_bounds__(256) void k_lu_laswp(const CellParams *cp, const int *cells, const int *prefix, int ncell,
                                                  int k, T *K, const int *ipiv) {
  const int ci = find_item_cell(prefix, ncell, blockIdx.x);
  const int c = cells[ci];
  const CellParams p = cp[c];
  const int chunk = blockIdx.x - prefix[ci];
  const int k0 = k * NB;
  const int col = k0 + NB + chunk * 256 + threadIdx.x;
  if (col >= p.npad) return;
  T *A = K + p.k_off;
  const int *pv = ipiv + p.sys_off;
  const long long ld = p.npad;
  for (int j = 0; j < 64; ++j) {
    const int r = pv[k0 + j];
    if (r != k0 + j) {
      const T t = A[(k0 + j) * ld + col];
      A[(k0 + j) * ld + col] = A[r * ld + col];
      A[r * ld + col] = t;
    }
  }
}
template <class T>
void launch_lu_laswp(hipStream_t s, const CellParams *cp, const WorkList &wl, int k, T *K, const int *ipiv) {
  if (wl.total <= 0) return;
  hipLaunchKernelGGL(k_lu_laswp<T>, dim3(wl.total), dim3(256), 0, s, cp, wl.cells, wl.prefix, wl.ncell, k, K, ipiv);
}

// ---------------------------------------------------------------------------
// triangular block solves: U_kj = L_kk^-1 A_kj (items [0, m)) and, for static
// pivoting, L_ik = A_ik U_kk^-1 (items [m, 2m)), m = nblk - k - 1
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_lu_trsm(const CellParams *cp, const int *cells, const int *prefix, int ncell,
                                                 int k, T *K) {
  constexpr int P = Pitch<T>::v;
  __shared__ __attribute__((aligned(16))) T d[64 * P];
  __shared__ __attribute__((aligned(16))) T b[64 * P];
  const int ci = find_item_cell(prefix, ncell, blockIdx.x);
  const int c = cells[ci];
  const CellParams p = cp[c];
  const int t = blockIdx.x - prefix[ci];
  const int m = p.nblk - k - 1;
  const bool upper = t < m;
  const int other = k + 1 + (upper ? t : t - m);
  T *A = K + p.k_off;
  const long long ld = p.npad;
  T *diag = A + (long long)(k * NB) * ld + k * NB;
  T *blk = upper ? A + (long long)(k * NB) * ld + other * NB : A + (long long)(other * NB) * ld + k * NB;
  load_block<T, P>(d, diag, ld);
  load_block<T, P>(b, blk, ld);
  __syncthreads();
  if (upper) {
    // forward substitution with the unit lower triangle, all 64 columns at once
    for (int i = 0; i < 63; ++i) {
      const int nr = 63 - i;
      for (int idx = threadIdx.x; idx < nr * 64; idx += 256) {
        const int r = i + 1 + (idx >> 6), cc = idx & 63;
        b[r * P + cc] -= d[r * P + i] * b[i * P + cc];
      }
      __syncthreads();
    }
  } else {
    // X U = B, column by column
    for (int j = 0; j < 64; ++j) {
      if (threadIdx.x < 64) b[threadIdx.x * P + j] /= d[j * P + j];
      __syncthreads();
      const int nc = 63 - j;
      for (int idx = threadIdx.x; idx < nc * 64; idx += 256) {
        const int r = idx / nc, cc = j + 1 + idx % nc;
        b[r * P + cc] -= b[r * P + j] * d[j * P + cc];
      }
      __syncthreads();
    }
  }
  store_block<T, P>(blk, b, ld);
}
template <class T>
void launch_lu_trsm(hipStream_t s, const CellParams *cp, const WorkList &wl, int k, T *K, int) {
  if (wl.total <= 0) return;
  hipLaunchKernelGGL(k_lu_trsm<T>, dim3(wl.total), dim3(256), 0, s, cp, wl.cells, wl.prefix, wl.ncell, k, K);
}

// ---------------------------------------------------------------------------
// trailing update on MFMA: C(64x64) -= L(64x64) U(64x64)
// 4 waves, each a 32 x 32 quadrant = 2 x 2 MFMA 16x16 tiles.
// Operand layout of v_mfma_*_16x16x4: lane l supplies A[l%16][l/16] and
// B[l/16][l%16].  Accumulator register r of lane l holds
//   f32: D[4(l/16) + r][l%16]      f64: D[l/16 + 4r][l%16]
// (both measured on gfx950 with row/column-tagged operands).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int acc_row(float, int lk, int r) { return 4 * lk + r; }
__device__ __forceinline__ int acc_row(double, int lk, int r) { return lk + 4 * r; }
__device__ __forceinline__ floatx4 mfma16(float a, float b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ doublex4 mfma16(double a, double b, doublex4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
template <class T> struct Acc4;
template <> struct Acc4<float> { typedef floatx4 type; };
template <> struct Acc4<double> { typedef doublex4 type; };

template <class T>
__global__ __launch_bounds__(256) void k_lu_gemm(const CellParams *cp, const int *cells, const int *prefix, int ncell,
                                                 int k, T *K) {
  constexpr int P = Pitch<T>::v;
  typedef typename Acc4<T>::type acc_t;
  __shared__ __attribute__((aligned(16))) T ls[64 * P];
  __shared__ __attribute__((aligned(16))) T us[64 * P];
  const int ci = find_item_cell(prefix, ncell, blockIdx.x);
  const int c = cells[ci];
  const CellParams p = cp[c];
  const int t = blockIdx.x - prefix[ci];
  const int m = p.nblk - k - 1;
  const int bi = k + 1 + t / m, bj = k + 1 + t % m;
  T *A = K + p.k_off;
  const long long ld = p.npad;
  load_block<T, P>(ls, A + (long long)(bi * NB) * ld + k * NB, ld);
  load_block<T, P>(us, A + (long long)(k * NB) * ld + bj * NB, ld);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  T *C = A + (long long)(bi * NB) * ld + bj * NB;
  acc_t acc[2][2];
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[ti][tj][r] = C[(long long)(wr + ti * 16 + acc_row(T(0), lk, r)) * ld + wc + tj * 16 + lr];
  __syncthreads();
#pragma unroll 4
  for (int kk = 0; kk < 64; kk += 4) {
    T a[2], b[2];
#pragma unroll
    for (int ti = 0; ti < 2; ++ti) a[ti] = -ls[(wr + ti * 16 + lr) * P + kk + lk];
#pragma unroll
    for (int tj = 0; tj < 2; ++tj) b[tj] = us[(kk + lk) * P + wc + tj * 16 + lr];
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
      for (int tj = 0; tj < 2; ++tj) acc[ti][tj] = mfma16(a[ti], b[tj], acc[ti][tj]);
  }
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) C[(long long)(wr + ti * 16 + acc_row(T(0), lk, r)) * ld + wc + tj * 16 + lr] = acc[ti][tj][r];
}
template <class T>
void launch_lu_gemm(hipStream_t s, const CellParams *cp, const WorkList &wl, int k, T *K) {
  if (wl.total <= 0) return;
  hipLaunchKernelGGL(k_lu_gemm<T>, dim3(wl.total), dim3(256), 0, s, cp, wl.cells, wl.prefix, wl.ncell, k, K);
}

// ---------------------------------------------------------------------------
// triangular solves with the factors, one right-hand side per cell, the
// whole vector in LDS; one workgroup (4 waves) per cell
// ---------------------------------------------------------------------------
template <class T, bool PIVOT>
__global__ __launch_bounds__(256) void k_lu_solve(const CellParams *cp, const CellState *cs, const int *cells,
                                                  const T *K, const int *ipiv, const double *rhs, double *delta,
                                                  int accumulate) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  T *y = reinterpret_cast<T *>(smem);
  const int c = cells[blockIdx.x];
  if (cs[c].ir_done) return;
  const CellParams p = cp[c];
  const T *A = K + p.k_off;
  const long long ld = p.npad;
  const double *r = rhs + (long long)p.vec_off;
  for (int i = threadIdx.x; i < p.npad; i += 256) y[i] = i < p.ndof ? (T)r[i] : T(0);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const bool w0 = threadIdx.x < 64;
  // forward: P, unit lower L (LINPACK order: block interchanges, then block)
  for (int kb = 0; kb < p.nblk; ++kb) {
    const int k0 = kb * NB;
    if (PIVOT) {
      if (threadIdx.x == 0) {
        const int *pv = ipiv + p.sys_off;
        for (int j = 0; j < 64; ++j) {
          const int q = pv[k0 + j];
          if (q != k0 + j) {
            const T tmp = y[k0 + j];
            y[k0 + j] = y[q];
            y[q] = tmp;
          }
        }
      }
      __syncthreads();
    }
    if (w0) {
      T bv = y[k0 + lane];
      const T *row = A + (long long)(k0 + lane) * ld + k0;
#pragma unroll 8
      for (int cc = 0; cc < 63; ++cc) {
        const T yc = __shfl(bv, cc, 64);
        if (lane > cc) bv -= row[cc] * yc;
      }
      y[k0 + lane] = bv;
    }
    __syncthreads();
    for (int rr = k0 + NB + threadIdx.x; rr < p.npad; rr += 256) {
      const T *row = A + (long long)rr * ld + k0;
      T sacc = T(0);
      if (sizeof(T) == 4) {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          const float4 l4 = reinterpret_cast<const float4 *>(row)[q];
          sacc += (T)l4.x * y[k0 + 4 * q] + (T)l4.y * y[k0 + 4 * q + 1] + (T)l4.z * y[k0 + 4 * q + 2] +
                  (T)l4.w * y[k0 + 4 * q + 3];
        }
      } else {
#pragma unroll
        for (int q = 0; q < 32; ++q) {
          const double2 l2 = reinterpret_cast<const double2 *>(row)[q];
          sacc += (T)l2.x * y[k0 + 2 * q] + (T)l2.y * y[k0 + 2 * q + 1];
        }
      }
      y[rr] -= sacc;
    }
    __syncthreads();
  }
  // backward: U with its diagonal
  for (int kb = p.nblk - 1; kb >= 0; --kb) {
    const int k0 = kb * NB;
    if (w0) {
      T bv = y[k0 + lane];
      const T *row = A + (long long)(k0 + lane) * ld + k0;
      const T dg = row[lane];
#pragma unroll 8
      for (int cc = 63; cc >= 0; --cc) {
        if (lane == cc) bv = bv / dg;
        const T yc = __shfl(bv, cc, 64);
        if (lane < cc) bv -= row[cc] * yc;
      }
      y[k0 + lane] = bv;
    }
    __syncthreads();
    for (int rr = threadIdx.x; rr < k0; rr += 256) {
      const T *row = A + (long long)rr * ld + k0;
      T sacc = T(0);
      if (sizeof(T) == 4) {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          const float4 l4 = reinterpret_cast<const float4 *>(row)[q];
          sacc += (T)l4.x * y[k0 + 4 * q] + (T)l4.y * y[k0 + 4 * q + 1] + (T)l4.z * y[k0 + 4 * q + 2] +
                  (T)l4.w * y[k0 + 4 * q + 3];
        }
      } else {
#pragma unroll
        for (int q = 0; q < 32; ++q) {
          const double2 l2 = reinterpret_cast<const double2 *>(row)[q];
          sacc += (T)l2.x * y[k0 + 2 * q] + (T)l2.y * y[k0 + 2 * q + 1];
        }
      }
      y[rr] -= sacc;
    }
    __syncthreads();
  }
  double *dd = delta + (long long)p.vec_off;
  for (int i = threadIdx.x; i < p.ndof; i += 256) dd[i] = accumulate ? dd[i] + (double)y[i] : (double)y[i];
}

size_t lu_solve_lds_bytes(int npad, int elem_size) { return (size_t)npad * elem_size; }

void lu_prepare_attributes(size_t max_lds) {
  hipFuncSetAttribute(reinterpret_cast<const void *>(&k_lu_solve<float, true>),
                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)max_lds);
  hipFuncSetAttribute(reinterpret_cast<const void *>(&k_lu_solve<float, false>),
                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)max_lds);
  hipFuncSetAttribute(reinterpret_cast<const void *>(&k_lu_solve<double, true>),
                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)max_lds);
  hipFuncSetAttribute(reinterpret_cast<const void *>(&k_lu_solve<double, false>),
                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)max_lds);
}

template <class T>
void launch_lu_solve(hipStream_t s, const CellParams *cp, const CellState *cs, const int *cells, int ncell,
                     const T *K, const int *ipiv, int pivoting, const double *r, double *delta, int max_npad,
                     int accumulate) {
  if (ncell <= 0) return;
  const size_t lds = lu_solve_lds_bytes(max_npad, sizeof(T));
  if (pivoting)
    hipLaunchKernelGGL((k_lu_solve<T, true>), dim3(ncell), dim3(256), lds, s, cp, cs, cells, K, ipiv, r, delta,
                       accumulate);
  else
    hipLaunchKernelGGL((k_lu_solve<T, false>), dim3(ncell), dim3(256), lds, s, cp, cs, cells, K, ipiv, r, delta,
                       accumulate);
}

#define WCFE_INST(T)                                                                                              \
  template void launch_lu_diag<T>(hipStream_t, const CellParams *, CellState *, const WorkList &, int, T *);       \
  template void launch_lu_panel_pp<T>(hipStream_t, const CellParams *, CellState *, const WorkList &, int, T *,    \
                                      int *);                                                                      \
  template void launch_lu_laswp<T>(hipStream_t, const CellParams *, const WorkList &, int, T *, const int *);      \
  template void launch_lu_trsm<T>(hipStream_t, const CellParams *, const WorkList &, int, T *, int);               \
  template void launch_lu_gemm<T>(hipStream_t, const CellParams *, const WorkList &, int, T *);                    \
  template void launch_lu_solve<T>(hipStream_t, const CellParams *, const CellState *, const int *, int, const T *, \
                                   const int *, int, const double *, double *, int, int);
WCFE_INST(float)
WCFE_INST(double)

}  // namespace wcfe
