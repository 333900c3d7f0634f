// Batched dense blocked LU for gfx950: right-looking, 64 x 64 blocks,
// row-major matrices padded to a multiple of 64 (identity on the padding).
//
// Every diagonal block k is factored once (L_kk \ U_kk stored in place) and
// the explicit inverses L_kk^-1, U_kk^-1 are kept in a side buffer (Dinv).
// That turns every block triangular solve of the factorisation into a
// 64 x 64 x 64 product on MFMA, and the diagonal solves of the triangular
// sweeps into small matrix-vector products.
//
// Per block step k, over every cell of the work list at once:
//   static pivoting, two block steps k, k+1 at a time (delayed update):
//                     [k = 0: k_lu_diag_inv]
//                     k_lu_trsm_mm   U_kj = L_kk^-1 A_kj, L_ik = A_ik U_kk^-1
//                     k_lu_gemm      strip: block column and row k+1 updated
//                                    with panel k only
//                     k_lu_trsm_mm   panel k+1
//                     k_lu_gemm      A_ij -= L_ik U_kj + L_i,k+1 U_k+1,j on
//                                    i, j >= k+2 (each C block read and
//                                    written once per two steps)
//                     The workgroup that finishes a diagonal block goes on
//                     to factor it and invert its triangles.
//                     Pivots below tau (64 eps max|K_ii|) are replaced by
//                     +-tau (static pivoting, Li & Demmel); FP64 iterative
//                     refinement removes the perturbation.
//   partial pivoting: k_lu_panel_pp  one workgroup factors the 64-wide
//                                    column panel with row interchanges
//                                    (LAPACK getf2 order inside the panel)
//                                    and inverts its diagonal triangles
//                     k_lu_laswp     interchanges on the trailing columns
//                                    only (LINPACK order)
//                     k_lu_trsm_mm   U_kj = L_kk^-1 A_kj
//                     k_lu_gemm
// MFMA: v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64.
// k_lu_solve applies the factors to one right-hand side per cell with the
// whole vector resident in LDS.
#include <type_traits>
#include <utility>

#include <hip/hip_ext.h>

#include "wcfe_internal.hpp"

namespace wcfe {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

// Work items: map[item] = (cell << 20) | index of the item within the cell.
// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  Remap the workgroup id so that every XCD receives one contiguous
// range of work items: the tiles of one cell and block row then share an L2.
__device__ __forceinline__ int xcd_item(int b, int n) {
  const int q = n >> 3, r = n & 7, x = b & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
}

// Dinv layout: per cell, per diagonal block k: L_kk^-1 (64 x 64 row-major,
// unit diagonal included) then U_kk^-1.
__device__ __forceinline__ long long dinv_off(const CellParams &p, int k) {
  return (long long)p.sys_off * 128 + (long long)k * 8192;
}

template <class T> struct Pitch { static constexpr int v = 68; };
template <> struct Pitch<double> { static constexpr int v = 66; };

// 64 x 64 block (row-major, leading dimension ld) <-> LDS tile with pitch P.
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

// Inverses of the unit lower and the upper triangle of the factored block in
// `a` (LDS), written row-major to Li and Ui (global).  Wave 0: column c of
// L^-1 by forward substitution; wave 1: column c of U^-1 by back substitution.
template <class T, int P>
__device__ __forceinline__ void invert_triangles(const T *a, T *Li, T *Ui) {
  const int w = threadIdx.x >> 6, c = threadIdx.x & 63;
  T x[64];
  if (w == 0) {
#pragma unroll
    for (int i = 0; i < 64; ++i) x[i] = (i == c) ? T(1) : T(0);
#pragma unroll
    for (int i = 1; i < 64; ++i) {
      T s = x[i];
#pragma unroll
      for (int m = 0; m < i; ++m) s -= a[i * P + m] * x[m];
      x[i] = s;
    }
#pragma unroll
    for (int i = 0; i < 64; ++i) Li[i * 64 + c] = x[i];
  } else if (w == 1) {
#pragma unroll
    for (int i = 63; i >= 0; --i) {
      T s = (i == c) ? T(1) : T(0);
#pragma unroll
      for (int m = i + 1; m < 64; ++m) s -= a[i * P + m] * x[m];
      x[i] = s / a[i * P + i];
    }
#pragma unroll
    for (int i = 0; i < 64; ++i) Ui[i * 64 + c] = x[i];
  }
}

template <class T, int PA, bool STORE = true>
__device__ __forceinline__ int invert_diag(T *a, T *Li, T *Ui, T tau);

// k = 0 of static pivoting: invert block (0, 0).
template <class T>
__global__ __launch_bounds__(256) void k_lu_diag_inv(const CellParams *cp, CellState *cs, const int *map, int k, T *K, T *Dinv) {
  constexpr int P = Pitch<T>::v;
  __shared__ T a[64 * P];
  const int v = map[blockIdx.x];
  const int c = v >> 20;
  const CellParams p = cp[c];
  const long long ld = p.npad;
  T *blk = K + p.k_off + (long long)(k * NB) * ld + k * NB;
  for (int idx = threadIdx.x; idx < 4096; idx += blockDim.x)
    a[(idx >> 6) * P + (idx & 63)] = blk[(idx >> 6) * ld + (idx & 63)];
  __syncthreads();
  T *D = Dinv + dinv_off(p, k);
  const int fixed = invert_diag<T, P>(a, D, D + 4096, (T)cs[c].tau);
  if (fixed) atomicAdd(&cs[c].n_static, fixed);
}
template <class T>
void launch_lu_diag_inv(hipStream_t s, const CellParams *cp, CellState *cs, const WorkList &wl, int k, T *K,
                        T *Dinv) {
  if (wl.total <= 0) return;
  hipLaunchKernelGGL(k_lu_diag_inv<T>, dim3(wl.total), dim3(256), 0, s, cp, cs, wl.map, k, K,
                     Dinv);
}

// ---------------------------------------------------------------------------
// partial pivoting: panel factorisation (one workgroup per cell), then the
// inverses of the diagonal triangles
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(512) void k_lu_panel_pp(const CellParams *cp, CellState *cs, const int *map, int k, T *K, int *ipiv, T *Dinv) {
  constexpr int P = 65;
  __shared__ T prow[64];
  __shared__ double rv[16];
  __shared__ int ri[16];
  __shared__ int s_piv;
  __shared__ T a[64 * P];
  const int v = map[blockIdx.x];
  const int c = v >> 20;
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
  T *blk = A + (long long)k0 * ld + k0;
  for (int idx = threadIdx.x; idx < 4096; idx += blockDim.x)
    a[(idx >> 6) * P + (idx & 63)] = blk[(idx >> 6) * ld + (idx & 63)];
  __syncthreads();
  T *D = Dinv + dinv_off(p, k);
  invert_triangles<T, P>(a, D, D + 4096);
}
template <class T>
void launch_lu_panel_pp(hipStream_t s, const CellParams *cp, CellState *cs, const WorkList &wl, int k, T *K,
                        int *ipiv, T *Dinv) {
  if (wl.total <= 0) return;
  hipLaunchKernelGGL(k_lu_panel_pp<T>, dim3(wl.total), dim3(512), 0, s, cp, cs, wl.map, k, K,
                     ipiv, Dinv);
}

// row interchanges of block k applied to the trailing columns [k0+64, npad)
template <class T>
__global__ __launch_bounds__(256) void k_lu_laswp(const CellParams *cp, const int *map,
                                                  int k, T *K, const int *ipiv) {
  const int v = map[blockIdx.x];
  const int c = v >> 20;
  const CellParams p = cp[c];
  const int chunk = v & 0xFFFFF;
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
  hipLaunchKernelGGL(k_lu_laswp<T>, dim3(wl.total), dim3(256), 0, s, cp, wl.map, k, K, ipiv);
}

// ---------------------------------------------------------------------------
// 64 x 64 x 64 products on MFMA.  4 waves, each a 32 x 32 quadrant = 2 x 2
// MFMA 16x16 tiles.  Operand layout of v_mfma_*_16x16x4: lane l supplies
// A[l%16][l/16] and B[l/16][l%16].  Accumulator register r of lane l holds
//   f32: D[4(l/16) + r][l%16]      f64: D[l/16 + 4r][l%16]
// (both measured on gfx950 with row/column-tagged operands).
// ---------------------------------------------------------------------------
__device__ __forceinline__ floatx4 mfma16(float a, float b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ doublex4 mfma16(double a, double b, doublex4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ int acc_row(float, int lk, int r) { return 4 * lk + r; }
__device__ __forceinline__ int acc_row(double, int lk, int r) { return lk + 4 * r; }
template <class T> struct Acc4;
template <> struct Acc4<float> { typedef floatx4 type; };
template <> struct Acc4<double> { typedef doublex4 type; };

// Software pipeline of the LDS-operand MFMA loops: the operands of k step
// kk + 4 are read while the MFMAs of step kk run (two register stages; the
// scheduling barriers keep the compiler from sinking the reads to their
// first use, which costs one LDS latency per step).  The MFMA sequence of
// every accumulator is unchanged.
#define WCFE_SCHED_FENCE() __builtin_amdgcn_sched_barrier(0)

// acc += sign * X(64x64, LDS pitch PX) Y(64x64, LDS pitch PY) for this wave's quadrant
template <class T, int PX, int PY>
__device__ __forceinline__ void mma_64(const T *xs, const T *ys, typename Acc4<T>::type acc[2][2], T sign) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  T a[2][2], b[2][2];
  auto load = [&](int s, int kk) {
#pragma unroll
    for (int ti = 0; ti < 2; ++ti) a[s][ti] = xs[(wr + ti * 16 + lr) * PX + kk + lk];
#pragma unroll
    for (int tj = 0; tj < 2; ++tj) b[s][tj] = ys[(kk + lk) * PY + wc + tj * 16 + lr];
  };
  load(0, 0);
#pragma unroll
  for (int kk = 0; kk < 64; kk += 4) {
    const int s = (kk >> 2) & 1;
    if (kk + 4 < 64) load(s ^ 1, kk + 4);
    WCFE_SCHED_FENCE();
#pragma unroll
    for (int ti = 0; ti < 2; ++ti) {
      const T x = sign * a[s][ti];
#pragma unroll
      for (int tj = 0; tj < 2; ++tj) acc[ti][tj] = mfma16(x, b[s][tj], acc[ti][tj]);
    }
    WCFE_SCHED_FENCE();
  }
}

// ---------------------------------------------------------------------------
// Diagonal block: its inverse by Gauss-Jordan elimination without
// interchanges (pivots below tau replaced by +-tau), in 16 x 16 sub-blocks.
// Per sub-block step kb: wave 0 inverts the diagonal sub-block P in place
// (cross-lane reads), then on MFMA
//   row kb:       A_kb,j <- P^-1 A_kb,j                 (j != kb)
//   other rows i: A_i,j  <- A_i,j - A_i,kb A_kb,j  (j != kb),  A_i,kb <- -A_i,kb P^-1
// one wave per block row, so each row's old A_i,kb is read before it is
// replaced.  The block is taken as factored A_kk = I A_kk: the stored pair is
// Li = I and Ui = A_kk^-1, so the band LU's off-diagonal factors are
// L_ik = A_ik A_kk^-1 and U_kj = A_kj (block LDU form), and every solve that
// applies Li then Ui to a diagonal block applies A_kk^-1.
// a: the block in LDS (pitch PA), inverted in place; Li, Ui: 64 x 64
// row-major outputs (STORE).  All 256 threads call it (uniform barriers).
// Returns the pivots this thread replaced.
// ---------------------------------------------------------------------------
template <class T> __device__ __forceinline__ T rdlane(T v, int l);
template <> __device__ __forceinline__ float rdlane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
template <> __device__ __forceinline__ double rdlane(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __hiloint2double(hi, lo);
}

// reciprocal: the hardware's in FP32 (1 ulp, the factor is refined in FP64),
// IEEE division in FP64
__device__ __forceinline__ float recip(float v) { return __builtin_amdgcn_rcpf(v); }
__device__ __forceinline__ double recip(double v) { return 1.0 / v; }

// operand masks of the 16 x 16 products: as stored, unit lower, upper
enum : int { M_FULL = 0, M_UNIT_LOWER = 1, M_UPPER = 2 };
template <int M, class T> __device__ __forceinline__ T masked(T v, int r, int c) {
  if (M == M_UNIT_LOWER) return c < r ? v : (c == r ? T(1) : T(0));
  if (M == M_UPPER) return c >= r ? v : T(0);
  return v;
}

// acc += s X Y for 16 x 16 sub-blocks X, Y in LDS (pitch PA), one wave
template <class T, int PA, int MX, int MY>
__device__ __forceinline__ void mm16(const T *X, const T *Y, typename Acc4<T>::type &acc, T s) {
  const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
  T xv[4], yv[4];  // every operand read before the four chained MFMAs
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    xv[q] = X[lr * PA + 4 * q + lk];
    yv[q] = Y[(4 * q + lk) * PA + lr];
  }
  WCFE_SCHED_FENCE();
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const T x = masked<MX>(xv[q], lr, 4 * q + lk);
    const T y = masked<MY>(yv[q], 4 * q + lk, lr);
    acc = mfma16(s * x, y, acc);
  }
}
// acc[j] += s X Y_j for three 16 x 16 products with the same X (one wave):
// the three chains interleaved, every operand read first.  Each
// accumulator's MFMA sequence is mm16's.
template <class T, int PA>
__device__ __forceinline__ void mm16x3(const T *X, const T *Y0, const T *Y1, const T *Y2,
                                       typename Acc4<T>::type (&acc)[3], T s) {
  const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
  const T *Y[3] = {Y0, Y1, Y2};
  T xv[4], yv[3][4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    xv[q] = X[lr * PA + 4 * q + lk];
#pragma unroll
    for (int j = 0; j < 3; ++j) yv[j][q] = Y[j][(4 * q + lk) * PA + lr];
  }
  WCFE_SCHED_FENCE();
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const T x = s * xv[q];
#pragma unroll
    for (int j = 0; j < 3; ++j) acc[j] = mfma16(x, yv[j][q], acc[j]);
  }
}
template <class T, int PA>
__device__ __forceinline__ void acc16_load(typename Acc4<T>::type &acc, const T *Z) {
  const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int r = 0; r < 4; ++r) acc[r] = Z[acc_row(T(0), lk, r) * PA + lr];
}
template <class T, int PA>
__device__ __forceinline__ void acc16_store(const typename Acc4<T>::type &acc, T *Z) {
  const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int r = 0; r < 4; ++r) Z[acc_row(T(0), lk, r) * PA + lr] = acc[r];
}

// 16 x 16 sub-block at D (pitch PA) inverted in place, one wave: lane l
// holds row l & 15, columns 4 (l >> 4) .. +3; the pivot and the pivot row and
// column reach the other lanes by cross-lane reads.
// FP32: the pivot row by DPP row broadcast (row_newbcast: lane j of each
// 16-lane row) and the pivot column by permlane32 / permlane16 swaps (row gj
// of the wave copied to all four rows): no LDS round trip on the chain.
template <int J> __device__ __forceinline__ float row_bcast(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x150 + J, 0xF, 0xF, false));
}
template <int G> __device__ __forceinline__ float wave_row_bcast(float v) {
  const int x = __float_as_int(v);
  const auto h = __builtin_amdgcn_permlane32_swap(x, x, false, false);  // rows (0,1,0,1) | (2,3,2,3)
  const int y = G < 2 ? h[0] : h[1];
  const auto q = __builtin_amdgcn_permlane16_swap(y, y, false, false);  // even row | odd row, everywhere
  return __int_as_float(G % 2 == 0 ? q[0] : q[1]);
}
template <int J>
__device__ __forceinline__ void gj16_step(float (&v)[4], int lane, int r, int g, float tau, int &fixed) {
  constexpr int GJ = J >> 2, QJ = J & 3;
  const float p0 = rdlane(v[QJ], GJ * 16 + J);
  const bool small = !(fabsf(p0) >= tau);
  const float piv = small ? (p0 >= 0.0f ? tau : -tau) : p0;
  fixed += (small && lane == 0) ? 1 : 0;
  const float ip = recip(piv);
  const float f = wave_row_bcast<GJ>(v[QJ]);  // A[r][J]
  float rq[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) rq[q] = row_bcast<J>(v[q]) * ip;  // pivot row / pivot
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const bool cj = 4 * g + q == J;
    v[q] = (r == J) ? (cj ? ip : rq[q]) : (cj ? -f * ip : v[q] - f * rq[q]);
  }
}
template <int... J>
__device__ __forceinline__ void gj16_steps(float (&v)[4], int lane, int r, int g, float tau, int &fixed,
                                           std::integer_sequence<int, J...>) {
  (gj16_step<J>(v, lane, r, g, tau, fixed), ...);
}

template <class T, int PA> __device__ __forceinline__ int gj16(T *D, T tau) {
  const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
  if constexpr (sizeof(T) == 4) {
    float v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = D[r * PA + 4 * g + q];
    int fixed = 0;
    gj16_steps(v, lane, r, g, (float)tau, fixed, std::make_integer_sequence<int, 16>{});
#pragma unroll
    for (int q = 0; q < 4; ++q) D[r * PA + 4 * g + q] = v[q];
    return fixed;
  }
  T v[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) v[q] = D[r * PA + 4 * g + q];
  int fixed = 0;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int gj = j >> 2, qj = j & 3;
    const T p0 = rdlane(v[qj], gj * 16 + j);
    const bool small = !(fabs((double)p0) >= (double)tau);
    const T piv = small ? (p0 >= T(0) ? tau : -tau) : p0;
    fixed += (small && lane == 0) ? 1 : 0;
    const T ip = recip(piv);
    const T f = __shfl(v[qj], gj * 16 + r, 64);  // A[r][j]
    T rq[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) rq[q] = __shfl(v[q], g * 16 + j, 64) * ip;  // pivot row / pivot
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const bool cj = 4 * g + q == j;
      v[q] = (r == j) ? (cj ? ip : rq[q]) : (cj ? -f * ip : v[q] - f * rq[q]);
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) D[r * PA + 4 * g + q] = v[q];
  return fixed;
}

template <class T, int PA, bool STORE>
__device__ __forceinline__ int invert_diag(T *a, T *Li, T *Ui, T tau) {
  typedef typename Acc4<T>::type acc_t;
  const int wave = threadIdx.x >> 6;
  int fixed = 0;
  auto sub = [](T *base, int i, int j) { return base + 16 * i * PA + 16 * j; };
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) {
    if (wave == 0) fixed += gj16<T, PA>(sub(a, kb, kb), tau);
    __syncthreads();
    if (wave < 3) {  // row kb
      const int j = wave + (wave >= kb ? 1 : 0);
      acc_t acc;
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[r] = T(0);
      mm16<T, PA, M_FULL, M_FULL>(sub(a, kb, kb), sub(a, kb, j), acc, T(1));
      acc16_store<T, PA>(acc, sub(a, kb, j));
    }
    __syncthreads();
    if (wave < 3) {  // the other rows, one per wave
      const int i = wave + (wave >= kb ? 1 : 0);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (j == kb) continue;
        acc_t acc;
        acc16_load<T, PA>(acc, sub(a, i, j));
        mm16<T, PA, M_FULL, M_FULL>(sub(a, i, kb), sub(a, kb, j), acc, T(-1));
        acc16_store<T, PA>(acc, sub(a, i, j));
      }
      acc_t acc;
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[r] = T(0);
      mm16<T, PA, M_FULL, M_FULL>(sub(a, i, kb), sub(a, kb, kb), acc, T(-1));
      acc16_store<T, PA>(acc, sub(a, i, kb));
    }
    __syncthreads();
  }
  if (STORE) {
    for (int idx = threadIdx.x; idx < 4096; idx += blockDim.x) {
      const int r = idx >> 6, c = idx & 63;
      Li[idx] = r == c ? T(1) : T(0);
      Ui[idx] = a[r * PA + c];
    }
  }
  return fixed;
}

// invert_diag with work for the idle waves: wave 0 runs pre() and then the
// Gauss-Jordan inverse of sub-block kb while waves 1-3 run side(kb); every
// thread runs post(kb) right after that window's barrier.  The step's
// products take one phase (two barriers per sub-block step instead of
// three): with P the inverted pivot sub-block, wave 3 forms the new pivot
// row P A_kj in registers and stores it in the next window, while each other
// row's wave forms A'_ik = -A_ik P in place and then A_ij + A'_ik A_kj with
// the old pivot row -- the same elimination as invert_diag
// (A_ij - A_ik (P A_kj)), products associated the other way.
template <class T, int PA, bool STORE, class Pre, class Side, class Post>
__device__ __forceinline__ int invert_diag_ov(T *a, T *Li, T *Ui, T tau, Pre pre, Side side, Post post) {
  typedef typename Acc4<T>::type acc_t;
  const int wave = threadIdx.x >> 6;
  int fixed = 0;
  auto sub = [](T *base, int i, int j) { return base + 16 * i * PA + 16 * j; };
  acc_t rowk[3];  // wave 3: the new pivot row of the previous step
  auto store_row = [&](int kb) {
#pragma unroll
    for (int q = 0; q < 3; ++q) acc16_store<T, PA>(rowk[q], sub(a, kb, q + (q >= kb ? 1 : 0)));
  };
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) {
    if (wave == 0) {
      if (kb == 0) pre();
      fixed += gj16<T, PA>(sub(a, kb, kb), tau);
    } else {
      if (wave == 3 && kb > 0) store_row(kb - 1);
      side(kb);
    }
    __syncthreads();
    post(kb);
    const int j0 = kb == 0 ? 1 : 0, j1 = kb <= 1 ? 2 : 1, j2 = kb <= 2 ? 3 : 2;  // the columns j != kb
    if (wave < 3) {  // row i: A'_ik = -A_ik P in place, then A_ij += A'_ik A_kj (old row kb)
      const int i = wave + (wave >= kb ? 1 : 0);
      acc_t o[3];
      acc16_load<T, PA>(o[0], sub(a, i, j0));
      acc16_load<T, PA>(o[1], sub(a, i, j1));
      acc16_load<T, PA>(o[2], sub(a, i, j2));
      acc_t acc;
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[r] = T(0);
      mm16<T, PA, M_FULL, M_FULL>(sub(a, i, kb), sub(a, kb, kb), acc, T(-1));
      acc16_store<T, PA>(acc, sub(a, i, kb));
      // the wave reads back what its lanes wrote: LDS operations of one
      // wave are performed in order; keep the compiler from reordering them
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      mm16x3<T, PA>(sub(a, i, kb), sub(a, kb, j0), sub(a, kb, j1), sub(a, kb, j2), o, T(1));
      acc16_store<T, PA>(o[0], sub(a, i, j0));
      acc16_store<T, PA>(o[1], sub(a, i, j1));
      acc16_store<T, PA>(o[2], sub(a, i, j2));
    } else {  // wave 3: the new row kb, P A_kj, kept until the next window
#pragma unroll
      for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int r = 0; r < 4; ++r) rowk[q][r] = T(0);
      mm16x3<T, PA>(sub(a, kb, kb), sub(a, kb, j0), sub(a, kb, j1), sub(a, kb, j2), rowk, T(1));
    }
    __syncthreads();
  }
  if (wave == 3) store_row(3);
  __syncthreads();
  if (STORE) {
    for (int idx = threadIdx.x; idx < 4096; idx += blockDim.x) {
      const int r = idx >> 6, c = idx & 63;
      Li[idx] = r == c ? T(1) : T(0);
      Ui[idx] = a[r * PA + c];
    }
  }
  return fixed;
}

// acc += -X_I Y_J over k = 0..63 for the 16 x 16 output tile (I, J) of a
// 64 x 64 product, one wave; X, Y in LDS (pitch P).  The MFMA sequence per
// tile is mma_64's (same operands, same k order), so the tile comes out bit
// for bit as mma_64 computes it.
template <class T, int P>
__device__ __forceinline__ void mmneg_tile(const T *X, const T *Y, int I, int J, typename Acc4<T>::type &acc) {
  const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
#pragma unroll 4
  for (int kk = 0; kk < 64; kk += 4) {
    const T x = T(-1) * X[(I * 16 + lr) * P + kk + lk];
    const T y = Y[(kk + lk) * P + J * 16 + lr];
    acc = mfma16(x, y, acc);
  }
}

// block triangular solves as products with the stored pair (Li, Ui):
//   items [0, m):  U_kj = Li_k A_kj
//   items [m, 2m): L_ik = A_ik Ui_k   (static pivoting only)
// Under static pivoting the pair is (I, A_kk^-1) (block LDU form, see
// invert_diag), so U_kj = A_kj and L_ik = A_ik A_kk^-1; under partial
// pivoting (U items only) Li_k is the inverse of the unit lower factor.
// Symmetric cells (p.sym) have only the L items, [0, m): U_ki = A_ik^T is
// stored by transposition of the block before the product.
template <class T>
__global__ __launch_bounds__(256) void k_lu_trsm_mm(const CellParams *cp, const int *map, int k, T *K,
                                                    const T *Dinv) {
  constexpr int P = Pitch<T>::v;
  typedef typename Acc4<T>::type acc_t;
  __shared__ __attribute__((aligned(16))) T xs[64 * P];
  __shared__ __attribute__((aligned(16))) T ys[64 * P];
  const int item = xcd_item(blockIdx.x, gridDim.x);
  const int v = map[item];
  const int c = v >> 20;
  const CellParams p = cp[c];
  const int t = v & 0xFFFFF;
  const int m = min(p.nblk - k - 1, p.bw);
  const bool upper = !p.sym && t < m;
  const int other = k + 1 + ((upper || p.sym) ? t : t - m);
  const long long ld = p.npad;
  T *A = K + p.k_off;
  const T *D = Dinv + dinv_off(p, k);
  T *blk = upper ? A + (long long)(k * NB) * ld + other * NB : A + (long long)(other * NB) * ld + k * NB;
  if (upper) {
    load_block<T, P>(xs, D, 64);
    load_block<T, P>(ys, blk, ld);
  } else {
    load_block<T, P>(xs, blk, ld);
    load_block<T, P>(ys, D + 4096, 64);
  }
  __syncthreads();
  acc_t acc[2][2];
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[ti][tj][r] = T(0);
  mma_64<T, P, P>(xs, ys, acc, T(1));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wr = (w >> 1) * 32, wc = (w & 1) * 32, lr = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        blk[(long long)(wr + ti * 16 + acc_row(T(0), lk, r)) * ld + wc + tj * 16 + lr] = acc[ti][tj][r];
  if (!p.sym) return;
  // U_{k,other} = A_{other,k}^T, the block before the product (still in xs)
  T *U = A + (long long)(k * NB) * ld + other * NB;
  for (int idx = threadIdx.x; idx < 4096; idx += blockDim.x) {
    const int r = idx >> 6, cc = idx & 63;
    U[(long long)r * ld + cc] = xs[cc * P + r];
  }
}
template <class T>
void launch_lu_trsm(hipStream_t s, const CellParams *cp, CellState *, const WorkList &wl, int k, T *K,
                    const T *Dinv) {
  if (wl.total <= 0) return;
  hipLaunchKernelGGL(k_lu_trsm_mm<T>, dim3(wl.total), dim3(256), 0, s, cp, wl.map, k, K, Dinv);
}

// A 64 x 64 tile in registers: each thread holds TileShape::E elements,
// moved as 16-byte vectors.
template <class T> struct TileShape {
  static constexpr int VN = 16 / sizeof(T);     // elements per vector
  static constexpr int PER_ROW = 64 / VN;       // vectors per row
  static constexpr int N = 64 * PER_ROW / 256;  // vectors per thread
  static constexpr int E = N * VN;              // elements per thread
};
template <class T>
__device__ __forceinline__ void tile_load(T (&t)[TileShape<T>::E], const T *g, long long ld) {
  typedef TileShape<T> R;
#pragma unroll
  for (int j = 0; j < R::N; ++j) {
    const int idx = threadIdx.x + 256 * j, r = idx / R::PER_ROW, c = (idx % R::PER_ROW) * R::VN;
    if constexpr (R::VN == 4) {
      const float4 v = *reinterpret_cast<const float4 *>(g + r * ld + c);
      t[4 * j] = v.x; t[4 * j + 1] = v.y; t[4 * j + 2] = v.z; t[4 * j + 3] = v.w;
    } else {
      const double2 v = *reinterpret_cast<const double2 *>(g + r * ld + c);
      t[2 * j] = v.x; t[2 * j + 1] = v.y;
    }
  }
}
template <class T, int P>
__device__ __forceinline__ void tile_to_lds(const T (&t)[TileShape<T>::E], T *s) {
  typedef TileShape<T> R;
#pragma unroll
  for (int j = 0; j < R::N; ++j) {
    const int idx = threadIdx.x + 256 * j, r = idx / R::PER_ROW, c = (idx % R::PER_ROW) * R::VN;
    if constexpr (R::VN == 4) {
      *reinterpret_cast<float4 *>(s + r * P + c) = make_float4(t[4 * j], t[4 * j + 1], t[4 * j + 2], t[4 * j + 3]);
    } else {
      *reinterpret_cast<double2 *>(s + r * P + c) = make_double2(t[2 * j], t[2 * j + 1]);
    }
  }
}

// The fused block step's B operand (the U tile) is kept in LDS unpadded and
// swizzled: element (k, n) at 64 k + (n ^ 16 (k & 3)).  An MFMA B read takes
// rows kk .. kk+3 (lane group lk) and 16 consecutive columns (lr): the four
// row groups land on four different 16-bank groups, so the read is free of
// bank conflicts (a padded row pitch of 68 puts rows 4 words apart: up to
// four lanes on one bank).  The same values in the same MFMA order.
__device__ __forceinline__ int bswz(int k, int n) { return 64 * k + (n ^ ((k & 3) << 4)); }
template <class T>
__device__ __forceinline__ void tile_to_lds_bswz(const T (&t)[TileShape<T>::E], T *s) {
  typedef TileShape<T> R;
  static_assert(R::VN == 4, "FP32 tiles");
#pragma unroll
  for (int j = 0; j < R::N; ++j) {
    const int idx = threadIdx.x + 256 * j, r = idx / R::PER_ROW, c = (idx % R::PER_ROW) * R::VN;
    *reinterpret_cast<float4 *>(s + bswz(r, c)) = make_float4(t[4 * j], t[4 * j + 1], t[4 * j + 2], t[4 * j + 3]);
  }
}
template <class T, int P>
__device__ __forceinline__ void mmneg_tile_bswz(const T *X, const T *Y, int I, int J, typename Acc4<T>::type &acc) {
  const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int kk = 0; kk < 64; kk += 4) {
    const T x = T(-1) * X[(I * 16 + lr) * P + kk + lk];
    const T y = Y[bswz(kk + lk, J * 16 + lr)];
    acc = mfma16(x, y, acc);
  }
}
// The four tiles (I, 0..3) and the tile (0, J0) of the same product, one
// wave: each k step reads two X rows and four Y columns for five MFMAs on
// five independent accumulators (one tile at a time waits on its LDS reads
// every four MFMAs).  Each tile's MFMA sequence is mmneg_tile_bswz's.
template <class T, int P>
__device__ __forceinline__ void mmneg_row_bswz(const T *X, const T *Y, int I, int J0, typename Acc4<T>::type (&acc)[4],
                                               typename Acc4<T>::type &acc0) {
  const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
#pragma unroll 4
  for (int kk = 0; kk < 64; kk += 4) {
    const T x = T(-1) * X[(I * 16 + lr) * P + kk + lk];
    const T x0 = T(-1) * X[lr * P + kk + lk];
    T y[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) y[j] = Y[bswz(kk + lk, j * 16 + lr)];
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[j] = mfma16(x, y[j], acc[j]);
    acc0 = mfma16(x0, J0 == 1 ? y[1] : J0 == 2 ? y[2] : y[3], acc0);
  }
}
template <class T, int PX>
__device__ __forceinline__ void mma_64_bswz(const T *xs, const T *ys, typename Acc4<T>::type acc[2][2], T sign) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  T a[2][2], b[2][2];
  auto load = [&](int s, int kk) {
#pragma unroll
    for (int ti = 0; ti < 2; ++ti) a[s][ti] = xs[(wr + ti * 16 + lr) * PX + kk + lk];
#pragma unroll
    for (int tj = 0; tj < 2; ++tj) b[s][tj] = ys[bswz(kk + lk, wc + tj * 16 + lr)];
  };
  load(0, 0);
#pragma unroll
  for (int kk = 0; kk < 64; kk += 4) {
    const int s = (kk >> 2) & 1;
    if (kk + 4 < 64) load(s ^ 1, kk + 4);
    WCFE_SCHED_FENCE();
#pragma unroll
    for (int ti = 0; ti < 2; ++ti) {
      const T x = sign * a[s][ti];
#pragma unroll
      for (int tj = 0; tj < 2; ++tj) acc[ti][tj] = mfma16(x, b[s][tj], acc[ti][tj]);
    }
    WCFE_SCHED_FENCE();
  }
}

// Trailing update C = A_ij - sum_p L_{i,k+p} U_{k+p,j} over np consecutive
// panels k .. k+np-1 (C is read and written once for np panels); the next
// panel's tiles are loaded into registers while the current one is on MFMA.
//   square mode: (i, j) over [s0, nblk)^2 with s0 = k + np (lower triangle
//                i >= j for symmetric cells)
//   strip mode:  target s0 = k + np: block column s0 from row s0 down, then
//                (non-symmetric cells) block row s0 from column s0 + 1 on --
//                the blocks panel s0 needs
// With `next_diag` (static pivoting) the workgroup of the diagonal block
// (s0, s0) then factors that block and stores the inverses of its triangles.
template <class T, bool STRIP>
__global__ __launch_bounds__(256) void k_lu_gemm(const CellParams *cp, CellState *cs, const int *map, int k, T *K,
                                                 T *Dinv, int next_diag, int np) {
  constexpr int P = Pitch<T>::v;
  typedef typename Acc4<T>::type acc_t;
  __shared__ __attribute__((aligned(16))) T ls[64 * P];
  __shared__ __attribute__((aligned(16))) T us[64 * P];
  const int item = xcd_item(blockIdx.x, gridDim.x);
  const int v = map[item];
  const int c = v >> 20;
  const CellParams p = cp[c];
  const int t = v & 0xFFFFF;
  const int s0 = k + np;
  int bi, bj;
  if (STRIP) {
    const int m = min(p.nblk - s0, p.bw);
    if (t < m) { bi = s0 + t; bj = s0; }
    else { bi = s0; bj = s0 + 1 + (t - m); }
  } else if (p.sym) {
    // lower triangle (i >= j) of the trailing matrix, row by row
    int i = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
    while (i * (i + 1) / 2 > t) --i;
    while ((i + 1) * (i + 2) / 2 <= t) ++i;
    bi = s0 + i;
    bj = s0 + (t - i * (i + 1) / 2);
  } else {
    const int mt = min(p.nblk - s0, p.bw);
    bi = s0 + t / mt;
    bj = s0 + t % mt;
  }
  T *A = K + p.k_off;
  const long long ld = p.npad;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
  const int lk = lane >> 4, lr = lane & 15;
  T *C = A + (long long)(bi * NB) * ld + bj * NB;
  T lt[TileShape<T>::E], ut[TileShape<T>::E];
  tile_load<T>(lt, A + (long long)(bi * NB) * ld + k * NB, ld);
  tile_load<T>(ut, A + (long long)(k * NB) * ld + bj * NB, ld);
  acc_t acc[2][2];
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        acc[ti][tj][r] = C[(long long)(wr + ti * 16 + acc_row(T(0), lk, r)) * ld + wc + tj * 16 + lr];
  for (int pn = 0; pn < np; ++pn) {
    if (pn) __syncthreads();  // the previous panel's products are done with the LDS tiles
    tile_to_lds<T, P>(lt, ls);
    tile_to_lds<T, P>(ut, us);
    __syncthreads();
    if (pn + 1 < np) {
      tile_load<T>(lt, A + (long long)(bi * NB) * ld + (k + pn + 1) * NB, ld);
      tile_load<T>(ut, A + (long long)((k + pn + 1) * NB) * ld + bj * NB, ld);
    }
    mma_64<T, P, P>(ls, us, acc, T(-1));
  }
  const bool diag_next = next_diag && bi == s0 && bj == s0;
  if (!diag_next) {
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
      for (int tj = 0; tj < 2; ++tj)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          C[(long long)(wr + ti * 16 + acc_row(T(0), lk, r)) * ld + wc + tj * 16 + lr] = acc[ti][tj][r];
    return;
  }
  // diagonal block s0 is final: factor it here
  __syncthreads();
  constexpr int PA = P;
  T *a = ls;
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) a[(wr + ti * 16 + acc_row(T(0), lk, r)) * PA + wc + tj * 16 + lr] = acc[ti][tj][r];
  __syncthreads();
  T *D = Dinv + dinv_off(p, s0);
  const int fixed = invert_diag<T, PA>(a, D, D + 4096, (T)cs[c].tau);
  if (fixed) atomicAdd(&cs[c].n_static, fixed);
}
template <class T>
void launch_lu_gemm(hipStream_t s, const CellParams *cp, CellState *cs, const WorkList &wl, int k, T *K, T *Dinv,
                    int next_diag, int npanel, int strip, hipEvent_t ev_start, hipEvent_t ev_stop) {
  if (wl.total <= 0) return;
  if (ev_start) {  // profiling (never under graph capture): events at the kernel's own start and end
    if (strip)
      hipExtLaunchKernelGGL((k_lu_gemm<T, true>), dim3(wl.total), dim3(256), 0, s, ev_start, ev_stop, 0, cp, cs,
                            wl.map, k, K, Dinv, next_diag, npanel);
    else
      hipExtLaunchKernelGGL((k_lu_gemm<T, false>), dim3(wl.total), dim3(256), 0, s, ev_start, ev_stop, 0, cp, cs,
                            wl.map, k, K, Dinv, next_diag, npanel);
  } else if (strip) {
    hipLaunchKernelGGL((k_lu_gemm<T, true>), dim3(wl.total), dim3(256), 0, s, cp, cs, wl.map, k, K, Dinv, next_diag,
                       npanel);
  } else {
    hipLaunchKernelGGL((k_lu_gemm<T, false>), dim3(wl.total), dim3(256), 0, s, cp, cs, wl.map, k, K, Dinv, next_diag,
                       npanel);
  }
}

// ---------------------------------------------------------------------------
// Fused block step for narrow bands under static pivoting, one launch per
// block step k = -1 .. nblk-2 (s0 = k + 1): the trailing update of panel k,
// the inverse of diagonal block s0, and the block solves of panel s0.  Items
// of a cell (lu_step_items):
//   t = 0          diagonal block s0: invert, store the inverse
//   t = 1 .. w     L_{s0+t, s0} = A_{s0+t, s0} A_{s0,s0}^-1
//   t = w+1 .. 2w  U_{s0, s0+t-w} = A_{s0, s0+t-w} (non-symmetric; the block
//                  LDU form leaves the row as it is)
//   then           the other blocks of panel k's window
// with w = min(nblk - s0 - 1, bw).  Blocks inside panel k's reach (row and
// column <= k + bw) take the update A -= L_{i,k} U_{k,j} first.  Every L item
// computes the inverse of block s0 itself from the same inputs (so bit for
// bit the diagonal item's), so no workgroup waits for another; only the
// diagonal item stores it.  Diagonal blocks of K keep their unfactored
// values: nothing reads them afterwards.  Symmetric cells: items are the
// lower blocks; an L item also stores U_{s0,i} = A_{i,s0}^T.
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ void step_tile(typename Acc4<T>::type (&acc)[2][2], const Chain &v, const T *A,
                                          long long ld, int bi, int bj, int k, bool upd, T *ls, T *us) {
  constexpr int P = Pitch<T>::v;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
  const int lk = lane >> 4, lr = lane & 15;
  const T *C = A + v.off(bi, bj, ld);
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        acc[ti][tj][r] = C[(long long)(wr + ti * 16 + acc_row(T(0), lk, r)) * ld + wc + tj * 16 + lr];
  if (!upd) return;
  T lt[TileShape<T>::E], ut[TileShape<T>::E];
  tile_load<T>(lt, A + v.off(bi, k, ld), ld);
  tile_load<T>(ut, A + v.off(k, bj, ld), ld);
  __syncthreads();  // earlier users of ls / us are done
  tile_to_lds<T, P>(lt, ls);
  tile_to_lds_bswz<T>(ut, us);
  __syncthreads();
  mma_64_bswz<T, P>(ls, us, acc, T(-1));
}

__device__ __forceinline__ int lu_step_w(const Chain &v, int k) { return min(v.nblk - k - 2, v.bw); }
__device__ __forceinline__ int lu_step_np(const Chain &v, int k) {
  return k >= 0 ? max(0, min(v.nblk - 1, k + v.bw) - (k + 1)) : 0;
}

// acc <- the 64 x 64 block at C (leading dimension ld), MFMA accumulator layout
template <class T>
__device__ __forceinline__ void acc_load64(typename Acc4<T>::type (&acc)[2][2], const T *C, long long ld) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wr = (w >> 1) * 32, wc = (w & 1) * 32, lk = lane >> 4, lr = lane & 15;
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        acc[ti][tj][r] = C[(long long)(wr + ti * 16 + acc_row(T(0), lk, r)) * ld + wc + tj * 16 + lr];
}
template <class T, int PS>
__device__ __forceinline__ void acc_store64(const typename Acc4<T>::type (&acc)[2][2], T *C, long long ld) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wr = (w >> 1) * 32, wc = (w & 1) * 32, lk = lane >> 4, lr = lane & 15;
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) C[(long long)(wr + ti * 16 + acc_row(T(0), lk, r)) * ld + wc + tj * 16 + lr] = acc[ti][tj][r];
}

// Chain B of a symmetric two-chain cell updates the middle window's blocks
// below the diagonal in transposed position: its virtual lower block (i, j),
// i > j, is real block (rmax - i, rmax - j) above the diagonal, and
// L_jk U_ki = (L_ik U_kj)^T for a symmetric matrix, so B applies the update
// of virtual (j, i) = real lower block instead -- the blocks that chain A
// updates and factors.
__device__ __forceinline__ void mid_swap(const Chain &v, int &ti, int &tj) {
  if (v.flip && v.sym && tj >= v.elim && ti > tj) {
    const int x = ti;
    ti = tj;
    tj = x;
  }
}

// One item t of cell c in block step k (see the comment above step_tile).
// All threads of the workgroup call it with the same arguments (uniform
// barriers); ls, us, xs are the workgroup's three LDS tiles.
template <class T>
__device__ __forceinline__ void lu_step_item(const CellParams *cp, CellState *cs, int c, int t, int k, T *K, T *Dinv,
                                             int panel, StepFwd fwd, T *ls, T *us, T *xs) {
  constexpr int P = Pitch<T>::v, PA = P;
  typedef typename Acc4<T>::type acc_t;
  const CellParams p = cp[c];
  // the item's chain: A (or the cell's only chain), then B's items
  Chain v = chain_of(p, 0);
  if (p.tw_m > 0) {
    const int na = chain_items(v, k, panel);
    if (t >= na) {
      v = chain_of(p, 1);
      t -= na;
    }
  }
  const int s0 = k + 1, wl = lu_step_w(v, k), npl = lu_step_np(v, k);
  T *A = K + p.k_off;
  const long long ld = p.npad;
  // forward substitution of the first solve (right-hand side -G in the band
  // numbering) for the chain's block row s: y_s = base - sum_{j in [jlo,
  // jhi)} L_s,j y_j with L_s,j and y_j final since earlier steps; base = -G,
  // or (chain A's rows of a two-chain cell's window) y_s as chain B left it;
  // 4 lanes per row
  auto fwd_row = [&](int s, int jlo, int jhi, bool from_y) {
    float *y = fwd.y + p.sys_off;
    const int r = threadIdx.x >> 2, q = threadIdx.x & 3;
    float part = 0.0f;
    for (int j = jlo; j < jhi; ++j) {
      const T *Lr = A + v.off(s, j, ld) + (long long)r * ld + q * 16;
      const float *yj = y + v.rb(j) * NB + q * 16;
#pragma unroll
      for (int e = 0; e < 16; ++e) part += (float)Lr[e] * yj[e];
    }
    part += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(part), 0xB1, 0xF, 0xF, false));
    part += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(part), 0x4E, 0xF, 0xF, false));
    if (q == 0) {
      const int i = v.rb(s) * NB + r, dof = fwd.kinv[p.sys_off + i];
      const float base = from_y ? y[i] : (dof >= 0 ? (float)(-fwd.G[dof]) : 0.0f);
      y[i] = base - part;
    }
  };
  if (s0 == v.elim) {
    // update-only step of a chain that stops before the middle window:
    // block t of the window [s0, k + bw] takes panel k's update
    const int nu = min(v.nblk - 1, k + v.bw) - s0 + 1;
    if (t == (v.sym ? nu * (nu + 1) / 2 : nu * nu)) {
      // the window rows' forward with this chain's L blocks (columns before
      // the window); the other chain adds its own and finishes them
      if (fwd.y)
        for (int s = s0; s < v.nblk; ++s) fwd_row(s, max(0, s - v.bw), s0, false);
      return;
    }
    int ti, tj;
    if (v.sym) {
      int i = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
      while (i * (i + 1) / 2 > t) --i;
      while ((i + 1) * (i + 2) / 2 <= t) ++i;
      ti = s0 + i;
      tj = s0 + (t - i * (i + 1) / 2);
    } else {
      ti = s0 + t / nu;
      tj = s0 + t % nu;
    }
    mid_swap(v, ti, tj);
    acc_t acc[2][2];
    step_tile<T>(acc, v, A, ld, ti, tj, k, true, ls, us);
    acc_store64<T, 0>(acc, A + v.off(ti, tj, ld), ld);
    return;
  }
  if (k < 0) panel = 0;  // the first step always runs L items
  // panel mode (panel 1): item 0 inverts block s0 once and solves every L
  // block of the panel (no L items); one item takes all trailing blocks of
  // the window.  Split panel mode (panel 2): item 0 inverts block s0 and
  // solves the first L block, the others are L items (each inverts block s0
  // itself), one item takes the trailing blocks.
  const int nl = panel == 1 ? 0 : panel == 2 ? (wl > 1 ? wl - 1 : 0) : wl;
  const int nsolve = nl + (v.sym ? 0 : wl);
  // trailing tile q of the window (lower triangle for symmetric cells)
  auto plain_tile = [&](int q, int &ti, int &tj) {
    if (v.sym) {
      int i = (int)((sqrtf(8.0f * (float)q + 1.0f) - 1.0f) * 0.5f);
      while (i * (i + 1) / 2 > q) --i;
      while ((i + 1) * (i + 2) / 2 <= q) ++i;
      ti = s0 + 1 + i;
      tj = s0 + 1 + (q - i * (i + 1) / 2);
    } else {
      ti = s0 + 1 + q / npl;
      tj = s0 + 1 + q % npl;
    }
  };
  // 0 diagonal, 1 L block, 2 U block (stored as updated), 3 other, 4 every
  // other block of the window (panel mode: one item per cell)
  const int ntile = v.sym ? npl * (npl + 1) / 2 : npl * npl;
  // forward substitution of block s0 for the first solve (right-hand side
  // -G in the band numbering): y_s0 = (-G)_s0 - sum_j L_s0,j y_j over the
  // band, with L_s0,j and y_j final since earlier steps; 4 lanes per row.
  // The last item of the cell; in panel mode the trailing-block item runs it
  // after its blocks (no workgroup of its own in the many-cell launches).
  auto forward_block = [&]() {
    if (!fwd.y) return;
    fwd_row(s0, max(0, s0 - v.bw), s0, p.tw_m > 0 && !v.flip && s0 >= p.tw_m);
  };
  const int t_last = 1 + nsolve + (panel ? 0 : ntile);
  if (t == t_last && !(panel && ntile > 0)) {
    forward_block();
    return;
  }
  int role, bi, bj;
  if (t == 0) { role = 0; bi = panel && wl > 0 ? s0 + 1 : s0; bj = s0; }
  else if (t <= nl) { role = 1; bi = s0 + t + (panel == 2 ? 1 : 0); bj = s0; }
  else if (t <= nsolve) { role = 2; bi = s0; bj = s0 + (t - nl); }
  else if (panel) { role = 4; bi = s0; bj = s0; }
  else {
    role = 3;
    plain_tile(t - 1 - nsolve, bi, bj);
    mid_swap(v, bi, bj);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
  const int lk = lane >> 4, lr = lane & 15;
  const int reach = k + v.bw;  // panel k's last block row / column
  T *C = A + v.off(bi, bj, ld);
  acc_t acc[2][2];
  if (role == 4) {  // the window's tiles one after another, as items of role 3 take them
    for (int q = 0; q < ntile; ++q) {
      int ti, tj;
      plain_tile(q, ti, tj);
      mid_swap(v, ti, tj);
      step_tile<T>(acc, v, A, ld, ti, tj, k, k >= 0 && ti <= reach && tj <= reach, ls, us);
      acc_store64<T, 0>(acc, A + v.off(ti, tj, ld), ld);
    }
    forward_block();
    return;
  }
  if (role >= 2) {
    step_tile<T>(acc, v, A, ld, bi, bj, k, k >= 0 && bi <= reach && bj <= reach, ls, us);
    acc_store64<T, 0>(acc, C, ld);
    return;
  }
  const bool store = role == 0, own = role == 1 || (store && panel && wl > 0), upd = k >= 0,
             upd_own = own && upd && bi <= reach;
  if (upd) {
    // Panel k updates block s0 (and the own block): the updates run inside
    // the inverse.  Block s0 -= L_{s0,k} U_{k,s0}: wave 0 takes sub-block
    // (0,0) just before its Gauss-Jordan step, waves 1-3 the other 15
    // meanwhile.  The own block's 16 sub-blocks live in waves 1-3's
    // registers; they take the update A -= L_{bi,k} U_{k,s0} during steps
    // 1-3 and are staged in ls for the final product.  LDS: xs block s0
    // (then its inverse), ls L_{s0,k} (then L_{bi,k}, then the own block),
    // us U_{k,s0}.  Every tile's MFMA sequence is mma_64's.
    T lt[TileShape<T>::E], ld2[TileShape<T>::E], ut[TileShape<T>::E];
    tile_load<T>(lt, A + v.off(s0, s0, ld), ld);
    tile_load<T>(ld2, A + v.off(s0, k, ld), ld);
    tile_load<T>(ut, A + v.off(k, s0, ld), ld);
    acc_t ow[6];
    if (own && w > 0) {
#pragma unroll
      for (int sl = 0; sl < 6; ++sl) {
        const int tt = (w - 1) + 3 * sl;
        if (tt < 16) {
          const T *Ct = C + (long long)((tt >> 2) * 16) * ld + (tt & 3) * 16;
#pragma unroll
          for (int r = 0; r < 4; ++r) ow[sl][r] = Ct[(long long)acc_row(T(0), lk, r) * ld + lr];
        }
      }
    }
    tile_to_lds<T, P>(lt, xs);
    tile_to_lds<T, P>(ld2, ls);
    tile_to_lds_bswz<T>(ut, us);
    if (upd_own) tile_load<T>(lt, A + v.off(bi, k, ld), ld);
    // panel item: the second L block (and, inside panel k's reach, its L
    // block of column k, below in post) are loaded while the inverse runs
    const bool pf2 = store && panel == 1 && wl >= 2, pf2_upd = pf2 && s0 + 2 <= reach;
    acc_t pf[2][2];
    if (pf2) acc_load64<T>(pf, A + v.off(s0 + 2, s0, ld), ld);
    __syncthreads();
    auto pre = [&]() {
      acc_t d;
      acc16_load<T, PA>(d, xs);
      mmneg_tile_bswz<T, P>(ls, us, 0, 0, d);
      acc16_store<T, PA>(d, xs);
    };
    T *Ut = A + v.off(s0, bi, ld);  // U_{s0,bi} (symmetric cells)
    auto side = [&](int kb) {
      if (kb == 0) {  // wave w: row w of the sub-blocks and sub-block (0, w)
        acc_t d[4], d0;
#pragma unroll
        for (int J = 0; J < 4; ++J) acc16_load<T, PA>(d[J], xs + w * 16 * PA + J * 16);
        acc16_load<T, PA>(d0, xs + w * 16);
        mmneg_row_bswz<T, P>(ls, us, w, w, d, d0);
#pragma unroll
        for (int J = 0; J < 4; ++J) acc16_store<T, PA>(d[J], xs + w * 16 * PA + J * 16);
        acc16_store<T, PA>(d0, xs + w * 16);
        return;
      }
      if (!own) return;
#pragma unroll
      for (int sl = 2 * (kb - 1); sl < 2 * kb; ++sl) {
        const int tt = (w - 1) + 3 * sl;
        if (tt >= 16) continue;
        const int I = tt >> 2, J = tt & 3;
        if (upd_own) mmneg_tile_bswz<T, P>(ls, us, I, J, ow[sl]);
        if (v.sym) {  // U_{s0,bi} = A_{bi,s0}^T: four consecutive entries of one row per lane
          const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
          *reinterpret_cast<float4 *>(Ut + (long long)(J * 16 + lr) * ld + I * 16 + 4 * lk) =
              make_float4(ow[sl][0], ow[sl][1], ow[sl][2], ow[sl][3]);
        }
      }
    };
    auto post = [&](int kb) {
      if (kb == 0 && upd_own) tile_to_lds<T, P>(lt, ls);  // L_{s0,k} is done with
      if (kb == 0 && pf2_upd) tile_load<T>(lt, A + v.off(s0 + 2, k, ld), ld);
      if (kb == 3 && own && w > 0) {
#pragma unroll
        for (int sl = 0; sl < 6; ++sl) {
          const int tt = (w - 1) + 3 * sl;
          if (tt < 16) acc16_store<T, P>(ow[sl], ls + (tt >> 2) * 16 * P + (tt & 3) * 16);
        }
      }
    };
    T *D = Dinv + dinv_off(p, v.rb(s0));
    if (store) {
      const int fixed = invert_diag_ov<T, PA, true>(xs, D, D + 4096, (T)cs[c].tau, pre, side, post);
      if (fixed) atomicAdd(&cs[c].n_static, fixed);
    } else {
      (void)invert_diag_ov<T, PA, false>(xs, D, D + 4096, (T)cs[c].tau, pre, side, post);
    }
    if (!own) return;
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
      for (int tj = 0; tj < 2; ++tj)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[ti][tj][r] = T(0);
    mma_64<T, P, PA>(ls, xs, acc, T(1));
    acc_store64<T, 0>(acc, C, ld);
    if (!(store && panel == 1)) return;
    // panel item: L blocks s0+2 .. s0+w after the inverse, each as an L item
    // computes it (same products in the same order)
    for (int t2 = 2; t2 <= wl; ++t2) {
      const int b2 = s0 + t2;
      T *C2 = A + v.off(b2, s0, ld);
      if (t2 == 2) {
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
#pragma unroll
          for (int tj = 0; tj < 2; ++tj) acc[ti][tj] = pf[ti][tj];
      } else {
        acc_load64<T>(acc, C2, ld);
      }
      __syncthreads();  // the previous block's product is done with ls
      if (b2 <= reach) {
        if (t2 > 2) tile_load<T>(lt, A + v.off(b2, k, ld), ld);
        tile_to_lds<T, P>(lt, ls);
        __syncthreads();
        mma_64_bswz<T, P>(ls, us, acc, T(-1));  // us: U_{k,s0}
        __syncthreads();
      }
      if (v.sym) {  // U_{s0,b2} = A_{b2,s0}^T from the registers: four consecutive entries of one row per lane
        T *U2 = A + v.off(s0, b2, ld);
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
#pragma unroll
          for (int tj = 0; tj < 2; ++tj) {
            T *u2 = U2 + (long long)(wc + tj * 16 + lr) * ld + wr + ti * 16;
            if constexpr (sizeof(T) == 4) {
              *reinterpret_cast<float4 *>(u2 + 4 * lk) =
                  make_float4(acc[ti][tj][0], acc[ti][tj][1], acc[ti][tj][2], acc[ti][tj][3]);
            } else {
#pragma unroll
              for (int r = 0; r < 4; ++r) u2[acc_row(T(0), lk, r)] = acc[ti][tj][r];
            }
          }
      }
#pragma unroll
      for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj)
#pragma unroll
          for (int r = 0; r < 4; ++r) ls[(wr + ti * 16 + acc_row(T(0), lk, r)) * P + wc + tj * 16 + lr] = acc[ti][tj][r];
      __syncthreads();
#pragma unroll
      for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj)
#pragma unroll
          for (int r = 0; r < 4; ++r) acc[ti][tj][r] = T(0);
      mma_64<T, P, PA>(ls, xs, acc, T(1));
      acc_store64<T, 0>(acc, C2, ld);
    }
    return;
  }
  // first step (no panel before it): block s0 and the own block as they are
  acc_t dacc[2][2];
  acc_load64<T>(dacc, A + v.off(s0, s0, ld), ld);
  if (own) acc_load64<T>(acc, C, ld);
  // the own block staged in xs for the product after the inverse; for a
  // symmetric cell U_{s0,bi} = A_{bi,s0}^T is stored now (its stores overlap
  // the inverse)
  if (own) {
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
      for (int tj = 0; tj < 2; ++tj)
#pragma unroll
        for (int r = 0; r < 4; ++r) xs[(wr + ti * 16 + acc_row(T(0), lk, r)) * P + wc + tj * 16 + lr] = acc[ti][tj][r];
  }
  T *a = us;
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) a[(wr + ti * 16 + acc_row(T(0), lk, r)) * PA + wc + tj * 16 + lr] = dacc[ti][tj][r];
  __syncthreads();
  if (own && v.sym) {
    T *U = A + v.off(s0, bi, ld);
    for (int idx = threadIdx.x; idx < 4096; idx += blockDim.x) {
      const int r = idx >> 6, cc = idx & 63;
      U[(long long)r * ld + cc] = xs[cc * P + r];
    }
  }
  T *D = Dinv + dinv_off(p, v.rb(s0));
  if (!own) {
    const int fixed = invert_diag<T, PA, true>(a, D, D + 4096, (T)cs[c].tau);
    if (fixed) atomicAdd(&cs[c].n_static, fixed);
    return;
  }
  (void)invert_diag<T, PA, false>(a, D, D + 4096, (T)cs[c].tau);
  // L = A_{bi,s0} A_{s0,s0}^-1
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[ti][tj][r] = T(0);
  mma_64<T, P, PA>(xs, a, acc, T(1));
  acc_store64<T, 0>(acc, C, ld);
}

// One launch per block step k: every item of every cell in the launch.
template <class T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_lu_step(
    const CellParams *cp, CellState *cs, const int *map, int k, T *K, T *Dinv, int panel, StepFwd fwd) {
  constexpr int P = Pitch<T>::v;
  __shared__ __attribute__((aligned(16))) T ls[64 * P];
  __shared__ __attribute__((aligned(16))) T us[64 * P];
  __shared__ __attribute__((aligned(16))) T xs[64 * P];
  const int item = xcd_item(blockIdx.x, gridDim.x);
  const int v = map[item];
  lu_step_item<T>(cp, cs, v >> 20, v & 0xFFFFF, k, K, Dinv, panel, fwd, ls, us, xs);
}

// The whole band LU of a cell in one workgroup: block steps k = -1 .. nblk-2
// in order, each step's items one after another (the first step in the
// L-item form, the others in the panel form: one inverse per step), with a
// barrier between items.  The items of one step are independent (they run
// concurrently in k_lu_step), so any order inside a step gives the same
// results as the launch-per-step schedule in panel form.  One launch factors
// every cell of the batch; a cell never waits for another cell's steps, and
// no launch sits between two steps of the chain.  map: (cell << 20), largest
// cells first.
template <class T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_lu_cell(
    const CellParams *cp, CellState *cs, const int *map, T *K, T *Dinv, StepFwd fwd) {
  constexpr int P = Pitch<T>::v;
  __shared__ __attribute__((aligned(16))) T ls[64 * P];
  __shared__ __attribute__((aligned(16))) T us[64 * P];
  __shared__ __attribute__((aligned(16))) T xs[64 * P];
  const int c = map[xcd_item(blockIdx.x, gridDim.x)] >> 20;
  const int nblk = chain_steps(cp[c]);
  for (int k = -1; k <= nblk - 2; ++k) {
    const int panel = k >= 0 ? 1 : 0;
    const int nt = cell_step_items(cp[c], k, panel);
    for (int t = 0; t < nt; ++t) {
      lu_step_item<T>(cp, cs, c, t, k, K, Dinv, panel, fwd, ls, us, xs);
      __syncthreads();  // the item's global writes are visible to the next item
    }
  }
}

template <class T>
void launch_lu_cell(hipStream_t s, const CellParams *cp, CellState *cs, const WorkList &wl, T *K, T *Dinv,
                    StepFwd fwd, hipEvent_t ev_start, hipEvent_t ev_stop) {
  if (wl.total <= 0) return;
  if constexpr (sizeof(T) == 4) {
    if (ev_start)
      hipExtLaunchKernelGGL(k_lu_cell<T>, dim3(wl.total), dim3(256), 0, s, ev_start, ev_stop, 0, cp, cs, wl.map, K,
                            Dinv, fwd);
    else
      hipLaunchKernelGGL(k_lu_cell<T>, dim3(wl.total), dim3(256), 0, s, cp, cs, wl.map, K, Dinv, fwd);
  }
}

template <class T>
void launch_lu_step(hipStream_t s, const CellParams *cp, CellState *cs, const WorkList &wl, int k, T *K, T *Dinv,
                    int panel, StepFwd fwd, hipEvent_t ev_start, hipEvent_t ev_stop) {
  if (wl.total <= 0) return;
  // FP32 only: three 64 x 68 LDS tiles (the FP64 variant would exceed the
  // static LDS limit); the schedule never selects it for FP64 factors
  if constexpr (sizeof(T) == 4) {
    if (ev_start)  // profiling (never under graph capture): events at the kernel's own start and end
      hipExtLaunchKernelGGL(k_lu_step<T>, dim3(wl.total), dim3(256), 0, s, ev_start, ev_stop, 0, cp, cs, wl.map, k,
                            K, Dinv, panel, fwd);
    else
      hipLaunchKernelGGL(k_lu_step<T>, dim3(wl.total), dim3(256), 0, s, cp, cs, wl.map, k, K, Dinv, panel, fwd);
  }
}

// ---------------------------------------------------------------------------
// triangular solves with the factors, one right-hand side per cell, the
// whole vector in LDS; diagonal blocks through the stored inverses
// ---------------------------------------------------------------------------
constexpr int SOLVE_THREADS = 512;

// y[k0 .. k0+63] <- Minv (64 x 64, row-major) * y[k0 .. k0+63]; 8 threads per row
template <class T>
__device__ __forceinline__ void diag_matvec(T *y, int k0, const T *Minv, T *tmp) {
  const int r = threadIdx.x >> 3, q = threadIdx.x & 7;
  T s = T(0);
  if (r < 64) {
    const T *row = Minv + r * 64 + q * 8;
#pragma unroll
    for (int c = 0; c < 8; ++c) s += row[c] * y[k0 + q * 8 + c];
  }
  s += __shfl_xor(s, 1, 64);
  s += __shfl_xor(s, 2, 64);
  s += __shfl_xor(s, 4, 64);
  if (r < 64 && q == 0) tmp[r] = s;
  __syncthreads();
  if (threadIdx.x < 64) y[k0 + threadIdx.x] = tmp[threadIdx.x];
  __syncthreads();
}

// y[rr] -= A[rr, k0 .. k0+63] . y[k0 .. k0+63] for rr in [r0, r1): a group of
// G lanes reads one row segment (G x 16 bytes, one coalesced 256 / 512-byte
// run), the partial dots are reduced across the group with shuffles.  Four
// rows per group are loaded before any arithmetic so the loads overlap.
template <class T>
__device__ __forceinline__ void block_column_update(T *y, const T *A, long long ld, int k0, int r0, int r1) {
  constexpr int V = 16 / sizeof(T);   // elements per 16-byte load
  constexpr int G = 64 / V;           // lanes per row segment
  constexpr int ROWS = SOLVE_THREADS / G;
  const int q = threadIdx.x % G, g = threadIdx.x / G;
  T yk[V];
#pragma unroll
  for (int e = 0; e < V; ++e) yk[e] = y[k0 + q * V + e];
  for (int base = r0; base < r1; base += 4 * ROWS) {
    T a[4][V];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int rr = base + u * ROWS + g;
      if (rr < r1) {
        const T *src = A + (long long)rr * ld + k0 + q * V;
        if constexpr (V == 4) {
          const float4 v = *reinterpret_cast<const float4 *>(src);
          a[u][0] = v.x; a[u][1] = v.y; a[u][2] = v.z; a[u][3] = v.w;
        } else {
          const double2 v = *reinterpret_cast<const double2 *>(src);
          a[u][0] = (T)v.x; a[u][1] = (T)v.y;
        }
      } else {
#pragma unroll
        for (int e = 0; e < V; ++e) a[u][e] = T(0);
      }
    }
    T sacc[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      T t = T(0);
#pragma unroll
      for (int e = 0; e < V; ++e) t += a[u][e] * yk[e];
#pragma unroll
      for (int m = 1; m < G; m <<= 1) t += __shfl_xor(t, m, 64);
      sacc[u] = t;
    }
    if (q == 0)
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int rr = base + u * ROWS + g;
        if (rr < r1) y[rr] -= sacc[u];
      }
  }
}

template <class T, bool PIVOT>
__global__ __launch_bounds__(SOLVE_THREADS) void k_lu_solve(const CellParams *cp, const CellState *cs,
                                                           const int *cells, const T *K, const T *Dinv,
                                                           const int *ipiv, const double *rhs, double *delta,
                                                           int accumulate) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  __shared__ T tmp[64];
  T *y = reinterpret_cast<T *>(smem);
  const int c = cells[blockIdx.x];
  if (cs[c].ir_done) return;
  const CellParams p = cp[c];
  if (p.npad == 0) return;  // sparse cell: solved by PCG
  const T *A = K + p.k_off;
  const long long ld = p.npad;
  const double *r = rhs + (long long)p.vec_off;
  for (int i = threadIdx.x; i < p.npad; i += SOLVE_THREADS) y[i] = i < p.ndof ? (T)r[i] : T(0);
  __syncthreads();
  // forward: P, unit lower L (LINPACK order: block interchanges, then block)
  for (int kb = 0; kb < p.nblk; ++kb) {
    const int k0 = kb * NB;
    if (PIVOT) {
      if (threadIdx.x == 0) {
        const int *pv = ipiv + p.sys_off;
        for (int j = 0; j < 64; ++j) {
          const int q = pv[k0 + j];
          if (q != k0 + j) {
            const T tt = y[k0 + j];
            y[k0 + j] = y[q];
            y[q] = tt;
          }
        }
      }
      __syncthreads();
    }
    diag_matvec<T>(y, k0, Dinv + dinv_off(p, kb), tmp);
    block_column_update<T>(y, A, ld, k0, k0 + NB, min(p.npad, (kb + 1 + p.bw) * NB));  // rows in the band
    __syncthreads();
  }
  // backward: U with its diagonal
  for (int kb = p.nblk - 1; kb >= 0; --kb) {
    const int k0 = kb * NB;
    diag_matvec<T>(y, k0, Dinv + dinv_off(p, kb) + 4096, tmp);
    block_column_update<T>(y, A, ld, k0, max(0, (kb - p.bw) * NB), k0);
    __syncthreads();
  }
  double *dd = delta + (long long)p.vec_off;
  for (int i = threadIdx.x; i < p.ndof; i += SOLVE_THREADS) dd[i] = accumulate ? dd[i] + (double)y[i] : (double)y[i];
}

// Workgroup barrier for LDS only: outstanding global loads stay in flight
// (a full __syncthreads also waits for them, which would serialise the
// prefetch below).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// ---------------------------------------------------------------------------
// Triangular sweeps of a band-factored cell by one workgroup (static
// pivoting, block bandwidth bw <= BWM), left-looking, in the block LDU form
// of the static-pivoting factorisation (Li = I, Ui = A_bb^-1):
//   forward   y_b = r_b - sum_{j=b-bw}^{b-1} L_bj y_j
//   backward  x_b = Ui_b (y_b - sum_{j=b+1}^{b+bw} U_bj x_j)
// Thread (row r, part q) = (tid / 8, tid % 8) owns eight columns of row r
// of every 64-row block (scol: 4q .. 4q+3 and 32+4q .. 32+4q+3 in FP32).  The band blocks and the diagonal inverse of the
// next block are loaded into registers while the current block is solved
// (they do not depend on the solution), so each block step costs its
// arithmetic and two LDS barriers, not a memory round trip.
// ---------------------------------------------------------------------------
constexpr int BAND_THREADS = 512;

// Column of a sweep thread's e-th entry (e = 0..7) of a 64-wide block row.
// FP32: columns 4q .. 4q+3 and 32+4q .. 32+4q+3 (two float4), so each load
// instruction of a wave takes whole 128-byte row segments (with 8q .. 8q+7 a
// wave's instruction would take every other 16 bytes of 256 and the next
// one the rest); FP64: 8q .. 8q+7.
template <class T> __device__ __forceinline__ int scol(int q, int e) {
  if constexpr (sizeof(T) == 4) return e < 4 ? 4 * q + e : 28 + 4 * q + e;
  else return 8 * q + e;
}

template <class T, int BWM>
struct BandRegs {
  T a[BWM][8];
  T d[8];
};

template <class T, int BWM, bool FWD>
__device__ __forceinline__ void band_load(BandRegs<T, BWM> &R, const CellParams &p, const T *A, const T *Dinv, int b,
                                          int bw) {
  const int r = threadIdx.x >> 3, q = threadIdx.x & 7;
  const long long ld = p.npad;
  const bool ok = b >= 0 && b < p.nblk;
#pragma unroll
  for (int u = 0; u < BWM; ++u) {
    const int j = FWD ? b - bw + u : b + 1 + u;
    if (ok && u < bw && j >= 0 && j < p.nblk) {
      const T *src = A + (long long)(b * NB + r) * ld + j * NB;
#pragma unroll
      for (int e = 0; e < 8; ++e) R.a[u][e] = src[scol<T>(q, e)];
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) R.a[u][e] = T(0);
    }
  }
  if (ok && !FWD) {  // forward: the block-LDU lower factor has identity diagonal blocks
    const T *src = Dinv + dinv_off(p, b) + 4096 + r * 64;
#pragma unroll
    for (int e = 0; e < 8; ++e) R.d[e] = src[scol<T>(q, e)];
  }
}

// sum over the 8 lanes of each aligned group, in every lane.  FP32 by DPP
// (quad_perm 1,0,3,2 and 2,3,0,1, then row_half_mirror: lane i adds the other
// quad's sum); the partial sums and their order are the xor butterfly's, so
// the result is the same bit for bit.
template <class T> __device__ __forceinline__ T sum8(T v) {
  if constexpr (sizeof(T) == 4) {
    auto dpp = [](float x, auto ctrl) {
      return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), decltype(ctrl)::value, 0xF, 0xF, false));
    };
    v += dpp(v, std::integral_constant<int, 0xB1>{});
    v += dpp(v, std::integral_constant<int, 0x4E>{});
    v += dpp(v, std::integral_constant<int, 0x141>{});
    return v;
  } else {
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    v += __shfl_xor(v, 4, 64);
    return v;
  }
}

// a sweep thread's 8 entries (columns scol(q, e)) of a 64-entry LDS vector
// block at p: two 16-byte reads (FP32)
template <class T> __device__ __forceinline__ void lds8(T (&v)[8], const T *p, int q) {
  if constexpr (sizeof(T) == 4) {
    const float4 a = *reinterpret_cast<const float4 *>(p + 4 * q), b = *reinterpret_cast<const float4 *>(p + 32 + 4 * q);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
  } else {
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = p[8 * q + e];
  }
}

template <class T, int BWM, bool FWD>
__device__ __forceinline__ void band_step(const BandRegs<T, BWM> &R, T *y, T *sb, int b, int bw, int nb) {
  const int r = threadIdx.x >> 3, q = threadIdx.x & 7;
  T part = T(0);
#pragma unroll
  for (int u = 0; u < BWM; ++u) {
    const int j = FWD ? b - bw + u : b + 1 + u;
    if (u < bw && j >= 0 && j < nb) {
      T yv[8];
      lds8<T>(yv, y + j * NB, q);
#pragma unroll
      for (int e = 0; e < 8; ++e) part += R.a[u][e] * yv[e];
    }
  }
  part = sum8(part);
  if (FWD) {  // identity diagonal block (static pivoting stores Li = I)
    if (q == 0) y[b * NB + r] -= part;
    lds_barrier();
    return;
  }
  if (q == 0) sb[r] = y[b * NB + r] - part;
  lds_barrier();
  T t = T(0);
  T sv[8];
  lds8<T>(sv, sb, q);
#pragma unroll
  for (int e = 0; e < 8; ++e) t += R.d[e] * sv[e];
  t = sum8(t);
  if (q == 0) y[b * NB + r] = t;
  lds_barrier();
}

// Sweeps of a two-chain cell (tw_m > 0).  The factorization is the block
// LU of Q A Q^T, Q the order [top blocks 0 .. m-1, bottom blocks n-1 down
// to m+bw, middle window m .. m+bw-1]: top and bottom do not couple, so
//   forward   chain A: top rows 0 .. m-1; chain B: bottom rows (virtual
//             order), then B's L blocks of the middle rows (X_B y_B)
//   middle    the middle rows' forward with A's L blocks, then their
//             backward with A's U blocks inside the window
//   backward  chain A: top rows m-1 .. 0; chain B: bottom rows nb-1 .. 0
// Each chain runs in a workgroup of its own (a compute unit streams one
// chain's factors): a forward launch, then a backward launch in which both
// workgroups of the cell solve the middle window (the same operations on the
// same data: the same bits) and then their own chain.  Nothing waits across
// workgroups.  One job is one block row: virtual row `row` of a chain (real
// block rb(row)), column blocks j0 .. j0+nj-1 (virtual), and for the
// backward sweep the diagonal inverse.
struct SweepJob {
  int row, j0, nj, flip, rmax;
  __device__ int rb(int v) const { return flip ? rmax - v : v; }
};

template <class T, int BWM, bool FWD>
__device__ __forceinline__ void job_load(BandRegs<T, BWM> &R, const SweepJob &s, const CellParams &p, const T *A,
                                         const T *Dinv) {
  const int r = threadIdx.x >> 3, q = threadIdx.x & 7;
  const long long ld = p.npad;
#pragma unroll
  for (int u = 0; u < BWM; ++u) {
    if (s.row >= 0 && u < s.nj) {
      const T *src = A + (long long)(s.rb(s.row) * NB + r) * ld + s.rb(s.j0 + u) * NB;
#pragma unroll
      for (int e = 0; e < 8; ++e) R.a[u][e] = src[scol<T>(q, e)];
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) R.a[u][e] = T(0);
    }
  }
  if (s.row >= 0 && !FWD) {
    const T *src = Dinv + dinv_off(p, s.rb(s.row)) + 4096 + r * 64;
#pragma unroll
    for (int e = 0; e < 8; ++e) R.d[e] = src[scol<T>(q, e)];
  }
}

// steps 0 .. L-1 of one chain; job(s) gives step s's block row (row -1:
// none); a ring of PF block rows in registers as in the one-chain sweep
template <class T, int BWM, bool FWD, class J>
__device__ __forceinline__ void sweep_chain(const CellParams &p, const T *A, const T *Dinv, T *y, T *sb, int L,
                                            J job) {
  const int r = threadIdx.x >> 3, q = threadIdx.x & 7;
  constexpr int PF = (sizeof(T) == 4 && BWM <= 4) ? 4 : 2;
  BandRegs<T, BWM> R[PF];
#pragma unroll
  for (int i = 0; i < PF; ++i) job_load<T, BWM, FWD>(R[i], job(i), p, A, Dinv);
  lds_barrier();
  for (int s0 = 0; s0 < L; s0 += PF) {
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      const int st = s0 + i;
      if (st < L) {
        const SweepJob a = job(st);
        T part = T(0);
#pragma unroll
        for (int u = 0; u < BWM; ++u) {
          if (a.row >= 0 && u < a.nj) {
            T yv[8];
            lds8<T>(yv, y + a.rb(a.j0 + u) * NB, q);
#pragma unroll
            for (int e = 0; e < 8; ++e) part += R[i].a[u][e] * yv[e];
          }
        }
        part = sum8(part);
        if (a.row >= 0) {
          const int o = a.rb(a.row) * NB + r;
          if (FWD) {  // identity diagonal blocks (Li = I)
            if (q == 0) y[o] -= part;
            lds_barrier();
          } else {
            if (q == 0) sb[r] = y[o] - part;
            lds_barrier();
            T t = T(0);
            T sv[8];
            lds8<T>(sv, sb, q);
#pragma unroll
            for (int e = 0; e < 8; ++e) t += R[i].d[e] * sv[e];
            t = sum8(t);
            if (q == 0) y[o] = t;
            lds_barrier();
          }
        }
      }
      job_load<T, BWM, FWD>(R[i], job(st + PF), p, A, Dinv);
    }
  }
}

// the right-hand side in the cell's band numbering: node n's rows go to
// 2 kpos[n] (the padding rows stay zero)
template <class T>
__device__ __forceinline__ void band_rhs(const CellParams &p, const int *kpos, const double *rhs, T *y) {
  for (int i = threadIdx.x; i < p.npad; i += BAND_THREADS) y[i] = T(0);
  lds_barrier();
  for (int m = threadIdx.x; m < p.n_nodes; m += BAND_THREADS) {
    const int n = p.node_off + m, o = 2 * kpos[n];
    y[o] = (T)rhs[2LL * n];
    y[o + 1] = (T)rhs[2LL * n + 1];
  }
}

// back to the batch numbering for the nodes whose band rows lie in
// [row_lo, row_hi); accumulate: delta += the solution
template <class T>
__device__ __forceinline__ void band_delta(const CellParams &p, const int *kpos, const T *y, double *delta,
                                           int accumulate, int row_lo, int row_hi) {
  for (int m = threadIdx.x; m < p.n_nodes; m += BAND_THREADS) {
    const int n = p.node_off + m, o = 2 * kpos[n];
    if (o < row_lo || o >= row_hi) continue;
    double *d = delta + 2LL * n;
    const double d0 = (double)y[o], d1 = (double)y[o + 1];
    d[0] = accumulate ? d[0] + d0 : d0;
    d[1] = accumulate ? d[1] + d1 : d1;
  }
}

// the one-chain sweeps of a plain cell: forward (skipped when the fused
// block steps ran it), backward
template <class T, int BWM, bool FWD>
__device__ __forceinline__ void band_sweep_plain(const CellParams &p, const T *A, const T *Dinv, T *y, T *sb) {
  const int bw = p.bw, nb = p.nblk;
  // a ring of PF block rows in registers: the loads of block b + PF are in
  // flight while blocks b + 1 .. b + PF - 1 are applied
  constexpr int PF = (sizeof(T) == 4 && BWM <= 4) ? 4 : 2;
  BandRegs<T, BWM> R[PF];
  if (FWD) {
#pragma unroll
    for (int i = 0; i < PF; ++i) band_load<T, BWM, true>(R[i], p, A, Dinv, i, bw);
    lds_barrier();
    for (int b = 0; b < nb; b += PF) {
#pragma unroll
      for (int i = 0; i < PF; ++i) {
        if (b + i < nb) band_step<T, BWM, true>(R[i], y, sb, b + i, bw, nb);
        band_load<T, BWM, true>(R[i], p, A, Dinv, b + i + PF, bw);
      }
    }
  } else {
#pragma unroll
    for (int i = 0; i < PF; ++i) band_load<T, BWM, false>(R[i], p, A, Dinv, nb - 1 - i, bw);
    for (int b = nb - 1; b >= 0; b -= PF) {
#pragma unroll
      for (int i = 0; i < PF; ++i) {
        if (b - i >= 0) band_step<T, BWM, false>(R[i], y, sb, b - i, bw, nb);
        band_load<T, BWM, false>(R[i], p, A, Dinv, b - i - PF, bw);
      }
    }
  }
}

template <class T, int BWM>
__global__ __launch_bounds__(BAND_THREADS) void k_lu_solve_band(const CellParams *cp, const CellState *cs,
                                                                const int *cells, const T *K, const T *Dinv,
                                                                const int *kpos, const double *rhs, double *delta,
                                                                int accumulate, const float *yfwd) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  __shared__ __attribute__((aligned(32))) T sb[64];
  T *y = reinterpret_cast<T *>(smem);
  const int c = cells[blockIdx.x];
  if (cs[c].ir_done) return;
  const CellParams p = cp[c];
  if (p.npad == 0) return;  // sparse cell: solved by PCG
  const T *A = K + p.k_off;  // bw <= BWM (checked by the launcher's caller)
  // yfwd: the forward sweep already ran inside the fused block steps, y is its result
  if (yfwd) {
    for (int i = threadIdx.x; i < p.npad; i += BAND_THREADS) y[i] = (T)yfwd[p.sys_off + i];
  } else {
    band_rhs<T>(p, kpos, rhs, y);
    lds_barrier();
    band_sweep_plain<T, BWM, true>(p, A, Dinv, y, sb);
  }
  lds_barrier();
  band_sweep_plain<T, BWM, false>(p, A, Dinv, y, sb);
  band_delta<T>(p, kpos, y, delta, accumulate, 0, p.npad);
}

// Batches with two-chain cells: a forward launch and a backward launch, two
// workgroups per cell (the second returns at once for a one-chain cell); the
// forward result goes through ysw (per cell at sys_off, band numbering).  The
// first solve after the fused block steps has its forward result in yfwd
// (every cell, the window rows included) and no forward launch.
template <class T, int BWM>
__global__ __launch_bounds__(BAND_THREADS) void k_lu_sweep_fwd(const CellParams *cp, const CellState *cs,
                                                               const int *cells, const T *K, const T *Dinv,
                                                               const int *kpos, const double *rhs, T *ysw) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  __shared__ __attribute__((aligned(32))) T sb[64];
  T *y = reinterpret_cast<T *>(smem);
  const int c = cells[blockIdx.x >> 1], part = blockIdx.x & 1;
  if (cs[c].ir_done) return;
  const CellParams p = cp[c];
  if (p.npad == 0) return;
  if (p.tw_m == 0 && part) return;  // one-chain cell: one workgroup
  const T *A = K + p.k_off;
  band_rhs<T>(p, kpos, rhs, y);
  lds_barrier();
  int lo = 0, hi = p.npad;
  if (p.tw_m == 0) {
    band_sweep_plain<T, BWM, true>(p, A, Dinv, y, sb);
  } else {
    const int n = p.nblk, bw = p.bw, m = p.tw_m, nb = p.tw_nb;
    if (part == 0) {  // chain A: the top rows
      sweep_chain<T, BWM, true>(p, A, Dinv, y, sb, m, [&](int s) {
        if (s >= m) return SweepJob{-1, 0, 0, 0, 0};
        const int j0 = max(0, s - bw);
        return SweepJob{s, j0, s - j0, 0, 0};
      });
      hi = m * NB;
    } else {  // chain B: the bottom rows, then its L blocks of the middle rows
      sweep_chain<T, BWM, true>(p, A, Dinv, y, sb, nb + bw, [&](int s) {
        if (s >= nb + bw) return SweepJob{-1, 0, 0, 1, n - 1};
        const int j0 = max(0, s - bw);
        return SweepJob{s, j0, min(s, nb) - j0, 1, n - 1};
      });
      lo = m * NB;
    }
  }
  lds_barrier();
  T *g = ysw + p.sys_off;
  for (int i = lo + threadIdx.x; i < hi; i += BAND_THREADS) g[i] = y[i];
}

template <class T, int BWM>
__global__ __launch_bounds__(BAND_THREADS) void k_lu_sweep_bwd(const CellParams *cp, const CellState *cs,
                                                               const int *cells, const T *K, const T *Dinv,
                                                               const int *kpos, double *delta, int accumulate,
                                                               const float *yfwd, const T *ysw) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  __shared__ __attribute__((aligned(32))) T sb[64];
  T *y = reinterpret_cast<T *>(smem);
  const int c = cells[blockIdx.x >> 1], part = blockIdx.x & 1;
  if (cs[c].ir_done) return;
  const CellParams p = cp[c];
  if (p.npad == 0) return;
  if (p.tw_m == 0 && part) return;
  const T *A = K + p.k_off;
  if (yfwd) {  // the first solve: the fused block steps ran the forward sweep (the window rows' too)
    for (int i = threadIdx.x; i < p.npad; i += BAND_THREADS) y[i] = (T)yfwd[p.sys_off + i];
  } else {
    const T *g = ysw + p.sys_off;
    for (int i = threadIdx.x; i < p.npad; i += BAND_THREADS) y[i] = g[i];
  }
  lds_barrier();
  if (p.tw_m == 0) {
    band_sweep_plain<T, BWM, false>(p, A, Dinv, y, sb);
    band_delta<T>(p, kpos, y, delta, accumulate, 0, p.npad);
    return;
  }
  const int n = p.nblk, bw = p.bw, m = p.tw_m, nb = p.tw_nb, mid = m + bw;
  // the middle window (both workgroups): forward with A's L blocks (unless
  // the block steps did), then backward
  if (!yfwd) {
    sweep_chain<T, BWM, true>(p, A, Dinv, y, sb, bw, [&](int s) {
      if (s >= bw) return SweepJob{-1, 0, 0, 0, 0};
      const int i = m + s, j0 = max(0, i - bw);
      return SweepJob{i, j0, i - j0, 0, 0};
    });
    lds_barrier();
  }
  sweep_chain<T, BWM, false>(p, A, Dinv, y, sb, bw, [&](int s) {
    if (s >= bw) return SweepJob{-1, 0, 0, 0, 0};
    const int i = mid - 1 - s;
    return SweepJob{i, i + 1, min(i + bw, mid - 1) - i, 0, 0};
  });
  lds_barrier();
  if (part == 0) {  // chain A: top rows m-1 .. 0; writes the top and middle nodes
    sweep_chain<T, BWM, false>(p, A, Dinv, y, sb, m, [&](int s) {
      const int i = m - 1 - s;
      if (i < 0) return SweepJob{-1, 0, 0, 0, 0};
      return SweepJob{i, i + 1, min(i + bw, mid - 1) - i, 0, 0};
    });
    lds_barrier();
    band_delta<T>(p, kpos, y, delta, accumulate, 0, mid * NB);
  } else {  // chain B: bottom rows nb-1 .. 0 (virtual); writes the bottom nodes
    sweep_chain<T, BWM, false>(p, A, Dinv, y, sb, nb, [&](int s) {
      const int v = nb - 1 - s;
      if (v < 0) return SweepJob{-1, 0, 0, 1, n - 1};
      return SweepJob{v, v + 1, min(v + bw, nb + bw - 1) - v, 1, n - 1};
    });
    lds_barrier();
    band_delta<T>(p, kpos, y, delta, accumulate, mid * NB, p.npad);
  }
}

template <class T>
void launch_lu_solve_band(hipStream_t s, const CellParams *cp, const CellState *cs, const int *cells, int ncell,
                          const T *K, const T *Dinv, const int *kpos, const double *r, double *delta, int accumulate,
                          int max_npad, int max_bw, const float *yfwd, T *ysw) {
  if (ncell <= 0) return;
  const size_t lds = lu_solve_lds_bytes(max_npad, sizeof(T));
  if (ysw) {  // two-chain cells in the batch (FP32, bw <= 4); no forward launch after the fused block steps
    if (max_bw <= 2) {
      if (!yfwd)
        hipLaunchKernelGGL((k_lu_sweep_fwd<T, 2>), dim3(2 * ncell), dim3(BAND_THREADS), lds, s, cp, cs, cells, K,
                           Dinv, kpos, r, ysw);
      hipLaunchKernelGGL((k_lu_sweep_bwd<T, 2>), dim3(2 * ncell), dim3(BAND_THREADS), lds, s, cp, cs, cells, K, Dinv,
                         kpos, delta, accumulate, yfwd, ysw);
    } else {
      if (!yfwd)
        hipLaunchKernelGGL((k_lu_sweep_fwd<T, 4>), dim3(2 * ncell), dim3(BAND_THREADS), lds, s, cp, cs, cells, K,
                           Dinv, kpos, r, ysw);
      hipLaunchKernelGGL((k_lu_sweep_bwd<T, 4>), dim3(2 * ncell), dim3(BAND_THREADS), lds, s, cp, cs, cells, K, Dinv,
                         kpos, delta, accumulate, yfwd, ysw);
    }
    return;
  }
  if (max_bw <= 2)
    hipLaunchKernelGGL((k_lu_solve_band<T, 2>), dim3(ncell), dim3(BAND_THREADS), lds, s, cp, cs, cells, K, Dinv, kpos, r,
                       delta, accumulate, yfwd);
  else if (max_bw <= 4)
    hipLaunchKernelGGL((k_lu_solve_band<T, 4>), dim3(ncell), dim3(BAND_THREADS), lds, s, cp, cs, cells, K, Dinv, kpos, r,
                       delta, accumulate, yfwd);
  else
    hipLaunchKernelGGL((k_lu_solve_band<T, 8>), dim3(ncell), dim3(BAND_THREADS), lds, s, cp, cs, cells, K, Dinv, kpos, r,
                       delta, accumulate, yfwd);
}

// ---------------------------------------------------------------------------
// Triangular solves for static pivoting across many workgroups: one workgroup
// per 64-row block row, every block row of every cell in one launch.
//   forward   y_k = L_kk^-1 (b_k - sum_{j<k} L_kj y_j)
//   backward  x_k = U_kk^-1 (y_k - sum_{j>k} U_kj x_j)      (x overwrites y)
// A workgroup streams its block row while the blocks it depends on are still
// being solved.  The solution entries are exchanged as self-validating 64-bit
// words (32-bit payload, 32-bit epoch; two words per FP64 entry) written and
// read with agent-scope atomics: a reader spins on the words it needs until
// they carry the current epoch, so no flag, fence or cache maintenance sits
// on the dependency chain.  The epoch advances by 2 per solve (forward E,
// backward E + 1), so stale words are never mistaken for new ones.
// Workgroups take their work item from a ticket counter, so every item a
// workgroup waits for belongs to a workgroup that started earlier; every wait
// is bounded (SPIN_LIMIT polls), after which the cell is flagged
// (CellState::solve_stall) and the workgroup finishes without waiting again.
// sync: [0] forward ticket, [1] backward ticket, [2] epoch, [3] any cell of
//       the list still refining (0: both sweeps return at once).
// ---------------------------------------------------------------------------
constexpr int MW_THREADS = 256;
constexpr int SPIN_LIMIT = 1 << 18;

__device__ __forceinline__ int mw_ticket(int *ticket, int *sh) {
  if (threadIdx.x == 0) *sh = atomicAdd(ticket, 1);
  __syncthreads();
  return *sh;
}

template <class T> struct MwShape {
  static constexpr int V = 16 / sizeof(T);  // elements per 16-byte load
  static constexpr int G = 64 / V;          // lanes per 64-element row segment
  static constexpr int RP = MW_THREADS / G; // rows per pass
  static constexpr int NP = 64 / RP;        // passes per 64-row block
  static constexpr int WPE = sizeof(T) / 4; // 64-bit words per exchanged entry
};

// part[u] += M[row_u, 0..63] . v[0..63] for this thread's slice, M a 64 x 64
// block (leading dimension ld), rows row_u = g + RP u, columns q V .. q V + V-1
template <class T>
__device__ __forceinline__ void mw_load_block(T (&a)[MwShape<T>::NP][MwShape<T>::V], const T *M, long long ld) {
  typedef MwShape<T> S;
  const int q = threadIdx.x % S::G, g = threadIdx.x / S::G;
#pragma unroll
  for (int u = 0; u < S::NP; ++u) {
    const T *src = M + (long long)(g + S::RP * u) * ld + q * S::V;
    if constexpr (S::V == 4) {
      const float4 v = *reinterpret_cast<const float4 *>(src);
      a[u][0] = v.x; a[u][1] = v.y; a[u][2] = v.z; a[u][3] = v.w;
    } else {
      const double2 v = *reinterpret_cast<const double2 *>(src);
      a[u][0] = v.x; a[u][1] = v.y;
    }
  }
}

template <class T>
__device__ __forceinline__ T mw_group_sum(T t) {
#pragma unroll
  for (int m = 1; m < MwShape<T>::G; m <<= 1) t += __shfl_xor(t, m, 64);
  return t;
}

__device__ __forceinline__ unsigned long long mw_word_load(const unsigned long long *w) {
  return __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void mw_word_store(unsigned long long *w, unsigned payload, unsigned epoch) {
  __hip_atomic_store(w, ((unsigned long long)epoch << 32) | payload, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Word w of the exchange buffer, already loaded into `v`: spin until it
// carries `epoch` (bounded).
__device__ __forceinline__ void mw_settle(unsigned long long &v, const unsigned long long *w, unsigned epoch,
                                          bool &gave_up, int *stall) {
  for (int spins = 0; (unsigned)(v >> 32) != epoch && !gave_up; ++spins) {
    if (spins >= SPIN_LIMIT) {
      gave_up = true;
      *stall = 1;
      break;
    }
    __builtin_amdgcn_s_sleep(2);
    v = mw_word_load(w);
  }
}

template <class T>
__device__ __forceinline__ T mw_decode(const unsigned long long *w) {
  if constexpr (MwShape<T>::WPE == 1) {
    return __uint_as_float((unsigned)w[0]);
  } else {
    return __hiloint2double((int)(unsigned)w[1], (int)(unsigned)w[0]);
  }
}

// Entry e of the exchange buffer, waiting until it carries `epoch`.
template <class T>
__device__ __forceinline__ T mw_get(const unsigned long long *yw, long long e, unsigned epoch, bool &gave_up,
                                    int *stall) {
  constexpr int WPE = MwShape<T>::WPE;
  unsigned long long w[WPE];
#pragma unroll
  for (int k = 0; k < WPE; ++k) w[k] = mw_word_load(yw + e * WPE + k);
#pragma unroll
  for (int k = 0; k < WPE; ++k) mw_settle(w[k], yw + e * WPE + k, epoch, gave_up, stall);
  return mw_decode<T>(w);
}

template <class T>
__device__ __forceinline__ void mw_put(unsigned long long *yw, long long e, T v, unsigned epoch) {
  if constexpr (MwShape<T>::WPE == 1) {
    mw_word_store(yw + e, __float_as_uint(v), epoch);
  } else {
    mw_word_store(yw + 2 * e, (unsigned)__double2loint(v), epoch);
    mw_word_store(yw + 2 * e + 1, (unsigned)__double2hiint(v), epoch);
  }
}

// One workgroup solves two consecutive block rows of the sweep: `rf`, then
// `rs` (rs = rf + 1 forward, rf - 1 backward; absent past the end), so the
// dependency chain has one cross-workgroup hop per two blocks.
template <class T>
__device__ __forceinline__ void mw_fetch_step(unsigned long long (&w)[MwShape<T>::V * MwShape<T>::WPE],
                                              const unsigned long long *ycw, int j, unsigned epoch, bool valid) {
  typedef MwShape<T> S;
  constexpr int NW = S::V * S::WPE;
  const int q = threadIdx.x % S::G;
  const long long base = (long long)j * NB * S::WPE + q * NW;
#pragma unroll
  for (int k = 0; k < NW; ++k) w[k] = valid ? mw_word_load(ycw + base + k) : ((unsigned long long)epoch << 32);
}

template <class T, bool FORWARD>
__global__ __launch_bounds__(MW_THREADS) void k_lu_solve_mw(const CellParams *cp, CellState *cs, const int *map,
                                                            const T *K, const T *Dinv, const double *rhs,
                                                            unsigned long long *yw, double *delta, const int *sync_c,
                                                            int *tickets, int accumulate) {
  typedef MwShape<T> S;
  constexpr int NW = S::V * S::WPE;
  constexpr int UNR = sizeof(T) == 4 ? 2 : 1;  // steps in flight per thread
  __shared__ int sh_ticket;
  __shared__ T vin[64];
  __shared__ T yf[64];
  if (sync_c[3] == 0) return;  // no cell of the list is refining
  const unsigned epoch_f = (unsigned)sync_c[2], epoch = epoch_f + (FORWARD ? 0u : 1u);
  const int item = map[mw_ticket(tickets + (FORWARD ? 0 : 1), &sh_ticket)];
  const int c = item >> 20, pi = item & 0xFFFFF;
  if (cs[c].ir_done) return;
  const CellParams p = cp[c];
  const long long ld = p.npad;
  // processing position of the first row; block of step s
  const int pf = 2 * pi;
  auto blk = [&](int st) { return FORWARD ? st : p.nblk - 1 - st; };
  const int rf = blk(pf);
  const bool two = pf + 1 < p.nblk;
  const int rs = two ? blk(pf + 1) : rf;
  const T *Af = K + p.k_off + (long long)(rf * NB) * ld;
  const T *As = K + p.k_off + (long long)(rs * NB) * ld;
  unsigned long long *ycw = yw + (long long)p.sys_off * S::WPE;
  int *stall = &cs[c].solve_stall;
  const int q = threadIdx.x % S::G, g = threadIdx.x / S::G;
  bool gave_up = false;
  // loaded before any wait: both diagonal inverses and the block (rs, rf)
  T dinf[S::NP][S::V], dins[S::NP][S::V], xs[S::NP][S::V];
  mw_load_block<T>(dinf, Dinv + dinv_off(p, rf) + (FORWARD ? 0 : 4096), 64);
  mw_load_block<T>(dins, Dinv + dinv_off(p, rs) + (FORWARD ? 0 : 4096), 64);
  mw_load_block<T>(xs, As + rf * NB, ld);
  T pf_[S::NP], ps_[S::NP];
#pragma unroll
  for (int u = 0; u < S::NP; ++u) pf_[u] = ps_[u] = T(0);
  // the blocks before rf within the band, in the order they are solved (the
  // first of them is one block beyond rs's band: zero in the fill envelope)
  for (int st = max(0, pf - p.bw); st < pf; st += UNR) {
    T a0[UNR][S::NP][S::V], a1[UNR][S::NP][S::V];
    unsigned long long w[UNR][NW];
#pragma unroll
    for (int i = 0; i < UNR; ++i) {
      const bool ok = st + i < pf;
      const int j = blk(ok ? st + i : st);
      if (ok) {
        mw_load_block<T>(a0[i], Af + j * NB, ld);
        if (two) mw_load_block<T>(a1[i], As + j * NB, ld);
      }
      mw_fetch_step<T>(w[i], ycw, j, epoch, ok);
    }
#pragma unroll
    for (int i = 0; i < UNR; ++i) {
      if (st + i >= pf) continue;
      const long long base = (long long)blk(st + i) * NB * S::WPE + q * NW;
#pragma unroll
      for (int k = 0; k < NW; ++k) mw_settle(w[i][k], ycw + base + k, epoch, gave_up, stall);
      T yv[S::V];
#pragma unroll
      for (int e = 0; e < S::V; ++e) yv[e] = mw_decode<T>(&w[i][e * S::WPE]);
#pragma unroll
      for (int u = 0; u < S::NP; ++u)
#pragma unroll
        for (int e = 0; e < S::V; ++e) {
          pf_[u] += a0[i][u][e] * yv[e];
          if (two) ps_[u] += a1[i][u][e] * yv[e];
        }
    }
  }
  // first row: y_rf = D_rf^-1 (b_rf - sum)
  const int k0f = rf * NB, k0s = rs * NB;
#pragma unroll
  for (int u = 0; u < S::NP; ++u) {
    const T sum = mw_group_sum<T>(pf_[u]);
    const int row = g + S::RP * u;
    if (q == 0) {
      T b;
      if (FORWARD) b = (k0f + row < p.ndof) ? (T)rhs[(long long)p.vec_off + k0f + row] : T(0);
      else b = mw_get<T>(ycw, k0f + row, epoch_f, gave_up, stall);
      vin[row] = b - sum;
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < S::NP; ++u) {
    T t = T(0);
#pragma unroll
    for (int e = 0; e < S::V; ++e) t += dinf[u][e] * vin[q * S::V + e];
    t = mw_group_sum<T>(t);
    const int row = g + S::RP * u;
    if (q == 0) {
      mw_put<T>(ycw, k0f + row, t, epoch);
      yf[row] = t;
      if (!FORWARD && k0f + row < p.ndof) {
        double *dd = delta + (long long)p.vec_off + k0f + row;
        *dd = accumulate ? *dd + (double)t : (double)t;
      }
    }
  }
  if (!two) return;
  __syncthreads();
  // second row: y_rs = D_rs^-1 (b_rs - sum - A(rs, rf) y_rf)
#pragma unroll
  for (int u = 0; u < S::NP; ++u) {
#pragma unroll
    for (int e = 0; e < S::V; ++e) ps_[u] += xs[u][e] * yf[q * S::V + e];
    const T sum = mw_group_sum<T>(ps_[u]);
    const int row = g + S::RP * u;
    if (q == 0) {
      T b;
      if (FORWARD) b = (k0s + row < p.ndof) ? (T)rhs[(long long)p.vec_off + k0s + row] : T(0);
      else b = mw_get<T>(ycw, k0s + row, epoch_f, gave_up, stall);
      vin[row] = b - sum;
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < S::NP; ++u) {
    T t = T(0);
#pragma unroll
    for (int e = 0; e < S::V; ++e) t += dins[u][e] * vin[q * S::V + e];
    t = mw_group_sum<T>(t);
    const int row = g + S::RP * u;
    if (q == 0) {
      mw_put<T>(ycw, k0s + row, t, epoch);
      if (!FORWARD && k0s + row < p.ndof) {
        double *dd = delta + (long long)p.vec_off + k0s + row;
        *dd = accumulate ? *dd + (double)t : (double)t;
      }
    }
  }
}

// Reset the tickets, advance the epoch and record whether any cell of the
// list is still refining (one workgroup).
__global__ __launch_bounds__(256) void k_solve_prepare(int *sync, const CellState *cs, const int *cells, int ncell) {
  __shared__ int any;
  if (threadIdx.x == 0) any = 0;
  __syncthreads();
  int a = 0;
  for (int i = threadIdx.x; i < ncell; i += 256) a |= cs[cells[i]].ir_done ? 0 : 1;
  if (a) any = 1;
  __syncthreads();
  if (threadIdx.x == 0) {
    sync[0] = 0;
    sync[1] = 0;
    sync[2] += 2;
    sync[3] = any;
  }
}

template <class T>
void launch_lu_solve_mw(hipStream_t s, const CellParams *cp, CellState *cs, const int *cells, int ncell,
                        const WorkList &fwd, const WorkList &bwd, const T *K, const T *Dinv, const double *r,
                        double *delta, void *yw, int *sync, int accumulate) {
  if (fwd.total <= 0) return;
  unsigned long long *w = static_cast<unsigned long long *>(yw);
  hipLaunchKernelGGL(k_solve_prepare, dim3(1), dim3(256), 0, s, sync, cs, cells, ncell);
  hipLaunchKernelGGL((k_lu_solve_mw<T, true>), dim3(fwd.total), dim3(MW_THREADS), 0, s, cp, cs, fwd.map, K, Dinv, r,
                     w, delta, sync, sync, accumulate);
  hipLaunchKernelGGL((k_lu_solve_mw<T, false>), dim3(bwd.total), dim3(MW_THREADS), 0, s, cp, cs, bwd.map, K, Dinv,
                     r, w, delta, sync, sync, accumulate);
}

size_t lu_solve_exchange_bytes(long long rows, int elem_size) { return (size_t)rows * 2 * elem_size; }

size_t lu_solve_lds_bytes(int npad, int elem_size) { return (size_t)npad * elem_size; }

// Dynamic LDS available to k_lu_solve next to its own static LDS.
size_t lu_solve_max_dynamic(size_t max_lds) { return max_lds > 1024 ? max_lds - 1024 : 0; }

hipError_t lu_prepare_attributes(size_t max_lds) {
  const int dyn = (int)lu_solve_max_dynamic(max_lds);
  const void *fns[] = {reinterpret_cast<const void *>(&k_lu_solve<float, true>),
                       reinterpret_cast<const void *>(&k_lu_solve<float, false>),
                       reinterpret_cast<const void *>(&k_lu_solve<double, true>),
                       reinterpret_cast<const void *>(&k_lu_solve<double, false>),
                       reinterpret_cast<const void *>(&k_lu_solve_band<float, 2>),
                       reinterpret_cast<const void *>(&k_lu_solve_band<float, 4>),
                       reinterpret_cast<const void *>(&k_lu_solve_band<float, 8>),
                       reinterpret_cast<const void *>(&k_lu_solve_band<double, 2>),
                       reinterpret_cast<const void *>(&k_lu_solve_band<double, 4>),
                       reinterpret_cast<const void *>(&k_lu_solve_band<double, 8>),
                       reinterpret_cast<const void *>(&k_lu_sweep_fwd<float, 2>),
                       reinterpret_cast<const void *>(&k_lu_sweep_fwd<float, 4>),
                       reinterpret_cast<const void *>(&k_lu_sweep_fwd<double, 2>),
                       reinterpret_cast<const void *>(&k_lu_sweep_fwd<double, 4>),
                       reinterpret_cast<const void *>(&k_lu_sweep_bwd<float, 2>),
                       reinterpret_cast<const void *>(&k_lu_sweep_bwd<float, 4>),
                       reinterpret_cast<const void *>(&k_lu_sweep_bwd<double, 2>),
                       reinterpret_cast<const void *>(&k_lu_sweep_bwd<double, 4>)};
  for (const void *f : fns) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, dyn);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template <class T>
void launch_lu_solve(hipStream_t s, const CellParams *cp, const CellState *cs, const int *cells, int ncell,
                     const T *K, const T *Dinv, const int *ipiv, int pivoting, const double *r, double *delta,
                     int max_npad, int accumulate) {
  if (ncell <= 0) return;
  const size_t lds = lu_solve_lds_bytes(max_npad, sizeof(T));
  if (pivoting)
    hipLaunchKernelGGL((k_lu_solve<T, true>), dim3(ncell), dim3(SOLVE_THREADS), lds, s, cp, cs, cells, K, Dinv, ipiv,
                       r, delta, accumulate);
  else
    hipLaunchKernelGGL((k_lu_solve<T, false>), dim3(ncell), dim3(SOLVE_THREADS), lds, s, cp, cs, cells, K, Dinv,
                       ipiv, r, delta, accumulate);
}

#define WCFE_INST(T)                                                                                                 \
  template void launch_lu_solve_band<T>(hipStream_t, const CellParams *, const CellState *, const int *, int,        \
                                        const T *, const T *, const int *, const double *, double *, int, int, int,  \
                                        const float *, T *);                                                              \
  template void launch_lu_diag_inv<T>(hipStream_t, const CellParams *, CellState *, const WorkList &, int, T *, T *); \
  template void launch_lu_panel_pp<T>(hipStream_t, const CellParams *, CellState *, const WorkList &, int, T *,       \
                                      int *, T *);                                                                    \
  template void launch_lu_laswp<T>(hipStream_t, const CellParams *, const WorkList &, int, T *, const int *);         \
  template void launch_lu_trsm<T>(hipStream_t, const CellParams *, CellState *, const WorkList &, int, T *,           \
                                  const T *);                                                                         \
  template void launch_lu_gemm<T>(hipStream_t, const CellParams *, CellState *, const WorkList &, int, T *, T *,      \
                                  int, int, int, hipEvent_t, hipEvent_t);                                             \
  template void launch_lu_step<T>(hipStream_t, const CellParams *, CellState *, const WorkList &, int, T *, T *, int, \
                                  StepFwd,                                                                         \
                                  hipEvent_t, hipEvent_t);                                                            \
  template void launch_lu_cell<T>(hipStream_t, const CellParams *, CellState *, const WorkList &, T *, T *, StepFwd,  \
                                  hipEvent_t, hipEvent_t);                                                            \
  template void launch_lu_solve<T>(hipStream_t, const CellParams *, const CellState *, const int *, int, const T *,   \
                                   const T *, const int *, int, const double *, double *, int, int);                  \
  template void launch_lu_solve_mw<T>(hipStream_t, const CellParams *, CellState *, const int *, int,               \
                                      const WorkList &, const WorkList &, const T *, const T *, const double *,      \
                                      double *, void *, int *, int);
WCFE_INST(float)
WCFE_INST(double)

}  // namespace wcfe
