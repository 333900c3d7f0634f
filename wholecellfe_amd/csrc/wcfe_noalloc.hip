// calculateForce_FA on device-resident inputs with no device allocation.
//
// wcfe_force_fa (the reference-shaped call behind calculateForce_FA,
// UL_growth_explicit.cpp:1198 / UL1:1196) copies host arrays in and out and
// builds each node's one-ring triangle list on the host.  This variant serves
// callers that cannot allocate device memory (code under stream capture,
// serving loops, harnesses that forbid hipMalloc): the caller passes device
// pointers for the inputs and the output, and the one-ring lists are built on
// the device in a file-scope scratch array.  The force kernel and the order of
// each node's triangle list (increasing triangle index) are wcfe_force_fa's,
// so both calls produce the same bits.
//
// The scratch array is one per device for the life of the process.  Every
// call on a device shares it and leaves its contents behind (the counters are
// cleared by a fill kernel at the start of each call): two calls that overlap
// on one device race.
#include <mutex>

#include "../../include/wcfe.h"
#include "wcfe_internal.hpp"

namespace wcfe {

constexpr int NA_MAX_NODES = 262144;
constexpr int NA_MAX_TRIS = 524288;
// count (nodes) | start (nodes + 1) | ring (3 tris) | bad index counter (1)
constexpr long long NA_INTS = (long long)NA_MAX_NODES + (NA_MAX_NODES + 1) + 3LL * NA_MAX_TRIS + 1;

static __device__ __align__(256) int scratch[NA_INTS];

namespace {
constexpr int NA_MAX_DEVICES = 64;
std::mutex na_mutex;
int *na_ptr[NA_MAX_DEVICES] = {};  // device address of `scratch` per device

__global__ __launch_bounds__(256) void k_na_zero(int *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// triangles per node (atomics: the counts do not depend on the order)
__global__ __launch_bounds__(256) void k_na_count(int n_nodes, int n_tris, const int *tris, int *cnt, int *bad) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_tris) return;
  for (int a = 0; a < 3; ++a) {
    const int v = tris[3 * e + a];
    if (v < 0 || v >= n_nodes) {
      atomicAdd(bad, 1);
      continue;
    }
    atomicAdd(&cnt[v], 1);
  }
}

// start = exclusive prefix sum of cnt (one workgroup of 1024 threads, each a
// contiguous chunk), then cnt is cleared for use as the fill cursor
__global__ __launch_bounds__(1024) void k_na_scan(int n_nodes, int *cnt, int *start) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int chunk = (n_nodes + 1023) / 1024;
  const int b = min(n_nodes, t * chunk), e = min(n_nodes, b + chunk);
  int s = 0;
  for (int i = b; i < e; ++i) s += cnt[i];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {  // inclusive scan (Hillis-Steele)
    const int v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = t > 0 ? part[t - 1] : 0;
  for (int i = b; i < e; ++i) {
    start[i] = run;
    run += cnt[i];
    cnt[i] = 0;
  }
  if (t == 1023) start[n_nodes] = part[1023];
}

__global__ __launch_bounds__(256) void k_na_fill(int n_nodes, int n_tris, const int *tris, const int *start, int *cur,
                                                 int *ring) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_tris) return;
  for (int a = 0; a < 3; ++a) {
    const int v = tris[3 * e + a];
    if (v < 0 || v >= n_nodes) continue;
    ring[start[v] + atomicAdd(&cur[v], 1)] = e;
  }
}

// each node's list in increasing triangle index (the host lists' order)
__global__ __launch_bounds__(256) void k_na_sort(int n_nodes, const int *start, int *ring) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n_nodes) return;
  int *r = ring + start[v];
  const int n = start[v + 1] - start[v];
  for (int i = 1; i < n; ++i) {
    const int key = r[i];
    int j = i - 1;
    while (j >= 0 && r[j] > key) {
      r[j + 1] = r[j];
      --j;
    }
    r[j + 1] = key;
  }
}

__global__ void k_na_report(const int *bad, int *out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *out = *bad;
}
}  // namespace

// Device address of the scratch array on `device` (which must be current),
// looked up once per device.
hipError_t noalloc_init(int device) {
  if (device < 0 || device >= NA_MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(na_mutex);
  if (na_ptr[device]) return hipSuccess;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  na_ptr[device] = static_cast<int *>(p);
  return hipSuccess;
}

// Enqueue the no-allocation calculateForce_FA on stream s of `device`.
// Returns WCFE_OK or WCFE_ERR_ARG (message in err).
int noalloc_force_fa(int device, hipStream_t s, int n_nodes, int n_tris, const int *d_tris, const double *d_x,
                     const double *d_cC, const double *d_fau, int law, double kint, double *d_F, int *d_n_bad,
                     std::string &err) {
  if (n_nodes < 1 || n_tris < 1 || n_nodes > NA_MAX_NODES || n_tris > NA_MAX_TRIS) {
    err = "wcfe_force_fa_noalloc: " + std::to_string(n_nodes) + " nodes / " + std::to_string(n_tris) +
          " triangles outside 1.." + std::to_string(NA_MAX_NODES) + " / 1.." + std::to_string(NA_MAX_TRIS) +
          " (the scratch array's capacity)";
    return WCFE_ERR_ARG;
  }
  int *base = nullptr;
  {
    std::lock_guard<std::mutex> lock(na_mutex);
    base = (device >= 0 && device < NA_MAX_DEVICES) ? na_ptr[device] : nullptr;
  }
  if (!base) {
    err = "wcfe_force_fa_noalloc: scratch array of device " + std::to_string(device) + " not initialised";
    return WCFE_ERR_STATE;
  }
  int *cnt = base, *start = base + NA_MAX_NODES, *ring = start + NA_MAX_NODES + 1,
      *bad = ring + 3LL * NA_MAX_TRIS;
  const int gn = (n_nodes + 255) / 256, gt = (n_tris + 255) / 256;
  hipLaunchKernelGGL(k_na_zero, dim3(gn), dim3(256), 0, s, cnt, n_nodes);
  hipLaunchKernelGGL(k_na_zero, dim3(1), dim3(256), 0, s, bad, 1);
  hipLaunchKernelGGL(k_na_count, dim3(gt), dim3(256), 0, s, n_nodes, n_tris, d_tris, cnt, bad);
  hipLaunchKernelGGL(k_na_scan, dim3(1), dim3(1024), 0, s, n_nodes, cnt, start);
  hipLaunchKernelGGL(k_na_fill, dim3(gt), dim3(256), 0, s, n_nodes, n_tris, d_tris, start, cnt, ring);
  hipLaunchKernelGGL(k_na_sort, dim3(gn), dim3(256), 0, s, n_nodes, start, ring);
  launch_ref_force_fa(s, n_nodes, start, ring, d_tris, d_x, d_cC, d_fau, law, kint, d_F);
  if (d_n_bad) hipLaunchKernelGGL(k_na_report, dim3(1), dim3(64), 0, s, bad, d_n_bad);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    err = std::string("wcfe_force_fa_noalloc: ") + hipGetErrorString(e);
    return WCFE_ERR_HIP;
  }
  return WCFE_OK;
}

}  // namespace wcfe
