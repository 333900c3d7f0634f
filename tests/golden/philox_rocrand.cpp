// Prints rocRAND's Philox4x32-10 (host-callable ten_rounds of
// rocrand_device::philox4x32_10_engine) for counters/keys read from stdin,
// one "c0 c1 c2 c3 k0 k1" line each, as "o0 o1 o2 o3" lines.  Used only to
// write the fixture tests/golden/philox_rocrand.npz.
#include <hip/hip_runtime.h>
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
// ten_rounds is a protected member of the engine; a derived class exposes it.
struct Philox : rocrand_device::philox4x32_10_engine {
  using rocrand_device::philox4x32_10_engine::ten_rounds;
};
int main() {
  Philox eng;
  unsigned c0, c1, c2, c3, k0, k1;
  while (std::scanf("%u %u %u %u %u %u", &c0, &c1, &c2, &c3, &k0, &k1) == 6) {
    uint4 o = eng.ten_rounds(uint4{c0, c1, c2, c3}, uint2{k0, k1});
    std::printf("%u %u %u %u\n", o.x, o.y, o.z, o.w);
  }
  return 0;
}
