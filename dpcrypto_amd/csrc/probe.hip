// Latency probe for the serial hash chains: one wave, 32 lane-pair Keccak
// states, `perms` back-to-back permutations in the fully unrolled form the
// chain kernels squeeze with (keccak2.hpp).  The states are written out so
// nothing is folded away.  Timed by dpkem_probe_perm_latency (engine.cpp).
#include <hip/hip_runtime.h>

#include "keccak2.hpp"
#include "launch.hpp"

namespace dpkem {

__global__ void __launch_bounds__(64) probe_perm(int perms, uint32_t* __restrict__ out) {
  const uint32_t p = threadIdx.x & 1;
  uint32_t a[25];
#pragma unroll
  for (int i = 0; i < 25; ++i) a[i] = (uint32_t)threadIdx.x * 0x9E3779B9u + (uint32_t)i;
  for (int r = 0; r < perms; ++r) k2::keccakf2_unrolled(a, p);
  uint32_t x = 0;
#pragma unroll
  for (int i = 0; i < 25; ++i) x ^= a[i];
  out[threadIdx.x] = x;
}

void saber_probe_perm_launch(hipStream_t st, int perms, uint32_t* out) {
  hipLaunchKernelGGL(probe_perm, dim3(1), dim3(64), 0, st, perms, out);
}

}  // namespace dpkem
