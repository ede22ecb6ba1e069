// k_common.hpp — device helpers shared by the gfx950 kernels.
#pragma once

#include <climits>

#include "pnm_internal.hpp"

namespace pnm {

constexpr int kMaxSteps = 512;  // walk cap, pore_extraction.hpp:278 and :627

// Periodic +-1 neighbours; equal to the reference's ((x%N)+N)%N for x in [-1, N]
// (pore_extraction.hpp:59-64), including N = 1 and N = 2 where neighbours alias.
__device__ __forceinline__ int wm1(int x, int N) { return x > 0 ? x - 1 : N - 1; }
__device__ __forceinline__ int wp1(int x, int N) { return x + 1 < N ? x + 1 : 0; }

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// Wave-aggregated append: one atomic per wave. Every lane of the wave must call it.
// Returns the slot of a lane with pred set (undefined otherwise).
__device__ __forceinline__ uint64_t wave_append(unsigned long long* counter, bool pred,
                                                int lane) {
  const uint64_t m = __ballot(pred);
  if (m == 0) return 0;
  const int leader = __ffsll((unsigned long long)m) - 1;
  unsigned long long base = 0;
  if (lane == leader) base = atomicAdd(counter, (unsigned long long)__popcll(m));
  base = __shfl(base, leader, 64);
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  return base + (uint64_t)__popcll(m & below);
}

// Coherent (L1-bypassing) load for union-find parents that other waves are lowering.
__device__ __forceinline__ int ld_agent(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int64_t ld_agent64(const int64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

inline unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace pnm
