// k_common.hpp — device helpers shared by the gfx950 kernels.
#pragma once

#include <climits>

#include "pnm_internal.hpp"

namespace pnm {

constexpr int kMaxSteps = 512;  // walk cap, pore_extraction.hpp:278 and :627

// Segmentation tile: 64 (x) x TILE_Y (y) x TILE_Z (z) voxels per 512-thread block.
constexpr int TILE_Y = 8;
constexpr int TILE_Z = 16;

// Periodic +-1 neighbours; equal to the reference's ((x%N)+N)%N for x in [-1, N]
// (pore_extraction.hpp:59-64), including N = 1 and N = 2 where neighbours alias.
// Division of 32-bit values by a divisor fixed for a launch: q0 = mulhi(n, floor(2^32 / d)) is
// q or q - 1, one correction makes it exact (d >= 2; d == 1 passes magic 0 and is special-cased).
struct UDiv {
  uint32_t d, magic;
};
inline UDiv make_udiv(uint32_t d) {
  return UDiv{d, d >= 2 ? (uint32_t)((1ull << 32) / d) : 0u};
}
__device__ __forceinline__ uint32_t udiv(uint32_t n, UDiv v, uint32_t& rem) {
  if (v.d == 1u) {
    rem = 0u;
    return n;
  }
  uint32_t q = __umulhi(n, v.magic);
  uint32_t r = n - q * v.d;
  if (r >= v.d) {
    ++q;
    r -= v.d;
  }
  rem = r;
  return q;
}
__device__ __forceinline__ int wm1(int x, int N) { return x > 0 ? x - 1 : N - 1; }
__device__ __forceinline__ int wp1(int x, int N) { return x + 1 < N ? x + 1 : 0; }

// Any v: the reference's wrap. v in [-N, 2N): two compares.
__device__ __forceinline__ int wrap_any(int v, int N) { return ((v % N) + N) % N; }
__device__ __forceinline__ int wrap_near(int v, int N) {
  return v < 0 ? v + N : (v >= N ? v - N : v);
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// Set bits of the wave mask m in the lanes below this one: v_mbcnt_lo/hi, two instructions (a
// popcount of m & below-mask takes two ANDs and two bit counts).
__device__ __forceinline__ unsigned lanes_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Neighbouring lanes of a wave by a DPP wave shift (one VALU move, no LDS crossbar):
// lane_prev gives lane - 1's value (lane 0 gets `first`), lane_next lane + 1's (lane 63 gets
// `last`).
__device__ __forceinline__ uint32_t lane_prev(uint32_t v, uint32_t first) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)first, (int)v, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ uint32_t lane_next(uint32_t v, uint32_t last) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)last, (int)v, 0x130, 0xf, 0xf, false);
}

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
  return base + (uint64_t)lanes_below(m);
}

// Coherent (L1-bypassing) load for union-find parents that other waves are lowering.
__device__ __forceinline__ int ld_agent(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int64_t ld_agent64(const int64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

inline unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

// One step of the gradient walk (pore_extraction.hpp:280-304) over the voxel and its 26 periodic
// neighbours in dz -> dy -> dx order: ascent (kUp) takes a larger value, descent a smaller one,
// either takes an equal value with a larger index than the current best; NaN never wins.
template <bool kUp>
__device__ __forceinline__ int gradient_best(const float* __restrict__ sdf, int x, int y, int z,
                                             int nx, int ny, int nz, int ci, float cv) {
  const int xs[3] = {wm1(x, nx), x, wp1(x, nx)};
  const int ys[3] = {wm1(y, ny), y, wp1(y, ny)};
  const int zs[3] = {wm1(z, nz), z, wp1(z, nz)};
  int best = ci;
  float bv = cv;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const int row = (zs[a] * ny + ys[b]) * nx;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (a == 1 && b == 1 && c == 1) continue;
        const int ni = row + xs[c];
        const float nv = sdf[ni];
        if (kUp ? (nv > bv) : (nv < bv)) {
          bv = nv;
          best = ni;
        } else if (nv == bv && ni > best) {
          best = ni;
        }
      }
    }
  return best;
}

// The ascent step (lexicographic (sdf, flat index) maximum).
__device__ __forceinline__ int ascent_best(const float* __restrict__ sdf, int x, int y, int z,
                                           int nx, int ny, int nz, int ci, float cv) {
  return gradient_best<true>(sdf, x, y, z, nx, ny, nz, ci, cv);
}

// (value, flat index) packed into one ordered 64-bit key: the value's bits mapped to an unsigned
// order in the high word, the index in the low word. `key(nv,ni) > key(bv,best)` is exactly the
// reference's `nv > bv || (nv == bv && ni > best)` (pore_extraction.hpp:97-99, :296-298) for
// numbers: -0 is folded onto +0 (they compare equal) and NaN maps to 0, below every number, so a
// NaN neighbour never wins. A NaN centre needs its own test (nothing beats it).
__device__ __forceinline__ uint64_t lex_key(float v, uint32_t idx) {
  uint32_t u = __float_as_uint(v + 0.0f);  // -0 -> +0
  u ^= (uint32_t)((int32_t)u >> 31) | 0x80000000u;  // negative: all bits flipped; else the sign
  if (v != v) u = 0u;
  return ((uint64_t)u << 32) | idx;
}
// Inverse of lex_key's value mapping (-0 comes back as +0, NaN as a quiet NaN).
__device__ __forceinline__ float key_value(uint64_t k) {
  const uint32_t u = (uint32_t)(k >> 32);
  if (u == 0u) return __uint_as_float(0x7fc00000u);
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}
__device__ __forceinline__ uint64_t kmax(uint64_t a, uint64_t b) { return a > b ? a : b; }

// Walk record of an sdf>0 voxel after its ascent walk inside its tile: target (low word) is the
// root if the walk is final (fixed point, or the 512-step cap: `capped`), else the first voxel
// outside the tile; steps = moves made to reach it.
__device__ __forceinline__ uint64_t walk_record(int target, int steps, bool fin, bool capped) {
  return (uint64_t)(uint32_t)target | ((uint64_t)steps << 32) | ((uint64_t)fin << 48) |
         ((uint64_t)capped << 49);
}
__device__ __forceinline__ int walk_target(uint64_t r) { return (int)(uint32_t)r; }
__device__ __forceinline__ int walk_steps(uint64_t r) { return (int)((r >> 32) & 0xffffu); }
__device__ __forceinline__ bool walk_final(uint64_t r) { return (r >> 48) & 1u; }
__device__ __forceinline__ bool walk_capped(uint64_t r) { return (r >> 49) & 1u; }

// Lock-free union-find over int32 parents (root = smallest index of its tree): the larger root
// is hooked under the smaller with atomicMin; a lost race retries with the value it found.
// find() reads through the L1 and halves the path with plain stores: every value it can read or
// write is an ancestor of the node, so a stale read only shortens progress, never connectivity,
// and the atomicMin results (coherent) decide every hook.
__device__ __forceinline__ int uf_find(int* L, int x) {
  int p = L[x];
  while (p != x) {
    const int g = L[p];
    if (g != p) L[x] = g;
    x = p;
    p = g;
  }
  return x;
}

__device__ __forceinline__ void uf_unite(int* L, int a, int b) {
  while (true) {
    a = uf_find(L, a);
    b = uf_find(L, b);
    if (a == b) return;
    if (a > b) {
      const int t = a;
      a = b;
      b = t;
    }
    const int old = atomicMin(L + b, a);
    if (old == b) return;
    b = old;
  }
}

// Block-ordered staging slot for `pred`: items are ordered by wave, then lane. Every thread of
// the block calls it (one barrier); `wcnt` is an LDS array of NW words used by this call only.
template <int NW>
__device__ __forceinline__ unsigned block_stage_pos(bool pred, unsigned* wcnt, int wave, int lane,
                                                    unsigned& total) {
  const uint64_t m = __ballot(pred);
  if (lane == 0) wcnt[wave] = (unsigned)__popcll(m);
  __syncthreads();
  unsigned base = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < NW; ++k) {
    const unsigned cc = wcnt[k];
    if (k < wave) base += cc;
    tot += cc;
  }
  total = tot;
  return base + lanes_below(m);
}

// Wave-aggregated append to an LDS counter (batched kernels). Every lane must call it.
__device__ __forceinline__ unsigned wave_append_lds(unsigned* ctr, bool pred, int lane) {
  const uint64_t m = __ballot(pred);
  if (m == 0) return 0;
  const int leader = __ffsll((unsigned long long)m) - 1;
  unsigned base = 0;
  if (lane == leader) base = atomicAdd(ctr, (unsigned)__popcll(m));
  base = __shfl(base, leader, 64);
  return base + lanes_below(m);
}

// merge_markers (pore_extraction.hpp:198-224) over x-runs. Every marker carries the index of
// the start of its x-run within its 64-voxel row chunk. A wave owns one row chunk; for each of
// the forward rows of the 13-neighbourhood (:204-206) — (y+1, z), (y-1, z+1), (y, z+1),
// (y+1, z+1) — it unites each of its runs with every run of that row it touches (26-adjacency:
// columns x-1, x, x+1). One union per touching pair suffices: at column x where both rows are
// markers (skipped when column x-1 gave the same pair), and at a run's first / last column for
// the diagonal-only contact. The +x neighbour across chunk (and periodic row) boundaries is the
// remaining forward relation. The resulting components, hence the component minima, are those of
// the voxel-level relation.
// Called by a whole wave for one 64-voxel row chunk (x = chunk start + lane).
__device__ __forceinline__ void runs_union_chunk(int* L, int nx, int ny, int nz, int x, int y,
                                                 int z, int lane) {
  const bool in = (x < nx) && (y < ny);
  const int i = in ? (z * ny + y) * nx + x : 0;
  const int my = in ? L[i] : -1;
  const bool mk = my != -1;
  const uint64_t m = __ballot(mk);
  if (!__any(mk)) return;  // wave-uniform
  const bool run_start = mk && (lane == 0 || !((m >> (lane - 1)) & 1ull));
  const bool run_end = mk && (lane == 63 || !((m >> (lane + 1)) & 1ull));
  const int xm = in ? wm1(x, nx) : 0, xp = in ? wp1(x, nx) : 0;
  // +x relation across the chunk start; at x = 0 this is the periodic wrap from x = nx-1
  if (mk && lane == 0) {
    const int nl = L[(z * ny + y) * nx + xm];
    if (nl != -1) uf_unite(L, my, nl);
  }
  const int yp = in ? wp1(y, ny) : 0, ym = in ? wm1(y, ny) : 0, zp = wp1(z, nz);
  const int rows[4] = {(z * ny + yp) * nx, (zp * ny + ym) * nx, (zp * ny + y) * nx,
                       (zp * ny + yp) * nx};
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int c = in ? L[rows[r] + x] : -1;
    const bool cm = c != -1;
    // previous lane's pair for the duplicate test at column x
    const int prev_c = __shfl_up(c, 1, 64);
    if (mk && cm) {
      const bool dup = lane > 0 && ((m >> (lane - 1)) & 1ull) && prev_c == c;
      if (!dup) uf_unite(L, my, c);
    }
    if (mk && !cm) {
      if (run_start) {
        const int a = L[rows[r] + xm];
        if (a != -1) uf_unite(L, my, a);
      }
      if (run_end) {
        const int b = L[rows[r] + xp];
        if (b != -1) uf_unite(L, my, b);
      }
    }
  }
}

}  // namespace pnm
