// k_util.hip — generic device utilities for gfx950 (wave64): fills, readback, exclusive scan,
// stable LSD radix sort (8-bit digits, wave-ballot ranking), unique of a sorted list.
#include "k_common.hpp"

namespace pnm {

// ---------------- fills ----------------
template <class T>
__global__ void k_fill(T* __restrict__ p, T v, int64_t count) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) p[i] = v;
}

static unsigned fill_blocks(int64_t count) {
  int64_t b = (count + 255) / 256;
  if (b > 65536) b = 65536;
  return (unsigned)(b < 1 ? 1 : b);
}

// 16-byte stores for an aligned body, the tail (fewer than 4 words) by the first threads.
__global__ void k_fill_u32x4(uint32_t* __restrict__ p, uint32_t v, int64_t count) {
  const int64_t n4 = count >> 2;
  uint4* __restrict__ p4 = reinterpret_cast<uint4*>(p);
  const uint4 vv = make_uint4(v, v, v, v);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int64_t i = t; i < n4; i += stride) p4[i] = vv;
  if (t < (count & 3)) p[(n4 << 2) + t] = v;
}

void fill_u32(Ctx* c, uint32_t* p, uint32_t v, int64_t count) {
  if (count <= 0) return;
  if ((reinterpret_cast<uintptr_t>(p) & 15u) == 0 && count >= 1024) {
    PNM_LAUNCH(c, "k_fill", k_fill_u32x4, dim3(fill_blocks((count + 3) / 4)), dim3(256), 0, p, v,
               count);
    return;
  }
  PNM_LAUNCH(c, "k_fill", k_fill<uint32_t>, dim3(fill_blocks(count)), dim3(256), 0, p, v,
                     count);
}

void fill_u64(Ctx* c, uint64_t* p, uint64_t v, int64_t count) {
  if (count <= 0) return;
  PNM_LAUNCH(c, "k_fill", k_fill<uint64_t>, dim3(fill_blocks(count)), dim3(256), 0, p, v,
                     count);
}

void readback(Ctx* c, const void* dev, int words, uint64_t* host_out) {
  if (words > Ctx::kPinnedWords) fail(PNM_EINVAL, "readback: too many words");
  PNM_HIP(hipMemcpyAsync(c->pinned, dev, sizeof(uint64_t) * words, hipMemcpyDeviceToHost,
                         c->stream));
  PNM_HIP(hipStreamSynchronize(c->stream));
  std::memcpy(host_out, c->pinned, sizeof(uint64_t) * words);
  if (c->err_host && *c->err_host)
    fail(PNM_EHIP, "internal error: a scan's look-back wait timed out");
}

// ---------------- exclusive scan (uint32) ----------------
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

// Block-wide exclusive scan of one value per thread (blockDim.x == 256); returns the prefix and
// writes the block total to *total.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* lds, uint32_t* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t inc = wave_incl_scan(v);
  if (lane == 63) lds[w] = inc;
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < SCAN_THREADS / 64; ++k) {
    const uint32_t s = lds[k];
    if (k < w) off += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return off + inc - v;
}

// Single-launch exclusive scan with decoupled look-back. Each block publishes its aggregate, then
// its inclusive prefix, in one 64-bit status word per block: value in the low half, (epoch << 2 |
// state) in the high half, so a reader gets value and state in one load and words left from an
// earlier call (another epoch) read as "not ready" — nothing is cleared between calls.
// Block order comes from a ticket taken when the block starts, so a block only waits for blocks
// that are already running and never for one that waits on it; the wait is still bounded (a
// timed-out wait sets the context's error word, which the next readback reports).
// The block that finishes last resets the ticket and done counters and advances the epoch for the
// next call on the same stream.
constexpr uint32_t LB_AGG = 1, LB_INC = 2;
constexpr uint32_t LB_SPIN_LIMIT = 1u << 18;

__device__ __forceinline__ uint64_t lb_word(uint32_t epoch, uint32_t st, uint32_t v) {
  return ((uint64_t)((epoch << 2) | st) << 32) | v;
}

// Thread 0 takes the block's ticket and reads the epoch into LDS; every thread returns after
// the barrier.
__device__ __forceinline__ void lb_begin(uint32_t* ctl, uint32_t& bid, uint32_t& epoch,
                                         uint32_t* s_bid, uint32_t* s_epoch) {
  if (threadIdx.x == 0) {
    *s_bid = atomicAdd(ctl + 0, 1u);
    *s_epoch = __hip_atomic_load(ctl + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  bid = *s_bid;
  epoch = *s_epoch;
}

// Wave 0 (all 64 lanes): publishes the block total `tot`, looks back over its predecessors and
// publishes the inclusive prefix; returns the block's exclusive prefix (in every lane).
__device__ __forceinline__ uint32_t lb_lookback(uint64_t* state, uint32_t bid, uint32_t epoch,
                                                uint32_t tot, uint32_t* err) {
  const int lane = threadIdx.x & 63;
  if (bid == 0) {
    if (lane == 0)
      __hip_atomic_store(state, lb_word(epoch, LB_INC, tot), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    return 0u;
  }
  if (lane == 0)
    __hip_atomic_store(state + bid, lb_word(epoch, LB_AGG, tot), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  uint32_t excl = 0;
  int64_t j = (int64_t)bid - 1;  // nearest predecessor not yet added
  uint32_t spins = 0;
  while (true) {
    const int64_t idx = j - lane;
    uint32_t st = LB_INC, v = 0;  // before block 0: an inclusive zero
    if (idx >= 0) {
      const uint64_t w = __hip_atomic_load(state + idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint32_t hi = (uint32_t)(w >> 32);
      st = (hi >> 2) == epoch ? (hi & 3u) : 0u;
      v = (uint32_t)w;
    }
    const uint64_t notready = __ballot(st == 0u);
    const uint64_t incl = __ballot(st == LB_INC);
    // lanes up to and including the nearest inclusive prefix (all 64 if there is none)
    const uint64_t upto = incl ? (((incl & (0ull - incl)) << 1) - 1ull) : ~0ull;
    if (notready & upto) {
      if (++spins > LB_SPIN_LIMIT) {
        if (lane == 0) *(volatile uint32_t*)err = 1u;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    uint32_t add = ((upto >> lane) & 1ull) ? v : 0u;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) add += __shfl_xor(add, o, 64);
    excl += add;
    if (incl) break;
    j -= 64;
  }
  if (lane == 0)
    __hip_atomic_store(state + bid, lb_word(epoch, LB_INC, excl + tot), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  return excl;
}

// Thread 0 of every block, at its very end: the block that finishes last resets the ticket and
// done counters and advances the epoch for the next call on the same stream.
__device__ __forceinline__ void lb_end(uint32_t* ctl, uint32_t epoch, unsigned nblocks) {
  const uint32_t d = atomicAdd(ctl + 1, 1u);
  if (d == nblocks - 1) {
    const uint32_t ne = (epoch + 1u) & 0x3fffffffu;
    __hip_atomic_store(ctl + 0, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(ctl + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(ctl + 2, ne ? ne : 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// ITEMS consecutive values of one thread; 16-byte vector accesses when the run is whole and
// aligned (only a tail block takes the scalar path).
template <int ITEMS, class T>
__device__ __forceinline__ void load_run(const T* in, int64_t base, int64_t n, T (&v)[ITEMS]) {
  constexpr int PER = 16 / sizeof(T);
  if (base + ITEMS <= n && ((reinterpret_cast<uintptr_t>(in + base) & 15u) == 0)) {
    const uint4* p = reinterpret_cast<const uint4*>(in + base);
#pragma unroll
    for (int q = 0; q < ITEMS / PER; ++q) {
      const uint4 t = p[q];
      T u[PER];
      __builtin_memcpy(u, &t, 16);
#pragma unroll
      for (int e = 0; e < PER; ++e) v[q * PER + e] = u[e];
    }
  } else {
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) v[k] = (base + k < n) ? in[base + k] : (T)0;
  }
}

// Exclusive scan of in[0..n) (in == out allowed): per-thread runs of ITEMS values, block scan,
// look-back across blocks; the last block writes the total.
template <int ITEMS>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_lookback(
    const uint32_t* in, int64_t n, uint32_t* out, uint64_t* __restrict__ total_out,
    uint32_t* __restrict__ ctl, uint64_t* __restrict__ state, unsigned nblocks,
    uint32_t* __restrict__ err) {
  __shared__ uint32_t lds[SCAN_THREADS / 64];
  __shared__ uint32_t s_bid, s_epoch, s_excl;
  uint32_t bid, epoch;
  lb_begin(ctl, bid, epoch, &s_bid, &s_epoch);
  const int64_t base = ((int64_t)bid * SCAN_THREADS + threadIdx.x) * ITEMS;
  uint32_t vals[ITEMS];
  load_run<ITEMS>(in, base, n, vals);
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) s += vals[k];
  uint32_t tot;
  const uint32_t pre = block_excl_scan(s, lds, &tot);
  if (threadIdx.x < 64) {
    const uint32_t excl = lb_lookback(state, bid, epoch, tot, err);
    if (threadIdx.x == 0) s_excl = excl;
  }
  __syncthreads();
  uint32_t run = s_excl + pre;
  if (base + ITEMS <= n && ((reinterpret_cast<uintptr_t>(out + base) & 15u) == 0)) {
    uint4* p = reinterpret_cast<uint4*>(out + base);
#pragma unroll
    for (int q = 0; q < ITEMS / 4; ++q) {
      uint4 t;
      t.x = run; run += vals[4 * q];
      t.y = run; run += vals[4 * q + 1];
      t.z = run; run += vals[4 * q + 2];
      t.w = run; run += vals[4 * q + 3];
      p[q] = t;
    }
  } else {
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      if (base + k < n) out[base + k] = run;
      run += vals[k];
    }
  }
  if (threadIdx.x == 0) {
    if (total_out && bid == nblocks - 1) *total_out = (uint64_t)s_excl + tot;
    lb_end(ctl, epoch, nblocks);
  }
}

// Sorted index list of the set bits of a row-padded bitmap: word w covers voxels
// (w / nwx) * nx + (w % nwx) * 64 + bit, so word order is flat-index order. One launch: each
// thread popcounts ITEMS consecutive words, the look-back scan gives its rank, and it writes the
// indices of its set bits in order.
template <int ITEMS>
__global__ __launch_bounds__(SCAN_THREADS) void k_bitmap_list(
    const uint64_t* __restrict__ map, int64_t nwords, int64_t nwx, int64_t nx,
    uint32_t* __restrict__ out, uint64_t* __restrict__ total_out, uint32_t* __restrict__ ctl,
    uint64_t* __restrict__ state, unsigned nblocks, uint32_t* __restrict__ err) {
  __shared__ uint32_t lds[SCAN_THREADS / 64];
  __shared__ uint32_t s_bid, s_epoch, s_excl;
  uint32_t bid, epoch;
  lb_begin(ctl, bid, epoch, &s_bid, &s_epoch);
  const int64_t base = ((int64_t)bid * SCAN_THREADS + threadIdx.x) * ITEMS;
  uint64_t w[ITEMS];
  load_run<ITEMS>(map, base, nwords, w);
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) s += (uint32_t)__popcll(w[k]);
  uint32_t tot;
  const uint32_t pre = block_excl_scan(s, lds, &tot);
  if (threadIdx.x < 64) {
    const uint32_t excl = lb_lookback(state, bid, epoch, tot, err);
    if (threadIdx.x == 0) s_excl = excl;
  }
  __syncthreads();
  uint32_t r = s_excl + pre;
  for (int k = 0; k < ITEMS; ++k) {
    uint64_t m = w[k];
    if (!m) continue;
    const int64_t wi = base + k;
    const int64_t row = wi / nwx;
    const uint32_t rb = (uint32_t)(row * nx + (wi - row * nwx) * 64);
    while (m) {
      out[r++] = rb + (uint32_t)(__ffsll((unsigned long long)m) - 1);
      m &= m - 1;
    }
  }
  if (threadIdx.x == 0) {
    if (total_out && bid == nblocks - 1) *total_out = (uint64_t)s_excl + tot;
    lb_end(ctl, epoch, nblocks);
  }
}

// Rank directory of a uint32 bitmap: chunk_rank[c] = number of set bits in words [0, 32c).
// The rank of bit f is then chunk_rank[f >> 10] + the popcounts of the words of its chunk
// before word f >> 5 + the bits below f in that word. Each thread counts MAP_RANK_CPT
// consecutive chunks, so that few blocks are chained by the look-back.
constexpr int MAP_RANK_CPT = 4;
__global__ __launch_bounds__(SCAN_THREADS) void k_map_rank(
    const uint32_t* __restrict__ map, int64_t words, uint32_t* __restrict__ chunk_rank,
    uint64_t* __restrict__ total_out, uint32_t* __restrict__ ctl, uint64_t* __restrict__ state,
    unsigned nblocks, uint32_t* __restrict__ err) {
  constexpr int ITEMS = MAP_CHUNK_WORDS;
  __shared__ uint32_t lds[SCAN_THREADS / 64];
  __shared__ uint32_t s_bid, s_epoch, s_excl;
  uint32_t bid, epoch;
  lb_begin(ctl, bid, epoch, &s_bid, &s_epoch);
  const int64_t chunk0 = ((int64_t)bid * SCAN_THREADS + threadIdx.x) * MAP_RANK_CPT;
  uint32_t cc[MAP_RANK_CPT];
  uint32_t s = 0;
#pragma unroll
  for (int q = 0; q < MAP_RANK_CPT; ++q) {
    uint32_t w[ITEMS];
    load_run<ITEMS>(map, (chunk0 + q) * ITEMS, words, w);
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) c += (uint32_t)__popc(w[k]);
    cc[q] = c;
    s += c;
  }
  uint32_t tot;
  const uint32_t pre = block_excl_scan(s, lds, &tot);
  if (threadIdx.x < 64) {
    const uint32_t excl = lb_lookback(state, bid, epoch, tot, err);
    if (threadIdx.x == 0) s_excl = excl;
  }
  __syncthreads();
  uint32_t run = s_excl + pre;
#pragma unroll
  for (int q = 0; q < MAP_RANK_CPT; ++q) {
    if ((chunk0 + q) * ITEMS < words) chunk_rank[chunk0 + q] = run;
    run += cc[q];
  }
  if (threadIdx.x == 0) {
    if (total_out && bid == nblocks - 1) *total_out = (uint64_t)s_excl + tot;
    lb_end(ctl, epoch, nblocks);
  }
}

// Single block: scan of up to SCAN_TILE values, optional total.
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_single(const uint32_t* in, int64_t n,
                                                              uint32_t* out,
                                                              uint64_t* __restrict__ total_out) {
  __shared__ uint32_t lds[SCAN_THREADS / 64];
  const int64_t base = (int64_t)threadIdx.x * SCAN_ITEMS;
  uint32_t vals[SCAN_ITEMS];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    vals[k] = (base + k < n) ? in[base + k] : 0u;
    s += vals[k];
  }
  uint32_t tot;
  uint32_t run = block_excl_scan(s, lds, &tot);
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    if (base + k < n) out[base + k] = run;
    run += vals[k];
  }
  if (total_out && threadIdx.x == 0) *total_out = tot;
}

static void lb_reserve(Ctx* c, int64_t nb) {
  if (!c->lb_ctl) {
    PNM_HIP(hipMalloc(reinterpret_cast<void**>(&c->lb_ctl), 256));
    const uint32_t init[4] = {0u, 0u, 1u, 0u};  // ticket, done, epoch
    PNM_HIP(hipMemcpyAsync(c->lb_ctl, init, sizeof(init), hipMemcpyHostToDevice, c->stream));
    PNM_HIP(hipStreamSynchronize(c->stream));
  }
  if (nb > 0x7fffffff) fail(PNM_EINVAL, "scan: too many blocks");
  if (nb <= c->lb_cap) return;
  if (c->lb_fixed)
    fail(PNM_ECAPACITY, "scan: " + std::to_string(nb) + " blocks exceed the fixed look-back state (" +
                            std::to_string(c->lb_cap) + ")");
  const int64_t cap = std::max<int64_t>(nb, std::max<int64_t>(4096, 2 * c->lb_cap));
  if (c->lb_state) {
    PNM_HIP(hipStreamSynchronize(c->stream));
    PNM_HIP(hipFree(c->lb_state));
    c->lb_state = nullptr;
    c->lb_cap = 0;
  }
  PNM_HIP(hipMalloc(reinterpret_cast<void**>(&c->lb_state), sizeof(uint64_t) * cap));
  // epoch 0 is never used, so zeroed words read as "not ready"
  PNM_HIP(hipMemsetAsync(c->lb_state, 0, sizeof(uint64_t) * cap, c->stream));
  PNM_HIP(hipStreamSynchronize(c->stream));
  c->lb_cap = cap;
}

// Long inputs take 32 values per thread, so that fewer blocks are chained by the look-back.
constexpr int64_t LB_LONG = (int64_t)SCAN_THREADS * 8 * 256;

void bitmap_list(Ctx* c, const uint64_t* map, int64_t nwords, int64_t nwx, int64_t nx,
                 uint32_t* out, uint64_t* total_dev) {
  if (nwords <= 0) {
    fill_u64(c, total_dev, 0, 1);
    return;
  }
  if (nwords > LB_LONG) {
    const int64_t nb = cdiv(nwords, SCAN_THREADS * 32);
    lb_reserve(c, nb);
    PNM_LAUNCH(c, "k_bitmap_list", k_bitmap_list<32>, dim3((unsigned)nb), dim3(SCAN_THREADS), 0,
               map, nwords, nwx, nx, out, total_dev, c->lb_ctl, c->lb_state, (unsigned)nb,
               c->err_dev);
  } else {
    const int64_t nb = cdiv(nwords, SCAN_THREADS * 8);
    lb_reserve(c, nb);
    PNM_LAUNCH(c, "k_bitmap_list", k_bitmap_list<8>, dim3((unsigned)nb), dim3(SCAN_THREADS), 0,
               map, nwords, nwx, nx, out, total_dev, c->lb_ctl, c->lb_state, (unsigned)nb,
               c->err_dev);
  }
}

void map_rank(Ctx* c, const uint32_t* map, int64_t words, uint32_t* chunk_rank,
              uint64_t* total_dev) {
  if (words <= 0) {
    if (total_dev) fill_u64(c, total_dev, 0, 1);
    return;
  }
  const int64_t nb = cdiv(words, (int64_t)SCAN_THREADS * MAP_CHUNK_WORDS * MAP_RANK_CPT);
  lb_reserve(c, nb);
  PNM_LAUNCH(c, "k_map_rank", k_map_rank, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, map, words,
             chunk_rank, total_dev, c->lb_ctl, c->lb_state, (unsigned)nb, c->err_dev);
}

void scan_exclusive_u32(Ctx* c, const uint32_t* in, uint32_t* out, int64_t n,
                        uint64_t* total_dev) {
  if (n <= 0) {
    if (total_dev) fill_u64(c, total_dev, 0, 1);
    return;
  }
  if (n <= SCAN_TILE) {
    PNM_LAUNCH(c, "k_scan_single", k_scan_single, dim3(1), dim3(SCAN_THREADS), 0, in, n, out,
               total_dev);
    return;
  }
  if (n > LB_LONG) {
    const int64_t nb = cdiv(n, SCAN_THREADS * 32);
    lb_reserve(c, nb);
    PNM_LAUNCH(c, "k_scan_lookback", k_scan_lookback<32>, dim3((unsigned)nb),
               dim3(SCAN_THREADS), 0, in, n, out, total_dev, c->lb_ctl, c->lb_state,
               (unsigned)nb, c->err_dev);
  } else {
    const int64_t nb = cdiv(n, SCAN_THREADS * 8);
    lb_reserve(c, nb);
    PNM_LAUNCH(c, "k_scan_lookback", k_scan_lookback<8>, dim3((unsigned)nb), dim3(SCAN_THREADS),
               0, in, n, out, total_dev, c->lb_ctl, c->lb_state, (unsigned)nb, c->err_dev);
  }
}

// ---------------- stable LSD radix sort ----------------
constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int RS_WAVES = RS_THREADS / 64;

// DB-bit digits (8 or 9): NB = 2^DB buckets, each thread owning NB / RS_THREADS of them.
template <class K, int DB>
__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const K* __restrict__ keys, int64_t n,
                                                        int shift, uint32_t* __restrict__ gh,
                                                        unsigned nblocks) {
  constexpr int NB = 1 << DB;
  __shared__ uint32_t h[NB];
  for (int d = threadIdx.x; d < NB; d += RS_THREADS) h[d] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll 4
  for (int r = 0; r < RS_ITEMS; ++r) {
    const int64_t i = base + r * RS_THREADS + threadIdx.x;
    if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & (NB - 1u)], 1u);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < NB; d += RS_THREADS)
    gh[(int64_t)d * nblocks + blockIdx.x] = h[d];
}

// Each block re-reads its tile in 16 rounds of 256 keys (in input order). Within a round, a
// wave finds the lanes sharing its digit with DB + 1 ballots (the extra bit marks keys past the
// end); per-wave digit counts are prefixed across the 4 waves in wave order, so the scatter is
// stable.
template <class K, bool VALS, int DB>
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(
    const K* __restrict__ keys, const uint32_t* __restrict__ vals, K* __restrict__ okeys,
    uint32_t* __restrict__ ovals, int64_t n, int shift, const uint32_t* __restrict__ ghs,
    unsigned nblocks) {
  constexpr int NB = 1 << DB;
  // per-wave digit counts and bases, double-buffered by round parity: a round's prefix zeroes
  // the other parity's counts (last read by the previous round's prefix, before this round's
  // first barrier), so each round takes two barriers
  __shared__ uint32_t gbase[NB];
  __shared__ uint32_t cnt[2][RS_WAVES][NB];
  __shared__ uint32_t wbase[2][RS_WAVES][NB];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  for (int d = t; d < NB; d += RS_THREADS) {
    gbase[d] = ghs[(int64_t)d * nblocks + blockIdx.x];
#pragma unroll
    for (int q = 0; q < RS_WAVES; ++q) cnt[0][q][d] = 0;
  }
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
  for (int r = 0; r < RS_ITEMS; ++r) {
    const int p = r & 1;
    const int64_t i = base + r * RS_THREADS + t;
    const bool valid = i < n;
    const K k = valid ? keys[i] : (K)0;
    const uint32_t d = valid ? ((uint32_t)(k >> shift) & (NB - 1u)) : (uint32_t)NB;
    uint64_t peers = ~0ull;
#pragma unroll
    for (int b = 0; b <= DB; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const uint32_t rank = lanes_below(peers);
    if (valid && rank == 0) cnt[p][w][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    for (int dd = t; dd < NB; dd += RS_THREADS) {
      uint32_t s = gbase[dd];
#pragma unroll
      for (int q = 0; q < RS_WAVES; ++q) {
        wbase[p][q][dd] = s;
        s += cnt[p][q][dd];
        cnt[p ^ 1][q][dd] = 0;
      }
      gbase[dd] = s;
    }
    __syncthreads();
    if (valid) {
      const uint32_t pos = wbase[p][w][d] + rank;
      okeys[pos] = k;
      if (VALS) ovals[pos] = vals[i];
    }
  }
}

template <class K, int DB>
static void radix_pass(Ctx* c, const K* ks, const uint32_t* vs, K* kd, uint32_t* vd, int64_t n,
                       int shift, uint32_t* gh, unsigned nb) {
  PNM_LAUNCH(c, "k_rs_hist", (k_rs_hist<K, DB>), dim3(nb), dim3(RS_THREADS), 0, ks, n, shift, gh,
             nb);
  scan_exclusive_u32(c, gh, gh, ((int64_t)1 << DB) * nb);
  if (vs)
    PNM_LAUNCH(c, "k_rs_scatter", (k_rs_scatter<K, true, DB>), dim3(nb), dim3(RS_THREADS), 0, ks,
               vs, kd, vd, n, shift, gh, nb);
  else
    PNM_LAUNCH(c, "k_rs_scatter", (k_rs_scatter<K, false, DB>), dim3(nb), dim3(RS_THREADS), 0,
               ks, vs, kd, vd, n, shift, gh, nb);
}

// LSD radix sort over bits [begin_bit, end_bit): 9-bit digits when that takes fewer passes than
// 8-bit ones (a 34-bit key: 4 passes instead of 5).
template <class K>
static void radix_sort_impl(Ctx* c, K* keys, uint32_t* vals, int64_t n, int begin_bit,
                            int end_bit) {
  if (n <= 1 || end_bit <= begin_bit) return;
  const int nbits = end_bit - begin_bit;
  const int db = (nbits + 8) / 9 < (nbits + 7) / 8 ? 9 : 8;
  const unsigned nb = (unsigned)((n + RS_TILE - 1) / RS_TILE);
  K* kalt = c->arena.get_n<K>((size_t)n);
  uint32_t* valt = vals ? c->arena.get_n<uint32_t>((size_t)n) : nullptr;
  uint32_t* gh = c->arena.get_n<uint32_t>(((size_t)1 << db) * nb);
  K* ks = keys;
  K* kd = kalt;
  uint32_t* vs = vals;
  uint32_t* vd = valt;
  int passes = 0;
  for (int shift = begin_bit; shift < end_bit; shift += db) {
    if (db == 9)
      radix_pass<K, 9>(c, ks, vs, kd, vd, n, shift, gh, nb);
    else
      radix_pass<K, 8>(c, ks, vs, kd, vd, n, shift, gh, nb);
    std::swap(ks, kd);
    std::swap(vs, vd);
    ++passes;
  }
  if (passes & 1) {
    PNM_HIP(hipMemcpyAsync(keys, ks, sizeof(K) * n, hipMemcpyDeviceToDevice, c->stream));
    if (vals)
      PNM_HIP(hipMemcpyAsync(vals, vs, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, c->stream));
  }
}

void radix_sort_u64(Ctx* c, uint64_t* keys, uint32_t* vals, int64_t n, int begin_bit,
                    int end_bit) {
  radix_sort_impl<uint64_t>(c, keys, vals, n, begin_bit, end_bit);
}
void radix_sort_u32(Ctx* c, uint32_t* keys, uint32_t* vals, int64_t n, int begin_bit,
                    int end_bit) {
  radix_sort_impl<uint32_t>(c, keys, vals, n, begin_bit, end_bit);
}

// ---------------- block-staged compaction ----------------
// Producer kernels write the k-th item of segment b to stage[b * cap + k] and the item count to
// counts[b]; this gathers them in segment order (deterministic, no contended atomics). Several
// lists with the same segment layout go in one pass: their counts are contiguous (list k at
// counts + k * nseg), one exclusive scan covers all of them, one gather launch moves every
// segment to its list (offsets relative to the list's first segment).
//
// One wave per GATHER_SEGS consecutive segments of one list: the lanes read the segments' counts
// and offsets together, then the wave walks the segments' 64-item chunks in batches of
// GATHER_BATCH, issuing every load of a batch (clamped, in-bounds addresses) before its stores,
// so a batch costs one memory round trip. The wave holding a list's last segment writes the
// list's total.
constexpr int GATHER_SEGS = 16;
constexpr int GATHER_BATCH = 8;

template <class T>
struct GatherLists {
  const T* stage[4];
  T* out[4];
  uint64_t* total[4];
};

template <class T>
__global__ __launch_bounds__(256) void k_gather_lists(GatherLists<T> L,
                                                      const uint32_t* __restrict__ counts,
                                                      const uint32_t* __restrict__ offs,
                                                      int64_t nseg, int64_t cap,
                                                      int64_t waves_per_list, int K) {
  const int lane = threadIdx.x & 63;
  const int64_t wv = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int k = (int)(wv / waves_per_list);
  if (k >= K) return;  // wave-uniform
  const int64_t s0 = (wv - (int64_t)k * waves_per_list) * GATHER_SEGS;  // first segment in list
  const int64_t cb = (int64_t)k * nseg;
  const int nq = (int)(nseg - s0 < GATHER_SEGS ? nseg - s0 : GATHER_SEGS);
  const T* __restrict__ st = L.stage[k];
  T* __restrict__ o = L.out[k];
  uint32_t c = 0, of = 0;
  if (lane < nq) {
    c = counts[cb + s0 + lane];
    of = offs[cb + s0 + lane] - offs[cb];
  }
  if (s0 + nq == nseg) {  // this wave holds the list's last segment
    const uint32_t cl = (uint32_t)__builtin_amdgcn_readlane((int)c, nq - 1);
    const uint32_t ol = (uint32_t)__builtin_amdgcn_readlane((int)of, nq - 1);
    if (lane == 0) *L.total[k] = (uint64_t)ol + cl;
  }
  int q = 0;
  uint32_t j = 0;  // next item of segment q (wave-uniform)
  while (true) {
    T v[GATHER_BATCH];
    uint32_t dst[GATHER_BATCH];
    bool ok[GATHER_BATCH];
#pragma unroll
    for (int bt = 0; bt < GATHER_BATCH; ++bt) {
      while (q < nq && j >= (uint32_t)__builtin_amdgcn_readlane((int)c, q)) {
        ++q;
        j = 0;
      }
      const int qq = q < nq ? q : nq - 1;
      const uint32_t cq = q < nq ? (uint32_t)__builtin_amdgcn_readlane((int)c, qq) : 0u;
      const uint32_t oq = (uint32_t)__builtin_amdgcn_readlane((int)of, qq);
      const uint32_t idx = j + (uint32_t)lane;
      ok[bt] = idx < cq;
      dst[bt] = oq + idx;
      v[bt] = st[(s0 + qq) * cap + (idx < (uint32_t)cap ? idx : (uint32_t)cap - 1u)];
      j += 64u;
    }
#pragma unroll
    for (int bt = 0; bt < GATHER_BATCH; ++bt)
      if (ok[bt]) o[dst[bt]] = v[bt];
    if (q >= nq) break;
  }
}

template <class T>
static void gather_lists(Ctx* c, int K, const T* const* stage, const uint32_t* counts,
                         int64_t nseg, int64_t cap, T* const* out, uint64_t* const* total_dev,
                         bool scan_total) {
  if (cap < 64) fail(PNM_EINVAL, "compaction: segment capacity below 64");
  GatherLists<T> L{};
  for (int k = 0; k < K; ++k) {
    L.stage[k] = stage[k];
    L.out[k] = out[k];
    L.total[k] = total_dev[k];
  }
  if (nseg <= 0) {
    for (int k = 0; k < K; ++k) fill_u64(c, total_dev[k], 0, 1);
    return;
  }
  uint32_t* offs = c->arena.get_n<uint32_t>((size_t)(K * nseg));
  scan_exclusive_u32(c, counts, offs, K * nseg, scan_total ? total_dev[0] : nullptr);
  const int64_t wpl = (nseg + GATHER_SEGS - 1) / GATHER_SEGS;
  PNM_LAUNCH(c, "k_gather_staged", k_gather_lists<T>, dim3(cdiv(K * wpl, 4)), dim3(256), 0, L,
             counts, offs, nseg, cap, wpl, K);
}

void compact_staged(Ctx* c, const uint32_t* stage, const uint32_t* counts, int64_t nblocks,
                    int64_t cap, uint32_t* out, uint64_t* total_dev) {
  gather_lists<uint32_t>(c, 1, &stage, counts, nblocks, cap, &out, &total_dev, false);
}

void compact_staged_multi(Ctx* c, int K, const uint32_t* const* stage, const uint32_t* counts,
                          int64_t nseg, int64_t cap, uint32_t* const* out,
                          uint64_t* const* total_dev) {
  gather_lists<uint32_t>(c, K, stage, counts, nseg, cap, out, total_dev, false);
}

void compact_staged64(Ctx* c, const uint64_t* stage, const uint32_t* counts, int64_t nblocks,
                      int64_t cap, uint64_t* out, uint64_t* total_dev) {
  gather_lists<uint64_t>(c, 1, &stage, counts, nblocks, cap, &out, &total_dev, false);
}

// ---------------- unique of a sorted list ----------------
// One launch: each thread takes UQ_ITEMS consecutive keys and counts those that differ from
// their predecessor, the look-back scan gives its output position, and it writes them in
// order; the last block writes the count.
constexpr int UQ_ITEMS = 8;
__global__ __launch_bounds__(SCAN_THREADS) void k_unique(const uint64_t* __restrict__ k, int64_t n,
                                                        uint64_t* __restrict__ out,
                                                        uint64_t* __restrict__ total_out,
                                                        uint32_t* __restrict__ ctl,
                                                        uint64_t* __restrict__ state,
                                                        unsigned nblocks,
                                                        uint32_t* __restrict__ err) {
  __shared__ uint32_t lds[SCAN_THREADS / 64];
  __shared__ uint32_t s_bid, s_epoch, s_excl;
  uint32_t bid, epoch;
  lb_begin(ctl, bid, epoch, &s_bid, &s_epoch);
  const int64_t base = ((int64_t)bid * SCAN_THREADS + threadIdx.x) * UQ_ITEMS;
  uint64_t v[UQ_ITEMS];
  load_run<UQ_ITEMS>(k, base, n, v);
  const uint64_t before = (base > 0 && base <= n) ? k[base - 1] : 0ull;
  uint32_t f = 0;  // bit q: item q starts a new key
  uint32_t s = 0;
#pragma unroll
  for (int q = 0; q < UQ_ITEMS; ++q) {
    const int64_t i = base + q;
    const bool nw = i < n && (i == 0 || v[q] != (q == 0 ? before : v[q - 1]));
    f |= (uint32_t)nw << q;
    s += nw;
  }
  uint32_t tot;
  const uint32_t pre = block_excl_scan(s, lds, &tot);
  if (threadIdx.x < 64) {
    const uint32_t excl = lb_lookback(state, bid, epoch, tot, err);
    if (threadIdx.x == 0) s_excl = excl;
  }
  __syncthreads();
  uint32_t r = s_excl + pre;
#pragma unroll
  for (int q = 0; q < UQ_ITEMS; ++q)
    if ((f >> q) & 1u) out[r++] = v[q];
  if (threadIdx.x == 0) {
    if (bid == nblocks - 1) *total_out = (uint64_t)s_excl + tot;
    lb_end(ctl, epoch, nblocks);
  }
}

void unique_sorted_u64_dev(Ctx* c, const uint64_t* keys, uint64_t* out, int64_t n,
                           uint64_t* count_dev) {
  if (n <= 0) {
    fill_u64(c, count_dev, 0, 1);
    return;
  }
  const int64_t nb = cdiv(n, SCAN_THREADS * UQ_ITEMS);
  lb_reserve(c, nb);
  PNM_LAUNCH(c, "k_unique", k_unique, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, keys, n, out,
             count_dev, c->lb_ctl, c->lb_state, (unsigned)nb, c->err_dev);
}


}  // namespace pnm
