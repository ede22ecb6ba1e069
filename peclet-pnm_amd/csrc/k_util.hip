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

void fill_u32(Ctx* c, uint32_t* p, uint32_t v, int64_t count) {
  if (count <= 0) return;
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

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(const uint32_t* __restrict__ in,
                                                              int64_t n,
                                                              uint32_t* __restrict__ sums) {
  __shared__ uint32_t lds[SCAN_THREADS / 64];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k)
    if (base + k < n) s += in[base + k];
  uint32_t tot;
  block_excl_scan(s, lds, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_down(const uint32_t* in, int64_t n,
                                                            const uint32_t* __restrict__ offs,
                                                            uint32_t* out) {
  __shared__ uint32_t lds[SCAN_THREADS / 64];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
  uint32_t vals[SCAN_ITEMS];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    vals[k] = (base + k < n) ? in[base + k] : 0u;
    s += vals[k];
  }
  uint32_t tot;
  uint32_t run = block_excl_scan(s, lds, &tot) + (offs ? offs[blockIdx.x] : 0u);
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    if (base + k < n) out[base + k] = run;
    run += vals[k];
  }
}

void scan_exclusive_u32(Ctx* c, const uint32_t* in, uint32_t* out, int64_t n) {
  if (n <= 0) return;
  const int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (nb == 1) {
    PNM_LAUNCH(c, "k_scan_down", k_scan_down, dim3(1), dim3(SCAN_THREADS), 0, in, n,
                       (const uint32_t*)nullptr, out);
    return;
  }
  uint32_t* sums = c->arena.get_n<uint32_t>((size_t)nb);
  PNM_LAUNCH(c, "k_scan_reduce", k_scan_reduce, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, in, n,
                     sums);
  scan_exclusive_u32(c, sums, sums, nb);
  PNM_LAUNCH(c, "k_scan_down", k_scan_down, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, in, n,
                     (const uint32_t*)sums, out);
}

// ---------------- stable LSD radix sort ----------------
constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int RS_WAVES = RS_THREADS / 64;

template <class K>
__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const K* __restrict__ keys, int64_t n,
                                                        int shift, uint32_t* __restrict__ gh,
                                                        unsigned nblocks) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll 4
  for (int r = 0; r < RS_ITEMS; ++r) {
    const int64_t i = base + r * RS_THREADS + threadIdx.x;
    if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & 0xFFu], 1u);
  }
  __syncthreads();
  gh[(int64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Each block re-reads its tile in 16 rounds of 256 keys (in input order). Within a round, a
// wave finds the lanes sharing its digit with 9 ballots; per-wave digit counts are prefixed
// across the 4 waves in wave order, so the scatter is stable.
template <class K, bool VALS>
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(
    const K* __restrict__ keys, const uint32_t* __restrict__ vals, K* __restrict__ okeys,
    uint32_t* __restrict__ ovals, int64_t n, int shift, const uint32_t* __restrict__ ghs,
    unsigned nblocks) {
  __shared__ uint32_t gbase[256];
  __shared__ uint32_t cnt[RS_WAVES][256];
  __shared__ uint32_t wbase[RS_WAVES][256];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  gbase[t] = ghs[(int64_t)t * nblocks + blockIdx.x];
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
  for (int r = 0; r < RS_ITEMS; ++r) {
    const int64_t i = base + r * RS_THREADS + t;
    const bool valid = i < n;
    const K k = valid ? keys[i] : (K)0;
    const uint32_t d = valid ? ((uint32_t)(k >> shift) & 0xFFu) : 256u;
    uint64_t peers = ~0ull;
#pragma unroll
    for (int b = 0; b < 9; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & below);
#pragma unroll
    for (int q = 0; q < RS_WAVES; ++q) cnt[q][t] = 0;
    __syncthreads();
    if (valid && rank == 0) cnt[w][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    {
      uint32_t s = gbase[t];
#pragma unroll
      for (int q = 0; q < RS_WAVES; ++q) {
        wbase[q][t] = s;
        s += cnt[q][t];
      }
      gbase[t] = s;
    }
    __syncthreads();
    if (valid) {
      const uint32_t pos = wbase[w][d] + rank;
      okeys[pos] = k;
      if (VALS) ovals[pos] = vals[i];
    }
  }
}

template <class K>
static void radix_sort_impl(Ctx* c, K* keys, uint32_t* vals, int64_t n, int begin_bit,
                            int end_bit) {
  if (n <= 1 || end_bit <= begin_bit) return;
  const unsigned nb = (unsigned)((n + RS_TILE - 1) / RS_TILE);
  K* kalt = c->arena.get_n<K>((size_t)n);
  uint32_t* valt = vals ? c->arena.get_n<uint32_t>((size_t)n) : nullptr;
  uint32_t* gh = c->arena.get_n<uint32_t>((size_t)256 * nb);
  K* ks = keys;
  K* kd = kalt;
  uint32_t* vs = vals;
  uint32_t* vd = valt;
  int passes = 0;
  for (int shift = begin_bit; shift < end_bit; shift += 8) {
    PNM_LAUNCH(c, "k_rs_hist", k_rs_hist<K>, dim3(nb), dim3(RS_THREADS), 0, ks, n, shift, gh,
                       nb);
    scan_exclusive_u32(c, gh, gh, (int64_t)256 * nb);
    if (vals)
      PNM_LAUNCH(c, "k_rs_scatter", (k_rs_scatter<K, true>), dim3(nb), dim3(RS_THREADS), 0, ks,
                         vs, kd, vd, n, shift, gh, nb);
    else
      PNM_LAUNCH(c, "k_rs_scatter", (k_rs_scatter<K, false>), dim3(nb), dim3(RS_THREADS), 0, ks, vs, kd, vd, n, shift, gh, nb);
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
// Producer kernels write the k-th item of block b to stage[b * cap + k] and the item count to
// counts[b]; this gathers them in block order (deterministic, no contended atomics).
__global__ void k_gather_staged(const uint32_t* __restrict__ stage,
                                const uint32_t* __restrict__ counts,
                                const uint32_t* __restrict__ offs, int64_t cap,
                                uint32_t* __restrict__ out) {
  const int64_t b = blockIdx.x;
  const uint32_t cnt = counts[b];
  const uint32_t base = offs[b];
  for (uint32_t k = threadIdx.x; k < cnt; k += blockDim.x) out[base + k] = stage[b * cap + k];
}

__global__ void k_gather_staged64(const uint64_t* __restrict__ stage,
                                  const uint32_t* __restrict__ counts,
                                  const uint32_t* __restrict__ offs, int64_t cap,
                                  uint64_t* __restrict__ out) {
  const int64_t b = blockIdx.x;
  const uint32_t cnt = counts[b];
  const uint32_t base = offs[b];
  for (uint32_t k = threadIdx.x; k < cnt; k += blockDim.x) out[base + k] = stage[b * cap + k];
}

__global__ void k_staged_total(const uint32_t* __restrict__ counts,
                               const uint32_t* __restrict__ offs, int64_t nblocks,
                               uint64_t* __restrict__ total) {
  *total = (uint64_t)offs[nblocks - 1] + counts[nblocks - 1];
}

void compact_staged(Ctx* c, const uint32_t* stage, const uint32_t* counts, int64_t nblocks,
                    int64_t cap, uint32_t* out, uint64_t* total_dev) {
  if (nblocks <= 0) {
    fill_u64(c, total_dev, 0, 1);
    return;
  }
  uint32_t* offs = c->arena.get_n<uint32_t>((size_t)nblocks);
  scan_exclusive_u32(c, counts, offs, nblocks);
  PNM_LAUNCH(c, "k_gather_staged", k_gather_staged, dim3((unsigned)nblocks), dim3(64), 0, stage,
             counts, offs, cap, out);
  PNM_LAUNCH(c, "k_staged_total", k_staged_total, dim3(1), dim3(1), 0, counts, offs, nblocks,
             total_dev);
}

// Several staged lists with the same segment layout in one pass: their counts are contiguous
// (list k at counts + k * nseg), one exclusive scan covers all of them, one gather launch moves
// every segment to its list (offsets relative to the list's first segment).
struct MultiLists {
  const uint32_t* stage[4];
  uint32_t* out[4];
  uint64_t* total[4];
};
__global__ void k_gather_multi(MultiLists L, const uint32_t* __restrict__ counts,
                               const uint32_t* __restrict__ offs, int64_t nseg, int64_t cap) {
  const int64_t b = blockIdx.x;
  const int k = (int)(b / nseg);
  const int64_t sgm = b - (int64_t)k * nseg;
  const uint32_t cnt = counts[b];
  const uint32_t base = offs[b] - offs[(int64_t)k * nseg];
  const uint32_t* __restrict__ st = L.stage[k] + sgm * cap;
  uint32_t* __restrict__ o = L.out[k];
  for (uint32_t j = threadIdx.x; j < cnt; j += blockDim.x) o[base + j] = st[j];
}
__global__ void k_multi_totals(MultiLists L, const uint32_t* __restrict__ counts,
                               const uint32_t* __restrict__ offs, int64_t nseg, int K) {
  const int k = threadIdx.x;
  if (k >= K) return;
  const int64_t last = (int64_t)(k + 1) * nseg - 1;
  *L.total[k] = (uint64_t)(offs[last] + counts[last] - offs[(int64_t)k * nseg]);
}

void compact_staged_multi(Ctx* c, int K, const uint32_t* const* stage, const uint32_t* counts,
                          int64_t nseg, int64_t cap, uint32_t* const* out,
                          uint64_t* const* total_dev) {
  MultiLists L{};
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
  scan_exclusive_u32(c, counts, offs, K * nseg);
  PNM_LAUNCH(c, "k_gather_staged", k_gather_multi, dim3((unsigned)(K * nseg)), dim3(64), 0, L,
             counts, offs, nseg, cap);
  PNM_LAUNCH(c, "k_staged_total", k_multi_totals, dim3(1), dim3(64), 0, L, counts, offs, nseg, K);
}

void compact_staged64(Ctx* c, const uint64_t* stage, const uint32_t* counts, int64_t nblocks,
                      int64_t cap, uint64_t* out, uint64_t* total_dev) {
  if (nblocks <= 0) {
    fill_u64(c, total_dev, 0, 1);
    return;
  }
  uint32_t* offs = c->arena.get_n<uint32_t>((size_t)nblocks);
  scan_exclusive_u32(c, counts, offs, nblocks);
  PNM_LAUNCH(c, "k_gather_staged", k_gather_staged64, dim3((unsigned)nblocks), dim3(64), 0,
             stage, counts, offs, cap, out);
  PNM_LAUNCH(c, "k_staged_total", k_staged_total, dim3(1), dim3(1), 0, counts, offs, nblocks,
             total_dev);
}

// ---------------- unique of a sorted list ----------------
__global__ void k_unique_flags(const uint64_t* __restrict__ k, int64_t n,
                               uint32_t* __restrict__ f) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = (i == 0 || k[i] != k[i - 1]) ? 1u : 0u;
}
__global__ void k_unique_write(const uint64_t* __restrict__ k, int64_t n,
                               const uint32_t* __restrict__ f, const uint32_t* __restrict__ pos,
                               uint64_t* __restrict__ out, uint64_t* __restrict__ count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && f[i]) out[pos[i]] = k[i];
  if (i == n - 1) *count = (uint64_t)pos[i] + f[i];
}

int64_t unique_sorted_u64(Ctx* c, const uint64_t* keys, uint64_t* out, int64_t n) {
  if (n <= 0) return 0;
  uint32_t* f = c->arena.get_n<uint32_t>((size_t)n);
  uint32_t* pos = c->arena.get_n<uint32_t>((size_t)n);
  uint64_t* cnt = c->arena.get_n<uint64_t>(1);
  const unsigned nb = (unsigned)((n + 255) / 256);
  PNM_LAUNCH(c, "k_unique_flags", k_unique_flags, dim3(nb), dim3(256), 0, keys, n, f);
  scan_exclusive_u32(c, f, pos, n);
  PNM_LAUNCH(c, "k_unique_write", k_unique_write, dim3(nb), dim3(256), 0, keys, n, f, pos, out,
                     cnt);
  uint64_t h;
  readback(c, cnt, 1, &h);
  return (int64_t)h;
}

}  // namespace pnm
