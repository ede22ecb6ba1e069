// k_batch.hip — batched segment_volume: thousands of independent periodic volumes of uneven
// size segmented by ONE persistent kernel launch (replayed from a hipGraph).
//
// Each workgroup pulls volumes from an atomic work queue (volumes pre-ordered largest first,
// so the tail of the schedule holds the small ones) and runs the whole watershed of a volume
// with workgroup barriers between phases — no grid-wide synchronisation exists:
//   P1 one pass over the SDF: ascent successor, x-run marker labels, flood band, sdf>0 peaks
//   P2 unions between touching x-runs of forward rows           (merge_markers :198-224)
//   P3 flatten; component roots -> solid keys                     (flatten :183-194)
//   P4 Jacobi flood sweeps over the shrinking band               (flood :237-269)
//   P5 capped root chase + first-appearance minima               (:273-311, :331-337)
//   P6 key ranking by a single-wave stable radix sort            (:340-374)
//   P7 seg = id[key]                                             (:378-381)
// Every volume is its own periodic box and is numbered on its own, exactly as a separate
// segment_volume call (pore_extraction.hpp:389-397) would number it.
//
// The segmentation of a volume is written into its seg buffer, which doubles as the working
// label array. Per-workgroup scratch (successor/ids, minima, lists) is sized for the largest
// volume of the batch.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "k_common.hpp"

namespace pnm {

struct BatchVol {
  const float* sdf;
  int32_t* seg;
  int nx, ny, nz;
  int pad;
};

struct BatchScratch {
  int32_t* nxt;     // ascent successor, later the id table
  int32_t* mn;      // first-appearance minima at key positions
  uint32_t* band;   // flood band (ping)
  uint32_t* band2;  // flood band (pong)
  int32_t* newlab;  // per-band-entry new label
  uint32_t* pkeys;  // pore keys (sdf>0 peaks, or all roots after a capped walk)
  uint32_t* skeys;  // solid keys (component roots)
  uint32_t* tmp;    // sort ping-pong values
  int64_t stride;   // elements per workgroup slice
};

constexpr int kBatchCounters = 8;

// Single-wave stable LSD radix sort of (key, val) pairs (keys < 2^bits), ping-pong through
// (k2, v2). Called by every thread of the block (barriers), work done by wave 0.
__device__ void block_wave_sort(uint32_t* k, uint32_t* v, uint32_t* k2, uint32_t* v2,
                                unsigned K, int bits, unsigned* hist, int wave, int lane) {
  bool flipped = false;
  for (int shift = 0; shift < bits; shift += 8) {
    if (wave == 0)
      for (int b = lane; b < 256; b += 64) hist[b] = 0;
    __syncthreads();
    if (wave == 0)
      for (unsigned j = lane; j < K; j += 64) atomicAdd(&hist[(k[j] >> shift) & 0xFFu], 1u);
    __syncthreads();
    if (wave == 0) {  // exclusive scan of 256 bins: 4 per lane
      unsigned h0 = hist[4 * lane], h1 = hist[4 * lane + 1], h2 = hist[4 * lane + 2],
               h3 = hist[4 * lane + 3];
      unsigned s = h0 + h1 + h2 + h3;
      unsigned inc = s;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const unsigned t = __shfl_up(inc, o, 64);
        if (lane >= o) inc += t;
      }
      const unsigned ex = inc - s;
      hist[4 * lane] = ex;
      hist[4 * lane + 1] = ex + h0;
      hist[4 * lane + 2] = ex + h0 + h1;
      hist[4 * lane + 3] = ex + h0 + h1 + h2;
    }
    __syncthreads();
    if (wave == 0) {
      for (unsigned r0 = 0; r0 < K; r0 += 64) {
        const unsigned j = r0 + lane;
        const bool valid = j < K;
        const uint32_t key = valid ? k[j] : 0u;
        const uint32_t d = valid ? ((key >> shift) & 0xFFu) : 256u;
        uint64_t peers = ~0ull;
#pragma unroll
        for (int b = 0; b < 9; ++b) {
          const bool bit = (d >> b) & 1u;
          const uint64_t m = __ballot(bit);
          peers &= bit ? m : ~m;
        }
        const unsigned rank = lanes_below(peers);
        const unsigned base = valid ? hist[d] : 0u;
        if (valid) {
          k2[base + rank] = key;
          v2[base + rank] = v[j];
        }
        __builtin_amdgcn_wave_barrier();
        if (valid && rank == 0) hist[d] = base + (unsigned)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
      }
    }
    __syncthreads();
    uint32_t* t;
    t = k; k = k2; k2 = t;
    t = v; v = v2; v2 = t;
    flipped = !flipped;
  }
  if (flipped) {  // after an odd number of passes the result sits in the temporaries: copy back
    for (unsigned j = threadIdx.x; j < K; j += blockDim.x) {
      k2[j] = k[j];
      v2[j] = v[j];
    }
    __syncthreads();
  }
}

// q = a / d, r = a % d for 0 <= a, 0 < d: a float quotient estimate, then exact corrections.
__device__ __forceinline__ void fdivmod(int a, int d, float inv_d, int& q, int& r) {
  q = (int)((float)a * inv_d);
  r = a - q * d;
  while (r < 0) {
    --q;
    r += d;
  }
  while (r >= d) {
    ++q;
    r -= d;
  }
}

// runs_union_chunk (k_common.hpp) with the same union pairs, every partner label loaded before
// the first union: the forward rows at x, and at a run's first / last column their x-1 / x+1
// (used only where the forward row has no marker at x). The unions' finds then overlap the
// loads of the next pairs instead of waiting behind them. A label read before another union
// rewrote it is an ancestor of what it would read later, so the components are the same.
__device__ __forceinline__ void runs_union_chunk_hoist(int* L, int nx, int ny, int nz, int x,
                                                       int y, int z, int lane) {
  const bool in = (x < nx) && (y < ny);
  const int i = in ? (z * ny + y) * nx + x : 0;
  const int my = in ? L[i] : -1;
  const bool mk = my != -1;
  const uint64_t m = __ballot(mk);
  if (!__any(mk)) return;  // wave-uniform
  const bool run_start = mk && (lane == 0 || !((m >> (lane - 1)) & 1ull));
  const bool run_end = mk && (lane == 63 || !((m >> (lane + 1)) & 1ull));
  const int xm = in ? wm1(x, nx) : 0, xp = in ? wp1(x, nx) : 0;
  const int yp = in ? wp1(y, ny) : 0, ym = in ? wm1(y, ny) : 0, zp = wp1(z, nz);
  const int rows[4] = {(z * ny + yp) * nx, (zp * ny + ym) * nx, (zp * ny + y) * nx,
                       (zp * ny + yp) * nx};
  int c[4], pc[4], ca[4], cb[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) c[r] = in ? L[rows[r] + x] : -1;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    ca[r] = run_start ? L[rows[r] + xm] : -1;
    cb[r] = run_end ? L[rows[r] + xp] : -1;
  }
  // +x relation across the chunk start; at x = 0 this is the periodic wrap from x = nx-1
  const int nl = (mk && lane == 0) ? L[(z * ny + y) * nx + xm] : -1;
#pragma unroll
  for (int r = 0; r < 4; ++r) pc[r] = __shfl_up(c[r], 1, 64);
  if (nl != -1) uf_unite(L, my, nl);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    if (mk && c[r] != -1) {
      // one union per touching pair: skipped when column x-1 gave the same pair
      const bool dup = lane > 0 && ((m >> (lane - 1)) & 1ull) && pc[r] == c[r];
      if (!dup) uf_unite(L, my, c[r]);
    }
    if (mk && c[r] == -1) {  // diagonal-only contacts at the run's ends
      if (ca[r] != -1) uf_unite(L, my, ca[r]);
      if (cb[r] != -1) uf_unite(L, my, cb[r]);
    }
  }
}

template <int BS>
__global__ __launch_bounds__(BS) void k_seg_batch(const BatchVol* __restrict__ vols,
                                                  const uint32_t* __restrict__ order, int nvol,
                                                  unsigned* queue, BatchScratch S, float thr,
                                                  int32_t* __restrict__ out_np,
                                                  int32_t* __restrict__ out_ns) {
  constexpr int NW = BS / 64;
  __shared__ unsigned ctr[kBatchCounters];
  __shared__ unsigned hist[256];
  __shared__ int s_q;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t off = (int64_t)blockIdx.x * S.stride;
  int32_t* nxt = S.nxt + off;
  int32_t* mn = S.mn + off;
  uint32_t* band = S.band + off;
  uint32_t* band2 = S.band2 + off;
  int32_t* newlab = S.newlab + off;
  uint32_t* pkeys = S.pkeys + off;
  uint32_t* skeys = S.skeys + off;
  uint32_t* tmp = S.tmp + off;

  for (;;) {
    if (tid == 0) s_q = (int)atomicAdd(queue, 1u);
    if (tid < kBatchCounters) ctr[tid] = 0;
    __syncthreads();
    const int q = s_q;
    if (q >= nvol) break;
    const int vid = (int)order[q];
    const BatchVol V = vols[vid];
    const float* __restrict__ sdf = V.sdf;
    int32_t* W = V.seg;
    const int nx = V.nx, ny = V.ny, nz = V.nz;
    const int n = nx * ny * nz;
    const int cx = (nx + 63) >> 6;
    const int nchunks = nz * ny * cx;

    // ---- P1: successor, x-run marker labels, band, sdf>0 peaks ----
    for (int ch = wave; ch < nchunks; ch += NW) {
      const int xc = ch % cx, row = ch / cx;
      const int y = row % ny, z = row / ny;
      const int x = (xc << 6) + lane;
      const bool in = x < nx;
      const int i = in ? (z * ny + y) * nx + x : 0;
      const float cv = in ? sdf[i] : 0.0f;
      int nx_i = i;
      bool pos_peak = false;
      if (in && cv > 0.0f) {
        nx_i = ascent_best(sdf, x, y, z, nx, ny, nz, i, cv);
        pos_peak = nx_i == i;
      }
      const bool mk = in && (cv < thr);
      const uint64_t m = __ballot(mk);
      const uint64_t starts = m & ~(m << 1);
      const uint64_t upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1);
      const uint64_t sm = starts & upto;
      const int s = sm ? 63 - __clzll((long long)sm) : lane;
      if (in) {
        W[i] = mk ? i - (lane - s) : -1;
        nxt[i] = nx_i;
        if (pos_peak) mn[i] = i;  // a basin's first voxel is never after its peak
      }
      const unsigned ps = wave_append_lds(&ctr[0], pos_peak, lane);
      if (pos_peak) pkeys[ps] = (uint32_t)i;
      const bool bnd = in && !(cv >= 0.0f) && !mk;
      const unsigned bs = wave_append_lds(&ctr[1], bnd, lane);
      if (bnd) band[bs] = (uint32_t)i;
    }
    __syncthreads();

    // ---- P2: run unions ----
    for (int ch = wave; ch < nchunks; ch += NW) {
      const int xc = ch % cx, row = ch / cx;
      runs_union_chunk_hoist(W, nx, ny, nz, (xc << 6) + lane, row % ny, row / ny, lane);
    }
    __syncthreads();

    // ---- P3: flatten; roots -> solid keys ----
    // four voxels per thread, their root chases advanced side by side
    for (int base = 0; base < n; base += BS * 4) {
      int l[4], p[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int i = base + c * BS + tid;
        l[c] = i < n ? W[i] : -1;
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) p[c] = l[c] != -1 ? W[l[c]] : -1;
      while (true) {
        bool more = false;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          if (l[c] != -1 && p[c] != l[c]) {
            l[c] = p[c];
            more = true;
          }
        }
        if (!more) break;
#pragma unroll
        for (int c = 0; c < 4; ++c) p[c] = l[c] != -1 ? W[l[c]] : -1;
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int i = base + c * BS + tid;
        bool root = false;
        if (i < n && l[c] != -1) {
          W[i] = l[c];
          root = l[c] == i;
          if (root) mn[i] = i;  // the component's smallest marker; flood may lower it
        }
        const unsigned sidx = wave_append_lds(&ctr[2], root, lane);
        if (root) skeys[sidx] = (uint32_t)i;
      }
    }
    __syncthreads();

    // ---- P4: Jacobi flood over the band ----
    {
      const float inv_nx = 1.0f / (float)nx, inv_ny = 1.0f / (float)ny;
      unsigned cnt = ctr[1];
      uint32_t* cur = band;
      uint32_t* nb2 = band2;
      while (cnt > 0) {
        if (tid == 0) ctr[3] = 0;
        __syncthreads();
        for (unsigned base = 0; base < cnt; base += BS) {
          const unsigned j = base + tid;
          bool keep = false;
          uint32_t i = 0;
          if (j < cnt) {
            i = cur[j];
            int row, x, z, y;
            fdivmod((int)i, nx, inv_nx, row, x);
            fdivmod(row, ny, inv_ny, z, y);
            const int xs[3] = {wm1(x, nx), x, wp1(x, nx)};
            const int ys[3] = {wm1(y, ny), y, wp1(y, ny)};
            const int zs[3] = {wm1(z, nz), z, wp1(z, nz)};
            int best = -1;
            for (int a = 0; a < 3; ++a)
              for (int b = 0; b < 3; ++b) {
                const int r = (zs[a] * ny + ys[b]) * nx;
                for (int c = 0; c < 3; ++c) {
                  if (a == 1 && b == 1 && c == 1) continue;
                  const int nl = W[r + xs[c]];
                  if (nl != -1 && (best == -1 || nl < best)) best = nl;
                }
              }
            newlab[j] = best;
            keep = best == -1;
          }
          const unsigned ks = wave_append_lds(&ctr[3], keep, lane);
          if (keep) nb2[ks] = i;
        }
        __syncthreads();
        for (unsigned j = tid; j < cnt; j += BS) {
          const int b = newlab[j];
          if (b != -1) {
            const int i = (int)cur[j];
            W[i] = b;
            if (i < b) atomicMin(mn + b, i);
          }
        }
        __syncthreads();
        const unsigned kept = ctr[3];
        __syncthreads();
        if (kept == cnt) break;
        uint32_t* t = cur;
        cur = nb2;
        nb2 = t;
        cnt = kept;
      }
    }

    // ---- P5: roots into W (pore voxels), pore first-appearance minima (voxels before their
    // peak only; solid minima were completed by the flood) ----
    // four voxels per thread, their capped chases advanced side by side
    constexpr int kCh = 4;
    for (int base = 0; base < n; base += BS * kCh) {
      int w[kCh], st[kCh];
      bool go[kCh], pore[kCh];
#pragma unroll
      for (int c = 0; c < kCh; ++c) {
        const int i = base + c * BS + tid;
        pore[c] = i < n && sdf[i] > 0.0f;
        go[c] = pore[c];
        w[c] = i;
        st[c] = 0;
      }
      while (true) {
        int t[kCh];
#pragma unroll
        for (int c = 0; c < kCh; ++c) t[c] = go[c] ? nxt[w[c]] : 0;
        bool more = false;
#pragma unroll
        for (int c = 0; c < kCh; ++c) {
          if (!go[c]) continue;
          if (t[c] == w[c]) {
            go[c] = false;
          } else {
            w[c] = t[c];
            if (++st[c] == kMaxSteps) go[c] = false;
          }
          more |= go[c];
        }
        if (!more) break;
      }
#pragma unroll
      for (int c = 0; c < kCh; ++c) {
        const int i = base + c * BS + tid;
        int pk = -1;
        if (pore[c]) {
          if (st[c] == kMaxSteps && nxt[w[c]] != w[c]) ctr[4] = 1;
          W[i] = w[c];
          pk = w[c];
        }
        const int ppk = __shfl_up(pk, 1, 64);
        if (pk >= 0 && i < pk && (lane == 0 || ppk != pk)) atomicMin(mn + pk, i);
      }
    }
    __syncthreads();
    unsigned npk = ctr[0];
    if (ctr[4]) {  // capped walk: rebuild the pore minima and the pore key set by full scans
      for (int i = tid; i < n; i += BS)
        if (sdf[i] > 0.0f) mn[i] = INT_MAX;
      __syncthreads();
      for (int i = tid; i < n; i += BS)
        if (sdf[i] > 0.0f) atomicMin(mn + W[i], i);
      __syncthreads();
      if (tid == 0) ctr[5] = 0;
      __syncthreads();
      for (int base = 0; base < n; base += BS) {
        const int i = base + tid;
        const bool key = i < n && sdf[i] > 0.0f && mn[i] != INT_MAX;
        const unsigned ks = wave_append_lds(&ctr[5], key, lane);
        if (key) pkeys[ks] = (uint32_t)i;
      }
      __syncthreads();
      npk = ctr[5];
    }
    const unsigned nsk = ctr[2];

    // ---- P6: ranking (ids written into nxt at key positions) ----
    int bits = 1;
    while (bits < 32 && (1u << bits) <= (unsigned)n) ++bits;
    // pores: sort (first voxel, key)
    for (unsigned j = tid; j < npk; j += BS) {
      band[j] = (uint32_t)mn[pkeys[j]];
      band2[j] = pkeys[j];
    }
    __syncthreads();
    block_wave_sort(band, band2, reinterpret_cast<uint32_t*>(newlab), tmp, npk, bits, hist, wave,
                    lane);
    for (unsigned j = tid; j < npk; j += BS) nxt[band2[j]] = (int32_t)(j + 1);
    __syncthreads();
    // solids: sort, then count only labels whose first voxel has sdf <= 0
    for (unsigned j = tid; j < nsk; j += BS) {
      band[j] = (uint32_t)mn[skeys[j]];
      band2[j] = skeys[j];
    }
    __syncthreads();
    block_wave_sort(band, band2, reinterpret_cast<uint32_t*>(newlab), tmp, nsk, bits, hist, wave,
                    lane);
    if (wave == 0) {
      unsigned carry = 0;
      for (unsigned r0 = 0; r0 < nsk; r0 += 64) {
        const unsigned j = r0 + lane;
        const bool valid = j < nsk && sdf[band[j]] <= 0.0f;
        const uint64_t m = __ballot(valid);
        const unsigned rank = carry + lanes_below(m);
        if (j < nsk) nxt[band2[j]] = valid ? -(int32_t)(rank + 1) : 0;
        carry += (unsigned)__popcll(m);
      }
      if (lane == 0) {
        ctr[6] = carry;
      }
    }
    __syncthreads();

    // ---- P7: seg = id[key] ----
    for (int i = tid; i < n; i += BS) {
      const int k = W[i];
      W[i] = k == -1 ? 0 : nxt[k];
    }
    if (tid == 0) {
      if (out_np) out_np[vid] = (int32_t)npk;
      if (out_ns) out_ns[vid] = (int32_t)ctr[6];
    }
    __syncthreads();
  }
}

__global__ void k_reset_queue(unsigned* q) { *q = 0u; }

// ------------------------------------------------------------------------------------------

struct BatchConfig {
  int block;       // 256, 512 or 1024
  int groups;      // persistent workgroups (before the memory and volume-count caps)
};

struct BatchPlan {
  BatchConfig cfg;
  int groups;
  int64_t stride;
};

static void launch_batch(Ctx* c, const BatchPlan& P, const BatchVol* vols_d,
                         const uint32_t* order_d, int nvol, unsigned* queue, BatchScratch S,
                         float thr, int32_t* out_np, int32_t* out_ns) {
  PNM_LAUNCH(c, "k_reset_queue", k_reset_queue, dim3(1), dim3(1), 0, queue);
  const dim3 grid((unsigned)P.groups);
  switch (P.cfg.block) {
    case 256:
      PNM_LAUNCH(c, "k_seg_batch", k_seg_batch<256>, grid, dim3(256), 0, vols_d, order_d, nvol,
                 queue, S, thr, out_np, out_ns);
      break;
    case 512:
      PNM_LAUNCH(c, "k_seg_batch", k_seg_batch<512>, grid, dim3(512), 0, vols_d, order_d, nvol,
                 queue, S, thr, out_np, out_ns);
      break;
    default:
      PNM_LAUNCH(c, "k_seg_batch", k_seg_batch<1024>, grid, dim3(1024), 0, vols_d, order_d,
                 nvol, queue, S, thr, out_np, out_ns);
      break;
  }
}

int batch_cu_count(Ctx* c) {
  hipDeviceProp_t p;
  PNM_HIP(hipGetDeviceProperties(&p, c->device));
  return p.multiProcessorCount;
}

// Scratch: 8 arrays of `stride` 4-byte elements per workgroup.
static BatchScratch make_scratch(Ctx* c, int groups, int64_t stride) {
  BatchScratch S{};
  S.stride = stride;
  const size_t each = (size_t)groups * (size_t)stride;
  S.nxt = c->arena.get_n<int32_t>(each);
  S.mn = c->arena.get_n<int32_t>(each);
  S.band = c->arena.get_n<uint32_t>(each);
  S.band2 = c->arena.get_n<uint32_t>(each);
  S.newlab = c->arena.get_n<int32_t>(each);
  S.pkeys = c->arena.get_n<uint32_t>(each);
  S.skeys = c->arena.get_n<uint32_t>(each);
  S.tmp = c->arena.get_n<uint32_t>(each);
  return S;
}

static uint64_t mix(uint64_t h, uint64_t v) {
  h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
  return h;
}

// Runs (or replays) one batched launch on the context's own stream.
static void run_batch(Ctx* c, const BatchPlan& P, const BatchVol* vols_d, const uint32_t* order_d,
                      int nvol, unsigned* queue, const BatchScratch& S, float thr,
                      int32_t* out_np, int32_t* out_ns, bool use_graph) {
  hipStream_t caller = c->stream;
  c->stream = c->own_stream;
  try {
    if (!use_graph) {
      launch_batch(c, P, vols_d, order_d, nvol, queue, S, thr, out_np, out_ns);
    } else {
      uint64_t key = 0;
      for (uint64_t v : {(uint64_t)(uintptr_t)vols_d, (uint64_t)(uintptr_t)order_d,
                         (uint64_t)(uintptr_t)queue, (uint64_t)(uintptr_t)S.nxt,
                         (uint64_t)S.stride, (uint64_t)nvol, (uint64_t)P.groups,
                         (uint64_t)P.cfg.block, (uint64_t)(uintptr_t)out_np,
                         (uint64_t)(uintptr_t)out_ns, (uint64_t)__builtin_bit_cast(uint32_t, thr)})
        key = mix(key, v);
      if (!c->batch_graph || c->batch_key != key) {
        if (c->batch_graph) (void)hipGraphExecDestroy((hipGraphExec_t)c->batch_graph);
        c->batch_graph = nullptr;
        hipGraph_t g = nullptr;
        const bool prof = c->prof;
        c->prof = false;  // no event records inside the captured region
        PNM_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        try {
          launch_batch(c, P, vols_d, order_d, nvol, queue, S, thr, out_np, out_ns);
        } catch (...) {
          (void)hipStreamEndCapture(c->stream, &g);
          if (g) (void)hipGraphDestroy(g);
          c->prof = prof;
          throw;
        }
        c->prof = prof;
        PNM_HIP(hipStreamEndCapture(c->stream, &g));
        hipGraphExec_t ge = nullptr;
        const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        PNM_HIP(e);
        c->batch_graph = ge;
        c->batch_key = key;
      }
      KTimer t(c, "k_seg_batch(graph)");
      PNM_HIP(hipGraphLaunch((hipGraphExec_t)c->batch_graph, c->stream));
    }
    PNM_HIP(hipStreamSynchronize(c->stream));
  } catch (...) {
    c->stream = caller;
    throw;
  }
  c->stream = caller;
}

void segment_batch_device(Ctx* c, int nvol, const float* const* sdf, int32_t* const* seg,
                          const int32_t* dims, const float spc[3], int flags,
                          int32_t* out_np, int32_t* out_ns) {
  std::vector<BatchVol> hv((size_t)nvol);
  std::vector<int64_t> sizes((size_t)nvol);
  int64_t maxn = 1, total = 0;
  for (int k = 0; k < nvol; ++k) {
    hv[k] = BatchVol{sdf[k], seg[k], dims[3 * k], dims[3 * k + 1], dims[3 * k + 2], 0};
    sizes[k] = (int64_t)dims[3 * k] * dims[3 * k + 1] * dims[3 * k + 2];
    maxn = std::max(maxn, sizes[k]);
    total += sizes[k];
  }
  std::vector<uint32_t> ord((size_t)nvol);
  std::iota(ord.begin(), ord.end(), 0u);
  std::stable_sort(ord.begin(), ord.end(),
                   [&](uint32_t a, uint32_t b) { return sizes[a] > sizes[b]; });
  const int64_t stride = (maxn + 63) & ~int64_t(63);
  const int cus = batch_cu_count(c);
  size_t free_b = 0, total_b = 0;
  PNM_HIP(hipMemGetInfo(&free_b, &total_b));
  // workgroup scratch may use up to half of the memory available to this call (8 arrays of
  // `stride` words each): the free memory plus what the context's arena already holds, which
  // this call reuses. (Free memory alone shrinks after the first call has allocated its scratch,
  // which planned fewer workgroups on every later call than the autotuner had timed, and
  // re-captured the graph.)
  const int64_t avail = (int64_t)free_b + c->arena.bytes();
  const int64_t per_group = stride * 4 * 8;
  const int max_groups = (int)std::max<int64_t>(1, (avail / 2) / per_group);

  BatchVol* vols_d = c->arena.get_n<BatchVol>((size_t)nvol);
  uint32_t* order_d = c->arena.get_n<uint32_t>((size_t)nvol);
  unsigned* queue = c->arena.get_n<unsigned>(1);
  PNM_HIP(hipMemcpyAsync(vols_d, hv.data(), sizeof(BatchVol) * nvol, hipMemcpyHostToDevice,
                         c->stream));
  PNM_HIP(hipMemcpyAsync(order_d, ord.data(), sizeof(uint32_t) * nvol, hipMemcpyHostToDevice,
                         c->stream));
  PNM_HIP(hipStreamSynchronize(c->stream));
  const float mn = std::min(spc[0], std::min(spc[1], spc[2]));
  const float thr = -1.5f * mn;

  auto plan_for = [&](BatchConfig cfg) {
    int groups = std::min(std::max(1, cfg.groups), max_groups);
    groups = std::min(groups, std::max(1, nvol));
    return BatchPlan{cfg, groups, stride};
  };
  // Candidate launch shapes: threads per workgroup x persistent workgroups, the workgroup count
  // searched in steps of half a workgroup per CU (measured: with 512 threads, 3.6 workgroups
  // per CU ran this batch 1.4x faster than 3.85, so whole numbers per CU are not enough).
  std::vector<BatchConfig> cands;
  for (int half = 4; half <= 8; ++half) cands.push_back({512, cus * half / 2});
  for (int per = 4; per <= 8; per += 2) cands.push_back({256, cus * per});
  for (int per = 1; per <= 2; ++per) cands.push_back({1024, cus * per});
  // The scratch is sized for the largest plan considered, once per batch shape: a later call of
  // the same shape keeps that size while it still fits (the memory estimate moves a little from
  // call to call; a larger request would re-allocate ~100 GB of arena slots and re-capture the
  // graph, 0.46 s measured on the 4096-crop batch).
  const bool tune = (flags & PNM_BATCH_AUTOTUNE) != 0;
  const uint64_t shape_key = mix(mix(mix(0, (uint64_t)nvol), (uint64_t)total), (uint64_t)maxn);
  int groups_max = 1;
  for (const BatchConfig& cand : cands) groups_max = std::max(groups_max, plan_for(cand).groups);
  if (c->batch_scratch_shape == shape_key && c->batch_scratch_groups > 0 &&
      c->batch_scratch_groups <= max_groups)
    groups_max = c->batch_scratch_groups;
  c->batch_scratch_groups = groups_max;
  c->batch_scratch_shape = shape_key;
  auto plan_in = [&](BatchConfig cfg) {  // a plan that fits the scratch
    BatchPlan P = plan_for(cfg);
    P.groups = std::min(P.groups, groups_max);
    return P;
  };
  BatchConfig cfg = c->batch_cfg_valid ? BatchConfig{c->batch_block, c->batch_groups}
                                       : BatchConfig{1024, cus * 2};
  BatchScratch S = make_scratch(c, groups_max, stride);
  if (tune && !(c->batch_cfg_valid && c->batch_shape == shape_key)) {
    float best = 1e30f;
    hipEvent_t e0, e1;
    PNM_HIP(hipEventCreate(&e0));
    PNM_HIP(hipEventCreate(&e1));
    int last_groups = -1, last_block = -1;
    for (const BatchConfig& cand : cands) {
      const BatchPlan P = plan_in(cand);
      if (P.groups == last_groups && cand.block == last_block) continue;  // capped duplicate
      last_groups = P.groups;
      last_block = cand.block;
      run_batch(c, P, vols_d, order_d, nvol, queue, S, thr, out_np, out_ns, false);  // warm
      PNM_HIP(hipEventRecord(e0, c->own_stream));
      run_batch(c, P, vols_d, order_d, nvol, queue, S, thr, out_np, out_ns, false);
      PNM_HIP(hipEventRecord(e1, c->own_stream));
      PNM_HIP(hipEventSynchronize(e1));
      float ms = 0.0f;
      PNM_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) {
        best = ms;
        cfg = BatchConfig{cand.block, P.groups};  // the count timed, kept for every later call
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->batch_cfg_valid = true;
    c->batch_shape = shape_key;
    c->batch_tuned_ms = best;
  }
  const BatchPlan P = plan_in(cfg);
  c->batch_groups = P.groups;
  c->batch_block = cfg.block;
  c->batch_per_cu = (P.groups + cus / 2) / std::max(cus, 1);
  run_batch(c, P, vols_d, order_d, nvol, queue, S, thr, out_np, out_ns,
            (flags & PNM_BATCH_GRAPH) != 0);
}

}  // namespace pnm
