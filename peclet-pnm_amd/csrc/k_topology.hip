// k_topology.hip — boundary-pair topology (extractTopologyView, pore_extraction.hpp:401-451).
//
// The reference appends one (min,max) pair per differing +x/+y/+z face (tens of millions at
// 512^3) and sorts them on the host. Here each block deduplicates the pairs of a 64x8x16 brick in
// an LDS hash set and appends only its distinct keys; the short candidate list is then radix
// sorted and made unique on the device. The result is the same sorted unique list.
//
// Key encoding: key = (uint64)(l1 - lo) << bits | (l2 - lo), with lo the smallest possible
// label, so unsigned key order is signed lexicographic (l1, l2) order.
#include "k_common.hpp"

namespace pnm {

constexpr int TOPO_BY = 8;          // rows per block, one wave each
constexpr int TOPO_ZB = 16;         // z-slices per block (a 64x8x16 brick)
constexpr int TOPO_THREADS = 64 * TOPO_BY;
constexpr int TOPO_TABLE = 512;     // LDS hash slots (4 KiB): a brick has ~80 distinct pairs
constexpr int TOPO_PROBES = 32;     // linear-probe limit before a key bypasses the table
constexpr int TOPO_FACES = 256;     // per-wave list of differing faces (fused kernel)
constexpr uint64_t kEmpty = ~0ull;

// Slot of a pair key in the brick's table (9 bits), from full-rate 24-bit multiplies only (a
// 32- or 64-bit integer multiply issues at a quarter of the VALU rate). The top bits of a
// 24 x 24-bit product's low word depend on all 24 bits of each factor. __umul24 returns a
// signed int: the product is taken as unsigned before the shift, which must be logical (an
// arithmetic shift would give a negative slot, an index outside the table).
static_assert(TOPO_TABLE == 512, "slot_hash returns 9 bits");
__device__ __forceinline__ uint32_t slot_hash(uint64_t k) {
  uint32_t x = (uint32_t)k ^ (uint32_t)__umul24((uint32_t)(k >> 32), 0x9E3779u);
  x ^= x >> 16;
  return (uint32_t)__umul24(x, 0x85EBCBu) >> 23;
}

// Each block deduplicates the differing-face pairs of its brick: a lane whose key equals the
// previous lane's key for the same direction skips (rows along a boundary repeat the pair), the
// rest go through an LDS hash set; the set is staged per block (no global counter). Keys that
// find no slot within the probe limit are appended to a small overflow list.
// Source of the labels: a finished seg array, or (SegIds) the segmentation's last step fused in
// — relabel_seg (:378-381) computed from the seg source words (kPoreFlag) for the brick and its
// +x/+y/+z halo into LDS, the brick's part written out as seg.
struct SegIds {
  const int32_t* segsrc;
  const int32_t* pore_id;
  const int32_t* solid_id;
  int32_t* seg_out;
};
constexpr int TB_X = 65, TB_Y = TOPO_BY + 1, TB_Z = TOPO_ZB + 1;  // brick + forward halo

template <bool FUSED>
__global__ __launch_bounds__(TOPO_THREADS) void k_topo_cand(const int32_t* __restrict__ seg, SegIds ids,
                                                   int nx, int ny, int nz, int32_t lo, int bits,
                                                   uint64_t* __restrict__ stage,
                                                   uint32_t* __restrict__ scnt,
                                                   uint64_t* __restrict__ overflow,
                                                   unsigned long long* ocount) {
  __shared__ unsigned long long table[TOPO_TABLE];
  __shared__ int S[FUSED ? TB_X * (TB_Y * TB_Z + 1) : 1];  // + a spare row (see below)
  // flat offset of each region row; bit 31: a row of the brick inside the volume (writes seg)
  __shared__ uint32_t RBASE[FUSED ? TB_Y * TB_Z : 1];
  // wave counts of the staging, by round parity (a round writes the set the previous one did
  // not read, so block_stage_pos's one barrier orders it)
  __shared__ unsigned wc[2][TOPO_BY];
  // per wave: the brick's differing faces not yet looked up (S index | direction << 14)
  __shared__ uint16_t PF[FUSED ? TOPO_BY * TOPO_FACES : 1];
  const int tid = threadIdx.y * 64 + threadIdx.x;
  const int lane = threadIdx.x, wave = threadIdx.y;
  for (int s = tid; s < TOPO_TABLE; s += TOPO_THREADS) table[s] = kEmpty;
  const int x0 = blockIdx.x * 64, y0 = blockIdx.y * TOPO_BY, z0 = blockIdx.z * TOPO_ZB;
  constexpr int NROW = TB_Y * TB_Z;
  if (FUSED && tid < NROW) {  // wrapped flat offset of each row of the region, once per block
    const int ry = tid % TB_Y, rz = tid / TB_Y;
    const bool brick = ry < TB_Y - 1 && rz < TOPO_ZB && y0 + ry < ny && z0 + rz < nz;
    RBASE[tid] = ((uint32_t)wrap_any(z0 + rz, nz) * (uint32_t)ny + (uint32_t)wrap_any(y0 + ry, ny)) *
                     (uint32_t)nx |
                 (brick ? 0x80000000u : 0u);
  }
  if (FUSED) __syncthreads();
  if (FUSED) {
    // Rows of the brick + halo over the waves, lane = column; the halo column 64 of the NROW rows
    // is one more chain on threads 0..NROW-1, issued with the first batch. Positions past the volume
    // hold the wrapped voxel: the +x/+y/+z neighbour of the volume's last voxel is read there.
    // Brick voxels inside the volume also write seg.
    // Each seg source word gives the id (pore id of a flagged root, solid id of a label, 0 for
    // -1). Loads are issued in branch-free batches of RB rows into registers (rows past the
    // region re-read a valid row). Row offsets come from the table (flat offsets fit 32 bits:
    // volumes hold fewer than 2^31 voxels).
    constexpr int RB = 10, MAXR = (NROW + TOPO_BY - 1) / TOPO_BY;
    static_assert(NROW <= TOPO_THREADS, "one halo chain per thread");
    const int gx = wrap_any(x0 + lane, nx), gxh = wrap_any(x0 + 64, nx);
    auto row_base = [&](int row) -> int64_t { return (int64_t)(RBASE[row] & 0x7fffffffu); };
    auto id_of = [&](int k) {
      const int a = k < -1 ? ids.pore_id[(uint32_t)k & ~kPoreFlag]
                           : ids.solid_id[k >= 0 ? k : 0];
      return k == -1 ? 0 : a;
    };
    const int hrow = tid < NROW ? tid : 0;
    const int swave = __builtin_amdgcn_readfirstlane(wave);  // row math in scalar registers
#pragma unroll 1
    for (int j0 = 0; j0 < MAXR; j0 += RB) {
      const bool halo = j0 == 0;
      uint32_t rb[RB];  // the rows' table words (offset, brick flag)
      int k[RB];
#pragma unroll
      for (int j = 0; j < RB; ++j) {
        const int row0 = swave + (j0 + j) * TOPO_BY;
        rb[j] = RBASE[row0 < NROW ? row0 : swave];
        k[j] = ids.segsrc[(rb[j] & 0x7fffffffu) + (uint32_t)gx];
      }
      // both batches load it (branch-free; cached the second time)
      const int kh = ids.segsrc[row_base(hrow) + gxh];
#pragma unroll
      for (int j = 0; j < RB; ++j) k[j] = id_of(k[j]);
      const int sh = id_of(kh);
      // The LDS store is unconditional (rows past the region go to the spare row): a store under
      // a branch lets the compiler sink that row's loads into the branch, one wait per load.
#pragma unroll
      for (int j = 0; j < RB; ++j) {
        const int row = swave + (j0 + j) * TOPO_BY;
        S[(row < NROW ? row : NROW) * TB_X + lane] = k[j];
        // a brick row inside the volume is its own unwrapped row: its offset plus gx is the
        // voxel's flat index
        if (row < NROW && (rb[j] >> 31) && x0 + lane < nx)
          ids.seg_out[(rb[j] & 0x7fffffffu) + (uint32_t)gx] = k[j];
      }
      if (halo && tid < NROW) S[tid * TB_X + 64] = sh;
    }
  }
  __syncthreads();
  const int x = x0 + threadIdx.x;
  const int y = y0 + threadIdx.y;
  const bool in = x < nx && y < ny;
  const int xp = in ? wp1(x, nx) : 0, yp = in ? wp1(y, ny) : 0;
  // Given seg (topology alone): a face pair repeats along x (the previous lane) and along z
  // (this lane's previous slice): both repeats skip the hash set. The keys are computed
  // branch-free; a direction in which no lane of the wave has a new key is skipped as a whole.
  uint64_t last[3] = {kEmpty, kEmpty, kEmpty};
  auto insert_slow = [&](uint64_t key, uint32_t h) {
#pragma unroll 1
    for (int p = 1; p < TOPO_PROBES; ++p) {
      h = (h + 1) & (TOPO_TABLE - 1);
      const unsigned long long cur = table[h];
      if (cur == key) return;
      if (cur == kEmpty) {
        const unsigned long long prv = atomicCAS(&table[h], kEmpty, (unsigned long long)key);
        if (prv == kEmpty || prv == key) return;
      }
    }
    // crowded table: the global unique pass removes repeats
    const unsigned long long sl = atomicAdd(ocount, 1ull);
    overflow[sl] = key;
  };
  // The first probe of a key (the key found, or inserted into an empty slot) is straight-line
  // code; only lanes whose first slot held another key continue probing.
  auto insert = [&](uint64_t key, bool need) {
    const uint32_t h = slot_hash(key);
    bool placed = true;
    if (need) {
      const unsigned long long cur = table[h];
      placed = cur == key;
      if (cur == kEmpty) {
        const unsigned long long prv = atomicCAS(&table[h], kEmpty, (unsigned long long)key);
        placed = prv == kEmpty || prv == key;
      }
    }
    if (__any(!placed) && !placed) insert_slow(key, h);
  };
  if constexpr (FUSED) {
    // Only about one face in ten separates two labels. Each wave lists its differing faces by
    // ballot (S index and direction, in LDS) and computes the pair keys of the list 64 at a
    // time, when the list may not hold another slice's faces and at the end: the key work runs
    // on full waves. A listed face whose key equals the previous entry's (a boundary run along
    // x) skips the set; repeats along z find their key at the first probe.
    uint16_t* const pf = PF + wave * TOPO_FACES;
    unsigned np = 0;
    auto flush = [&]() {
      __builtin_amdgcn_wave_barrier();
      for (unsigned j0 = 0; j0 < np; j0 += 64) {
        const unsigned j = j0 + lane;
        uint64_t key = kEmpty;
        if (j < np) {
          const uint32_t f = pf[j];
          const int e = (int)(f & 0x3fffu), k = (int)(f >> 14);
          const int a = S[e], b = S[e + (k == 0 ? 1 : (k == 1 ? TB_X : TB_X * TB_Y))];
          const int l1 = a < b ? a : b, l2 = a < b ? b : a;
          key = ((uint64_t)((uint32_t)l1 - (uint32_t)lo) << bits) |
                (uint64_t)((uint32_t)l2 - (uint32_t)lo);
        }
        const uint64_t prev = ((uint64_t)lane_prev((uint32_t)(key >> 32), ~0u) << 32) |
                              lane_prev((uint32_t)key, ~0u);
        insert(key, key != kEmpty && prev != key);
      }
      __builtin_amdgcn_wave_barrier();
      np = 0;
    };
    static_assert(((TOPO_ZB - 1) * TB_Y + 3) * TB_X + 63 < (1 << 14), "S index in 14 bits");
    static_assert(TOPO_FACES >= 4 * 64, "room for a slice's faces after the flush test");
    for (int zz = 0; zz < TOPO_ZB; ++zz) {
      const bool act = in && z0 + zz < nz;
      const int e = (zz * TB_Y + wave) * TB_X + lane;
      int my = 0, nb[3] = {0, 0, 0};
      if (act) {
        my = S[e];
        nb[0] = S[e + 1];
        nb[1] = S[e + TB_X];
        nb[2] = S[e + TB_X * TB_Y];
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const bool d = act && my != nb[k];
        const uint64_t dm = __ballot(d);
        if (d) pf[np + lanes_below(dm)] = (uint16_t)(e | (k << 14));
        np += (unsigned)__popcll(dm);
      }
      if (np > (unsigned)(TOPO_FACES - 3 * 64)) flush();  // wave-uniform
    }
    flush();
  } else {
  for (int zz = 0; zz < TOPO_ZB; ++zz) {
    const int z = z0 + zz;
    const bool act = in && z < nz;
    int my = 0, nb[3] = {0, 0, 0};
    if (act) {
      if (FUSED) {
        const int e = (zz * TB_Y + wave) * TB_X + lane;
        my = S[e];
        nb[0] = S[e + 1];
        nb[1] = S[e + TB_X];
        nb[2] = S[e + TB_X * TB_Y];
      } else {
        const int zp = wp1(z, nz);
        my = seg[(z * ny + y) * nx + x];
        nb[0] = seg[(z * ny + y) * nx + xp];
        nb[1] = seg[(z * ny + yp) * nx + x];
        nb[2] = seg[(zp * ny + y) * nx + x];
      }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int nl = nb[k];
      const int l1 = my < nl ? my : nl, l2 = my < nl ? nl : my;
      const uint64_t kv = ((uint64_t)((uint32_t)l1 - (uint32_t)lo) << bits) |
                          (uint64_t)((uint32_t)l2 - (uint32_t)lo);
      const uint64_t key = act && my != nl ? kv : kEmpty;
      // the previous lane's key (kEmpty at lane 0)
      const uint64_t prev = ((uint64_t)lane_prev((uint32_t)(key >> 32), ~0u) << 32) |
                            lane_prev((uint32_t)key, ~0u);
      const bool need = key != kEmpty && key != last[k] && prev != key;
      last[k] = key;
      if (!__any(need)) continue;  // wave-uniform
      insert(key, need);
    }
  }
  }
  __syncthreads();
  const int64_t bid = ((int64_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
  unsigned sbase = 0;  // block-uniform
  for (int s0 = 0, par = 0; s0 < TOPO_TABLE; s0 += TOPO_THREADS, par ^= 1) {
    const uint64_t key = table[s0 + tid];
    const bool has = key != kEmpty;
    unsigned tot;
    const unsigned pos = block_stage_pos<TOPO_BY>(has, wc[par], wave, lane, tot);
    if (has) stage[bid * TOPO_TABLE + sbase + pos] = key;
    sbase += tot;
  }
  if (tid == 0) scnt[bid] = sbase;
}

// Appends the overflow keys after the staged ones (both counts on the device).
__global__ void k_topo_append(uint64_t* __restrict__ cand, const uint64_t* __restrict__ cnt,
                              const uint64_t* __restrict__ ovf) {
  const uint64_t base = cnt[0], nov = cnt[1];
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nov;
       j += (uint64_t)gridDim.x * blockDim.x)
    cand[base + j] = ovf[j];
}

// The unique count is read on the device (the grid covers every candidate), and nothing is
// written when it exceeds the caller's capacity.
__global__ void k_topo_decode(const uint64_t* __restrict__ keys,
                              const uint64_t* __restrict__ count_dev, int64_t cap, int32_t lo,
                              int bits, int32_t* __restrict__ pairs) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t count = (int64_t)*count_dev;
  if (count > cap || k >= count) return;
  const uint64_t key = keys[k];
  const uint64_t mask = (bits >= 64) ? ~0ull : ((1ull << bits) - 1);
  pairs[2 * k] = (int32_t)((uint32_t)(key >> bits) + (uint32_t)lo);
  pairs[2 * k + 1] = (int32_t)((uint32_t)(key & mask) + (uint32_t)lo);
}

static int nbits(uint64_t v) {
  int b = 0;
  while (b < 64 && (v >> b) != 0) ++b;
  return b;
}

static int64_t topology_impl(Ctx* c, const int32_t* seg, const SegIds* ids, Vol v,
                             int32_t lo_label, int32_t hi_label, bool labels_known,
                             int32_t* pairs_d, int64_t cap, int64_t* needed) {
  int32_t lo = INT_MIN;
  int bits = 32;
  if (labels_known) {
    lo = lo_label;
    bits = nbits((uint64_t)((int64_t)hi_label - (int64_t)lo_label));
    if (bits < 1) bits = 1;
  }
  const dim3 grid(cdiv(v.nx, 64), cdiv(v.ny, TOPO_BY), cdiv(v.nz, TOPO_ZB));
  const int64_t nblk = (int64_t)grid.x * grid.y * grid.z;
  const int64_t maxc = 3 * v.n;
  uint64_t* stage = c->arena.get_n<uint64_t>((size_t)(nblk * TOPO_TABLE));
  uint32_t* scnt = c->arena.get_n<uint32_t>((size_t)nblk);
  uint64_t* cand = c->arena.get_n<uint64_t>((size_t)maxc);
  uint64_t* ovf = c->arena.get_n<uint64_t>((size_t)maxc);
  uint64_t* cnt = c->arena.get_n<uint64_t>(2);
  fill_u64(c, cnt, 0, 2);
  // keys that find no slot in a brick's table go to `ovf`; the staged tables are compacted to
  // the head of `cand`, and the overflow (rare) is appended after them once the counts are read
  // back
  if (ids)
    PNM_LAUNCH(c, "k_seg_topo", k_topo_cand<true>, grid, dim3(64, TOPO_BY, 1), 0, seg, *ids, v.nx,
               v.ny, v.nz, lo, bits, stage, scnt, ovf,
               reinterpret_cast<unsigned long long*>(cnt + 1));
  else
    PNM_LAUNCH(c, "k_topo_cand", k_topo_cand<false>, grid, dim3(64, TOPO_BY, 1), 0, seg, SegIds{},
               v.nx, v.ny, v.nz, lo, bits, stage, scnt, ovf,
               reinterpret_cast<unsigned long long*>(cnt + 1));
  compact_staged64(c, stage, scnt, nblk, TOPO_TABLE, cand, cnt);
  uint64_t h[2];
  readback(c, cnt, 2, h);
  const int64_t ncand = (int64_t)h[0] + (int64_t)h[1];
  if (h[1] > 0)  // keys that found no slot in their brick's table (rare)
    PNM_LAUNCH(c, "k_topo_append", k_topo_append, dim3(64), dim3(256), 0, cand, cnt, ovf);
  if (ncand == 0) {
    *needed = 0;
    return 0;
  }
  radix_sort_u64(c, cand, nullptr, ncand, 0, 2 * bits);
  uint64_t* uniq = c->arena.get_n<uint64_t>((size_t)ncand);
  uint64_t* nu_dev = c->arena.get_n<uint64_t>(1);
  unique_sorted_u64_dev(c, cand, uniq, ncand, nu_dev);
  // decoded before the count is read back: one host round trip for both
  PNM_LAUNCH(c, "k_topo_decode", k_topo_decode, dim3(cdiv(ncand, 256)), dim3(256), 0, uniq,
             nu_dev, cap, lo, bits, pairs_d);
  uint64_t hn;
  readback(c, nu_dev, 1, &hn);
  const int64_t nu = (int64_t)hn;
  *needed = nu;
  return nu > cap ? -1 : nu;
}

int64_t topology(Ctx* c, const int32_t* seg, Vol v, int32_t lo_label, int32_t hi_label,
                 bool labels_known, int32_t* pairs_d, int64_t cap, int64_t* needed) {
  return topology_impl(c, seg, nullptr, v, lo_label, hi_label, labels_known, pairs_d, cap,
                       needed);
}

int64_t seg_topology(Ctx* c, const int32_t* segsrc, const int32_t* pore_id,
                     const int32_t* solid_id, int32_t* seg, Vol v, int32_t lo_label,
                     int32_t hi_label, int32_t* pairs_d, int64_t cap, int64_t* needed) {
  const SegIds ids{segsrc, pore_id, solid_id, seg};
  return topology_impl(c, nullptr, &ids, v, lo_label, hi_label, true, pairs_d, cap, needed);
}

}  // namespace pnm
