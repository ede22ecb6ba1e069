// k_segment.hip — marker-controlled watershed on gfx950 (segmentVolumeView,
// pore_extraction.hpp:157-386), reformulated around the order-free fixpoints:
//   CCL      : label = minimum flat index of the periodic 26-connected marker component;
//              lock-free union-find with atomicMin hooking (larger root under smaller).
//   flood    : Jacobi sweeps restricted to a compacted band list (the only voxels a sweep can
//              change), double-buffered through a per-band-entry label array.
//   roots    : chase the ascent successor array up to 512 steps (exactly the capped walk).
//   renumber : per-key first-appearance minimum by atomicMin, then a rank by radix sort of the
//              (few) keys instead of two full-volume scans.
#include <algorithm>

#include "k_common.hpp"

namespace pnm {

static dim3 vox_grid(Vol v) { return dim3(cdiv(v.nx, 64), cdiv(v.ny, 4), (unsigned)v.nz); }
static const dim3 kVoxBlock(64, 4, 1);

// ---------------- CCL ----------------
__global__ __launch_bounds__(256) void k_ccl_runs(int* L, int nx, int ny, int nz) {
  runs_union_chunk(L, nx, ny, nz, blockIdx.x * 64 + threadIdx.x, blockIdx.y * 4 + threadIdx.y,
                   blockIdx.z, threadIdx.x);
}

// flatten (pore_extraction.hpp:183-194). Component roots (= component minima) are staged per
// block as the solid keys; min_solid[root] starts at the root itself (the component's smallest
// marker); only flood-filled voxels with a smaller index can lower it (k_flood_apply).
__global__ __launch_bounds__(256) void k_ccl_flatten(int* L, int64_t n, int32_t* min_solid,
                                                     uint32_t* stage, uint32_t* cnt) {
  __shared__ unsigned wc[4];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  bool root = false;
  if (i < n) {
    int l = L[i];
    if (l >= 0) {  // a marker (every other voxel holds a negative code)
      int p = L[l];
      while (p != l) {
        l = p;
        p = L[l];
      }
      L[i] = l;
      root = (l == (int)i);
      if (root) min_solid[i] = (int)i;
    }
  }
  unsigned tot;
  const unsigned pos = block_stage_pos<4>(root, wc, wave, lane, tot);
  if (root) stage[(int64_t)blockIdx.x * 256 + pos] = (uint32_t)i;
  if (threadIdx.x == 0) cnt[blockIdx.x] = tot;
}

void launch_ccl_flatten(Ctx* c, Vol v, int32_t* labels, int32_t* min_solid, uint32_t* key_stage,
                        uint32_t* key_cnt) {
  PNM_LAUNCH(c, "k_ccl_flatten", k_ccl_flatten, dim3(cdiv(v.n, 256)), dim3(256), 0, labels, v.n,
             min_solid, key_stage, key_cnt);
}

void launch_ccl(Ctx* c, Vol v, int32_t* labels, int32_t* min_solid, uint32_t* key_stage,
                uint32_t* key_cnt) {
  PNM_LAUNCH(c, "k_ccl_runs", k_ccl_runs, vox_grid(v), kVoxBlock, 0, labels, v.nx, v.ny, v.nz);
  launch_ccl_flatten(c, v, labels, min_solid, key_stage, key_cnt);
}

// ---------------- flatten + first flood sweeps, tile by tile ----------------
// flatten (pore_extraction.hpp:183-194) and the first FLOOD_SWEEPS Jacobi sweeps of flood
// (:237-269) in one pass. A 512-thread block loads its 64x8x16 tile plus a halo of
// FLOOD_SWEEPS voxels (labels resolved to their component root, band candidates = unlabelled
// !(sdf >= 0)) into LDS and sweeps there: sweep k is exact on the region shrunk by k voxels, so
// after FLOOD_SWEEPS sweeps the tile holds exactly the global state after those sweeps. It writes
// the tile's labels to a second array (neighbouring blocks keep reading the unflooded labels of
// this tile as their halo), lowers the solid first-appearance minimum for voxels it labelled
// (relabel_min :335-336; every marker of a label comes after its root, which k_tile seeded),
// and stages the component roots (solid keys) and the voxels still unlabelled (the band of the
// next global sweep).
constexpr int FLOOD_SWEEPS = 1;
constexpr int FR_X = 64 + 2 * FLOOD_SWEEPS, FR_Y = TILE_Y + 2 * FLOOD_SWEEPS,
              FR_Z = TILE_Z + 2 * FLOOD_SWEEPS;
constexpr int FR_N = FR_X * FR_Y * FR_Z;
constexpr int FR_SEG = 64 * TILE_Z;  // staging slots per wave
constexpr int FLOOD_LIST = 3072;     // LDS list of band candidates per tile (3 blocks/CU)
// region word: >= 0 label, -1 band candidate, -2 other unlabelled, <= -3 labelled this sweep

__global__ __launch_bounds__(64 * TILE_Y) void k_flood_tile(const int* __restrict__ L,
                                                            int* __restrict__ Lout, int nx, int ny,
                                                            int nz, int32_t* __restrict__ min_solid,
                                                            uint32_t* __restrict__ key_stage,
                                                            uint32_t* __restrict__ key_cnt,
                                                            uint32_t* __restrict__ band_stage,
                                                            uint32_t* __restrict__ band_cnt,
                                                            const int32_t* __restrict__ xface,
                                                            bool write_all) {
  __shared__ int R[FR_N];
  __shared__ uint16_t cand_list[FLOOD_LIST];
  const int lane = threadIdx.x, wave = threadIdx.y;
  const int tid = wave * 64 + lane;
  constexpr int NT = 64 * TILE_Y;
  const int x0 = blockIdx.x * 64, y0 = blockIdx.y * TILE_Y, z0 = blockIdx.z * TILE_Z;
  const int nzt = min(TILE_Z, nz - z0);
  const int64_t seg_id =
      (((int64_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x) * TILE_Y + wave;
  constexpr int H = FLOOD_SWEEPS;
  // load: labels (a marker's component is L[L[i]]: its tile-local root points at the component
  // root) and band candidates (the tile pass wrote -1 for a voxel that is no marker and not
  // sdf >= 0, -2 for the other non-markers). Rows of the region are spread over the waves, a
  // lane holding
  // column x0 + lane (rows aligned to the chunk: two 128-byte lines per row and array). The
  // halo columns x0-1 and x0+64 are the neighbouring chunks' edge columns, whose tile labels
  // the tile pass wrote compactly (x-face pairs): one more load chain on the first threads,
  // issued with the first batch, with no value needed (halo entries are never swept, so an
  // unlabelled one is just "not a label"). A partial last chunk reads its right halo column
  // directly. Per batch: the raw labels, then the component roots, then one LDS store. Loads are issued in branch-free batches of RB rows into registers (rows past the
  // region re-read a valid row). Each region row's wrapped (z*ny + y) and flat offset come from
  // tables the block fills first, kept in the candidate list's LDS (not used before the sweep);
  // flat offsets fit 32 bits (volumes hold fewer than 2^31 voxels).
  static_assert(H == 1, "the x halo is one column on each side");
  const int rows = FR_Y * (nzt + 2 * H);
  const int gx0 = wrap_any(x0 + lane, nx);
  constexpr int RB = 12;
  constexpr int MAXR = (FR_Y * FR_Z + TILE_Y - 1) / TILE_Y;  // rows per wave, full tile
  uint32_t* rzy = reinterpret_cast<uint32_t*>(cand_list);
  uint32_t* rbase = rzy + FR_Y * FR_Z;
  static_assert(sizeof(cand_list) >= 2 * sizeof(uint32_t) * FR_Y * FR_Z, "row tables fit");
  if (tid < rows) {
    const int ry = tid % FR_Y, rz = tid / FR_Y;
    const uint32_t zy = (uint32_t)wrap_any(z0 - H + rz, nz) * (uint32_t)ny +
                        (uint32_t)wrap_any(y0 - H + ry, ny);
    rzy[tid] = zy;
    rbase[tid] = zy * (uint32_t)nx;
  }
  __syncthreads();
  auto row_base = [&](int row) -> int64_t { return (int64_t)rbase[row]; };
  auto region_word = [](int l, int root) { return l >= 0 ? root : l; };
  const int swave = __builtin_amdgcn_readfirstlane(wave);
  const int nhalo = rows * 2;  // <= 2 * FR_Y * FR_Z < NT
  const int hrow = tid < nhalo ? tid >> 1 : 0, hside = tid & 1;  // 0: x0-1, 1: x0+64
  const int cx = gridDim.x;
  const int64_t nzy = (int64_t)nz * ny;
  const int bxm = blockIdx.x > 0 ? (int)blockIdx.x - 1 : cx - 1;
  const int bxp = (int)blockIdx.x + 1 < cx ? (int)blockIdx.x + 1 : 0;
  const bool right_face = x0 + 64 < nx || nx % 64 == 0;  // x0+64 is chunk bxp's first column
  // x-face entry (chunk, z, y) holds (first column, last column): x0-1 is chunk bxm's last
  // column (for chunk 0 the volume's last column), x0+64 chunk bxp's first
  const bool from_face = hside == 0 || right_face;
  const int32_t* hsrc = from_face ? xface : L;
  const int64_t hidx = hside == 0 ? 2 * (bxm * nzy + rzy[hrow]) + 1
                                  : (right_face ? 2 * (bxp * nzy + rzy[hrow])
                                                : row_base(hrow) + wrap_any(x0 + 64, nx));
#pragma unroll 1
  for (int j0 = 0; j0 < MAXR; j0 += RB) {
    const bool halo = j0 == 0;
    int la[RB];
#pragma unroll
    for (int j = 0; j < RB; ++j) {
      const int row0 = swave + (j0 + j) * TILE_Y;
      la[j] = L[row_base(row0 < rows ? row0 : swave) + gx0];
    }
    const int lh = halo ? hsrc[hidx] : -1;
    int ra[RB];
#pragma unroll
    for (int j = 0; j < RB; ++j) ra[j] = L[la[j] >= 0 ? la[j] : 0];
    const int rh = halo ? L[lh >= 0 ? lh : 0] : -1;
#pragma unroll
    for (int j = 0; j < RB; ++j) {
      const int row = swave + (j0 + j) * TILE_Y;
      if (row < rows) R[row * FR_X + lane + 1] = region_word(la[j], ra[j]);
    }
    if (halo && tid < nhalo) R[hrow * FR_X + (hside ? 65 : 0)] = lh >= 0 ? rh : -2;
  }
  __syncthreads();
  // Band candidates: the tile's voxels that start unlabelled (with one sweep the exact region is
  // the tile itself). cand0 marks this thread's (flooded voxels lower the minimum); they are also
  // compacted into an LDS list, one segment per wave holding the wave's own row slice by slice,
  // filled by ballot without atomics, so that the sweep touches only them. A wave with more
  // candidates than its segment sweeps its row slice by slice instead (each wave sweeps and
  // later reads only its own row, so no barrier separates the listing from the sweep).
  static_assert(FLOOD_SWEEPS == 1, "the candidate region is the tile");
  constexpr int WL = FLOOD_LIST / TILE_Y;
  uint32_t cand0 = 0;
  unsigned nw = 0;
#pragma unroll
  for (int zl = 0; zl < TILE_Z; ++zl) {
    const int e = ((zl + H) * FR_Y + wave + H) * FR_X + lane + H;
    const bool cand = zl < nzt && R[e] == -1;
    if (cand) cand0 |= 1u << zl;
    const uint64_t m = __ballot(cand);
    const unsigned pos = nw + lanes_below(m);
    if (cand && pos < (unsigned)WL) cand_list[wave * WL + pos] = (uint16_t)e;
    nw += (unsigned)__popcll(m);
  }
  const bool listed = nw <= (unsigned)WL;  // wave-uniform
  // One sweep at region entry e: the minimum labelled neighbour. As unsigned numbers every
  // unlabelled state (negative) lies above every label (a flat index below 2^31), so the
  // minimum of the 26 words is the smallest label if there is one.
  auto sweep_at = [&](int e) {
    if (R[e] != -1) return;
    uint32_t best = ~0u;
#pragma unroll
    for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
      for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx) {
          if (dx == 0 && dy == 0 && dz == 0) continue;
          best = min(best, (uint32_t)R[e + (dz * FR_Y + dy) * FR_X + dx]);
        }
    if (best < 0x80000000u) R[e] = -3 - (int)best;  // labelled this sweep: reads as unlabelled
  };
  // the one sweep, exact on the tile (the region with its one-voxel halo), over the wave's row
  if (listed) {
    for (unsigned j = lane; j < nw; j += 64) sweep_at(cand_list[wave * WL + j]);
  } else {
    for (int zl = 0; zl < nzt; ++zl) sweep_at(((zl + H) * FR_Y + wave + H) * FR_X + lane + H);
  }
  __syncthreads();
  // the sweep's labels (-3 - label) are decoded where the outputs read them
  // tile outputs, staged per wave in (slice, lane) order
  const int x = x0 + lane, y = y0 + wave;
  const bool col_in = x < nx && y < ny;
  unsigned nk = 0, nb = 0;
#pragma unroll
  for (int zl = 0; zl < TILE_Z; ++zl) {
    if (zl >= nzt) break;
    const int e = ((zl + H) * FR_Y + wave + H) * FR_X + lane + H;
    const int i = ((z0 + zl) * ny + y) * nx + x;
    const int rv = col_in ? R[e] : -2;
    const int v = rv <= -3 ? -3 - rv : rv;  // labelled by the last sweep
    bool root = false;
    if (col_in) {
      // -2: sdf >= 0 and unlabelled, a voxel whose seg source word the tile pass or the root
      // resolver writes (unless write_all)
      if (write_all || v != -2) Lout[i] = v >= 0 ? v : -1;
      if (v >= 0 && ((cand0 >> zl) & 1u) && i < v) atomicMin(min_solid + v, i);
      root = v == i;
    }
    const bool band = v == -1;
    const uint64_t km = __ballot(root);
    if (root) key_stage[seg_id * FR_SEG + nk + lanes_below(km)] = (uint32_t)i;
    nk += (unsigned)__popcll(km);
    const uint64_t bm = __ballot(band);
    if (band) band_stage[seg_id * FR_SEG + nb + lanes_below(bm)] = (uint32_t)i;
    nb += (unsigned)__popcll(bm);
  }
  if (lane == 0) {
    key_cnt[seg_id] = nk;
    band_cnt[seg_id] = nb;
  }
}

void launch_flood_tile(Ctx* c, Vol v, const int32_t* labels_ccl, int32_t* labels,
                       int32_t* min_solid, uint32_t* key_stage, uint32_t* key_cnt,
                       uint32_t* band_stage, uint32_t* band_cnt, const int32_t* xface,
                       bool write_all) {
  PNM_LAUNCH(c, "k_flood_tile", k_flood_tile,
             dim3(cdiv(v.nx, 64), cdiv(v.ny, TILE_Y), cdiv(v.nz, TILE_Z)), dim3(64, TILE_Y), 0,
             labels_ccl, labels, v.nx, v.ny, v.nz, min_solid, key_stage, key_cnt, band_stage,
             band_cnt, xface, write_all);
}

// Marker labels with every negative code as -1 (the reference's labels after its CCL).
__global__ void k_label_codes(const int32_t* __restrict__ L, int64_t n, int32_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = max(L[i], -1);
}

void launch_label_codes(Ctx* c, Vol v, const int32_t* labels, int32_t* out) {
  PNM_LAUNCH(c, "k_label_codes", k_label_codes, dim3(cdiv(v.n, 256)), dim3(256), 0, labels, v.n,
             out);
}

__global__ void k_merge_roots(const float* __restrict__ sdf, int64_t n,
                              const int32_t* __restrict__ roots, int32_t* __restrict__ segsrc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && sdf[i] > 0.0f) segsrc[i] = (int32_t)((uint32_t)roots[i] | kPoreFlag);
}

void launch_merge_roots(Ctx* c, const float* sdf, Vol v, const int32_t* roots, int32_t* segsrc) {
  PNM_LAUNCH(c, "k_merge_roots", k_merge_roots, dim3(cdiv(v.n, 256)), dim3(256), 0, sdf, v.n,
             roots, segsrc);
}

// ---------------- flood (pore_extraction.hpp:237-269) ----------------
// One Jacobi sweep over the current band list, whose length the kernels read on the device (the
// host does not size the grids, so no read back precedes a sweep): a fixed grid of FS_BLOCKS
// blocks, block b taking the b-th contiguous share of the list. The new label is the minimum
// labelled neighbour of the previous state; voxels still unlabelled are staged in the block's
// segment (seg_cap slots, at least a share) for the next band list. Neighbour words are seg
// source words: a label, -1, or a flagged root (kPoreFlag), the last two above every label as
// unsigned numbers.
__global__ __launch_bounds__(256) void k_flood_sweep(const int32_t* __restrict__ L, int nx,
                                                     int ny, int nz, UDiv dx, UDiv dy,
                                                     const uint32_t* __restrict__ band,
                                                     const uint64_t* __restrict__ count,
                                                     int32_t* __restrict__ newlab,
                                                     uint32_t* __restrict__ stage,
                                                     uint32_t* __restrict__ cnt, int64_t seg_cap) {
  // wave counts of the staging, by round parity: a round writes the set the previous round did
  // not read, so its one barrier (inside block_stage_pos) orders it with both neighbours
  __shared__ unsigned wc[2][4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t nband = (int64_t)*count;
  const int64_t share = (nband + gridDim.x - 1) / gridDim.x;
  const int64_t j_begin = (int64_t)blockIdx.x * share;
  const int64_t j_end = min(nband, j_begin + share);
  unsigned base = 0;
  int par = 0;
  for (int64_t j0 = j_begin; j0 < j_end; j0 += 256, par ^= 1) {  // block-uniform
    const int64_t j = j0 + threadIdx.x;
    bool keep = false;
    uint32_t i = 0;
    if (j < j_end) {
      i = band[j];
      uint32_t x, y;
      const uint32_t r = udiv(i, dx, x);  // r = z * ny + y
      const uint32_t z = udiv(r, dy, y);
      // neighbour offsets along each axis: -1, 0, +1 steps, wrapped at the faces
      const int plane = nx * ny;
      const int ox[3] = {x > 0 ? -1 : nx - 1, 0, (int)x + 1 < nx ? 1 : 1 - nx};
      const int oy[3] = {y > 0 ? -nx : (ny - 1) * nx, 0, (int)y + 1 < ny ? nx : (1 - ny) * nx};
      const int oz[3] = {z > 0 ? -plane : (nz - 1) * plane, 0,
                         (int)z + 1 < nz ? plane : (1 - nz) * plane};
      int nl[26];
      int k = 0;
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
          for (int cc = 0; cc < 3; ++cc) {
            if (a == 1 && b == 1 && cc == 1) continue;
            nl[k++] = L[(int)i + oz[a] + oy[b] + ox[cc]];
          }
      // a label is a flat index (below 2^31); -1 (unlabelled) and flagged roots are not
      uint32_t best = ~0u;
#pragma unroll
      for (int q = 0; q < 26; ++q) best = min(best, (uint32_t)nl[q]);
      keep = best >= kPoreFlag;
      newlab[j] = keep ? -1 : (int)best;
    }
    unsigned tot;
    const unsigned pos = block_stage_pos<4>(keep, wc[par], wave, lane, tot);
    if (keep) stage[(int64_t)blockIdx.x * seg_cap + base + pos] = i;
    base += tot;
  }
  if (threadIdx.x == 0) cnt[blockIdx.x] = base;
}

// Applies a sweep; a newly labelled voxel before its label's first appearance lowers it
// (relabel_min for solids, :335-336 — every other voxel of the label comes after the root).
__global__ void k_flood_apply(int32_t* __restrict__ L, const uint32_t* __restrict__ band,
                              const int32_t* __restrict__ newlab,
                              const uint64_t* __restrict__ count, int32_t* min_solid) {
  const int64_t nband = (int64_t)*count;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nband;
       j += (int64_t)gridDim.x * blockDim.x) {
    const int32_t b = newlab[j];
    if (b != -1) {
      const uint32_t i = band[j];
      L[i] = b;
      if ((int32_t)i < b) atomicMin(min_solid + b, (int32_t)i);
    }
  }
}

// Staging segment of a sweep block: a share of the longest band, at least the 64 slots the
// compaction's gather needs.
int64_t flood_seg_cap(int64_t max_items) { return std::max<int64_t>(cdiv(max_items, FS_BLOCKS), 64); }

void flood_sweep(Ctx* c, Vol v, int32_t* labels, const uint32_t* band, const uint64_t* count,
                 int64_t max_items, uint32_t* band_next, uint64_t* count_next, int32_t* newlab,
                 uint32_t* stage, uint32_t* stage_cnt, int32_t* min_solid) {
  const int64_t seg_cap = flood_seg_cap(max_items);
  PNM_LAUNCH(c, "k_flood_sweep", k_flood_sweep, dim3(FS_BLOCKS), dim3(256), 0, labels, v.nx,
             v.ny, v.nz, make_udiv((uint32_t)v.nx), make_udiv((uint32_t)v.ny), band, count,
             newlab, stage, stage_cnt, seg_cap);
  compact_staged(c, stage, stage_cnt, FS_BLOCKS, seg_cap, band_next, count_next);
  PNM_LAUNCH(c, "k_flood_apply", k_flood_apply, dim3(FS_BLOCKS), dim3(256), 0, labels, band,
             newlab, count, min_solid);
}

// Capped-walk fallback: the root set is no longer the peak set, so redo the pore minima from a
// fully initialised array and collect the keys by a full scan.
__global__ void k_min_pore_full(const float* __restrict__ sdf, int64_t n,
                                const int32_t* __restrict__ segsrc, int32_t* min_pore) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && sdf[i] > 0.0f) atomicMin(min_pore + ((uint32_t)segsrc[i] & ~kPoreFlag), (int)i);
}

void launch_min_full(Ctx* c, const float* sdf, Vol v, const int32_t* segsrc, int32_t* min_pore) {
  fill_u32(c, reinterpret_cast<uint32_t*>(min_pore), (uint32_t)INT_MAX, v.n);
  PNM_LAUNCH(c, "k_min_pore_full", k_min_pore_full, dim3(cdiv(v.n, 256)), dim3(256), 0, sdf, v.n,
             segsrc, min_pore);
}

__global__ void k_collect_keys(const float* __restrict__ sdf, int64_t n,
                               const int32_t* __restrict__ min_pore, uint32_t* stage,
                               uint32_t* cnt) {
  __shared__ unsigned wc[4];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bool key = (i < n) && sdf[i] > 0.0f && min_pore[i] != INT_MAX;
  unsigned tot;
  const unsigned pos = block_stage_pos<4>(key, wc, wave, lane, tot);
  if (key) stage[(int64_t)blockIdx.x * 256 + pos] = (uint32_t)i;
  if (threadIdx.x == 0) cnt[blockIdx.x] = tot;
}

int64_t collect_keys_full(Ctx* c, const float* sdf, Vol v, const int32_t* min_pore,
                          uint32_t* key_list, uint64_t* counter) {
  const int64_t nb = cdiv(v.n, 256);
  uint32_t* stage = c->arena.get_n<uint32_t>((size_t)(nb * 256));
  uint32_t* cnt = c->arena.get_n<uint32_t>((size_t)nb);
  PNM_LAUNCH(c, "k_collect_keys", k_collect_keys, dim3((unsigned)nb), dim3(256), 0, sdf, v.n,
             min_pore, stage, cnt);
  compact_staged(c, stage, cnt, nb, 256, key_list, counter);
  uint64_t h;
  readback(c, counter, 1, &h);
  return (int64_t)h;
}

// ---------------- ids by first appearance (relabel_first/scans/assign, :340-374) ----------------
// Pores: every key sets the bit of its first-appearance voxel in an n-bit map, a rank directory
// of the map (map_rank) turns a bit into its rank, and a key's id is its rank + 1; pore keys that
// are not sdf>0 (NaN peaks of the extract_pores predicate) get no id. Solids: a key's first
// voxel with sdf <= 0 gives id -(rank + 1) among those voxels, any other solid key id 0. A volume
// has few solid components, so up to SOLID_RANK_MAX of them are ranked by direct comparison of
// their first voxels in one block; more go through a map of their own like the pores.
constexpr int SOLID_RANK_MAX = 4096;

__global__ void k_mark_first(const float* __restrict__ sdf, const uint32_t* __restrict__ keys,
                             int64_t nk, const int32_t* __restrict__ mins, bool solid,
                             uint32_t* __restrict__ map) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nk) return;
  const uint32_t r = keys[j];
  const uint32_t f = (uint32_t)mins[r];
  if (solid ? sdf[f] <= 0.0f : sdf[r] > 0.0f) atomicOr(map + (f >> 5), 1u << (f & 31));
}

// Rank of bit f in a bitmap with a rank directory (map_rank): the chunk's prefix, the words of
// the chunk before f's word, the bits below f in its word.
__device__ __forceinline__ uint32_t map_rank_of(const uint32_t* __restrict__ map,
                                                const uint32_t* __restrict__ crank, int64_t f) {
  const int64_t w = f >> 5, c = w / MAP_CHUNK_WORDS;
  uint32_t r = crank[c];
  for (int64_t q = c * MAP_CHUNK_WORDS; q < w; ++q) r += (uint32_t)__popc(map[q]);
  return r + (uint32_t)__popc(map[w] & ((1u << (f & 31)) - 1u));
}

// ids[key] still holds the key's first-appearance voxel and becomes its id.
__global__ void k_assign_ranks(const float* __restrict__ sdf, const uint32_t* __restrict__ keys,
                               int64_t nk, int32_t* __restrict__ ids, bool solid,
                               const uint32_t* __restrict__ map,
                               const uint32_t* __restrict__ crank) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nk) return;
  const uint32_t r = keys[j];
  const uint32_t f = (uint32_t)ids[r];
  if (!solid) {
    if (sdf[r] > 0.0f) ids[r] = (int32_t)(map_rank_of(map, crank, f) + 1);
  } else {
    ids[r] = sdf[f] <= 0.0f ? -(int32_t)(map_rank_of(map, crank, f) + 1) : 0;
  }
}

// All solid keys in one block: every first voxel is read before any id is written.
__global__ __launch_bounds__(1024) void k_solid_ranks(const float* __restrict__ sdf,
                                                      const uint32_t* __restrict__ keys,
                                                      int64_t nk, int32_t* __restrict__ ids,
                                                      unsigned long long* __restrict__ count) {
  __shared__ uint32_t F[SOLID_RANK_MAX];  // first voxel of a key with sdf <= 0, else ~0
  __shared__ unsigned nvalid;
  if (threadIdx.x == 0) nvalid = 0;
  __syncthreads();
  const int n = (int)nk;
  unsigned mine = 0;
  for (int j = threadIdx.x; j < n; j += 1024) {
    const uint32_t f = (uint32_t)ids[keys[j]];
    const bool ok = sdf[f] <= 0.0f;
    F[j] = ok ? f : ~0u;
    mine += ok ? 1u : 0u;
  }
  atomicAdd(&nvalid, mine);
  __syncthreads();
  for (int j = threadIdx.x; j < n; j += 1024) {
    const uint32_t f = F[j];
    int32_t id = 0;
    if (f != ~0u) {  // first voxels are distinct (a voxel carries one label)
      uint32_t r = 0;
      for (int q = 0; q < n; ++q) r += F[q] < f ? 1u : 0u;
      id = -(int32_t)(r + 1);
    }
    ids[keys[j]] = id;
  }
  if (threadIdx.x == 0) *count = nvalid;
}

// Both key sets; counts_dev[0] = number of pore ids, counts_dev[1] = solid ids (no
// synchronisation).
void assign_ids_bitmap(Ctx* c, const float* sdf, Vol v, const uint32_t* pkeys, int64_t npk,
                       int32_t* pore_ids, const uint32_t* skeys, int64_t nsk, int32_t* solid_ids,
                       uint64_t* counts_dev) {
  const int64_t words = (v.n + 31) / 32;
  auto by_map = [&](const uint32_t* keys, int64_t nk, int32_t* ids, bool solid, uint64_t* cnt) {
    uint32_t* map = c->arena.get_n<uint32_t>((size_t)words);
    uint32_t* crank = c->arena.get_n<uint32_t>((size_t)cdiv(words, MAP_CHUNK_WORDS));
    fill_u32(c, map, 0u, words);
    if (nk > 0)
      PNM_LAUNCH(c, "k_mark_first", k_mark_first, dim3(cdiv(nk, 256)), dim3(256), 0, sdf, keys,
                 nk, (const int32_t*)ids, solid, map);
    map_rank(c, map, words, crank, cnt);
    if (nk > 0)
      PNM_LAUNCH(c, "k_assign_ranks", k_assign_ranks, dim3(cdiv(nk, 256)), dim3(256), 0, sdf,
                 keys, nk, ids, solid, map, crank);
  };
  by_map(pkeys, npk, pore_ids, false, counts_dev);
  if (nsk > SOLID_RANK_MAX)
    by_map(skeys, nsk, solid_ids, true, counts_dev + 1);
  else
    PNM_LAUNCH(c, "k_solid_ranks", k_solid_ranks, dim3(1), dim3(1024), 0, sdf, skeys, nsk,
               solid_ids, reinterpret_cast<unsigned long long*>(counts_dev + 1));
}

// relabel_seg (pore_extraction.hpp:378-381) from the seg source words: a flagged root takes its
// pore id, a label its solid id, -1 gives 0.
__global__ __launch_bounds__(256) void k_seg(int64_t n, const int32_t* __restrict__ segsrc,
                                             const int32_t* __restrict__ pore_id,
                                             const int32_t* __restrict__ solid_id,
                                             int32_t* __restrict__ seg) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t k = segsrc[i];
  int32_t s = 0;
  if (k < -1)
    s = pore_id[(uint32_t)k & ~kPoreFlag];
  else if (k >= 0)
    s = solid_id[k];
  seg[i] = s;
}

void launch_seg(Ctx* c, Vol v, const int32_t* segsrc, const int32_t* pore_id,
                const int32_t* solid_id, int32_t* seg) {
  PNM_LAUNCH(c, "k_seg", k_seg, dim3(cdiv(v.n, 256)), dim3(256), 0, v.n, segsrc, pore_id,
             solid_id, seg);
}

}  // namespace pnm
