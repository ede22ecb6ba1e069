// k_tile.hip — the segmentation's one pass over the SDF, tile by tile in LDS.
//
// A 512-thread block owns a 64(x) x 8(y) x 16(z) tile. It marches the tile's slices in z,
// computing for every voxel the 26-neighbour lexicographic (value, index) maximum
// (pore_extraction.hpp:93-104, :280-304) as the maximum of ordered 64-bit keys (lex_key in
// k_common.hpp) over the 3x3 column neighbourhood of three slices, and keeps two tile-sized
// arrays in LDS:
//   LAB: marker labels (init_markers :177), first as the start of the voxel's x-run;
//   DIR: for sdf>0 voxels, the direction (0..26, 13 = itself) of the ascent successor.
// After the march it finishes, without leaving LDS,
//   - the CCL inside the tile (merge_markers :198-224 restricted to in-tile, non-wrapping
//     neighbour pairs; union-find with LDS atomicMin, root = smallest local index, which is the
//     smallest flat index since the tile is a box), and writes every marker's label as the flat
//     index of its tile-local root. k_ccl_seams then adds the relations that cross tiles or wrap;
//   - the ascent walk inside the tile (gradient_path :279-309): a walk that reaches a fixed point
//     or the 512-step cap inside the tile is final and its root is written; a walk that leaves
//     the tile records (first voxel outside, steps so far) and is queued for k_roots_resolve.
//     Walk records are also written for every sdf>0 voxel on the tile's boundary — exactly the
//     voxels a walk can enter the tile through — so a resolver hops tile to tile;
//   - the first-appearance minimum of each in-tile root over the tile's final walks (relabel_min
//     :333-334), stored at the root: this seeds min_pore for every sdf>0 peak.
// The extract_pores peak predicate (:87-105) goes into a bitmap of peak voxels; marker roots seed
// the solid first-appearance minimum.
#pragma clang fp contract(off)

#include "k_common.hpp"

namespace pnm {

constexpr int T_LX = 66, T_LY = TILE_Y + 2;
constexpr int T_VOX = 64 * TILE_Y * TILE_Z;
constexpr int T_SEG = 64 * TILE_Z;  // staging slots per wave
// DIR byte: the ascent direction of an sdf>0 voxel as (dz+1) << 4 | (dy+1) << 2 | (dx+1)
// (kSelf: the voxel itself), kNoDir for other voxels; the offsets come back by bit-field
// extraction.
constexpr uint8_t kNoDir = 63;
constexpr uint8_t kDirMask = 63;
constexpr int kSelf = (1 << 4) | (1 << 2) | 1;

__device__ __forceinline__ int lds_find(int* L, int x) {
  int p = L[x];
  while (p != x) {
    const int g = L[p];
    if (g != p) L[x] = g;
    x = p;
    p = g;
  }
  return x;
}

__device__ __forceinline__ void lds_unite(int* L, int a, int b) {
  while (true) {
    a = lds_find(L, a);
    b = lds_find(L, b);
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

// Unites every x-run of row mask `a` with every x-run of row mask `b` it touches under
// 26-adjacency (columns x-1, x, x+1 inside the 64-column chunk), each run named by the local
// index of its first voxel (row bases base_a / base_b). The runs and their contacts come from
// the masks alone (bit scans), so one lane handles a whole pair of rows.
__device__ __forceinline__ void unite_row_runs(int* LAB, uint64_t a, uint64_t b, int base_a,
                                               int base_b) {
  while (a) {
    const int s = __ffsll((unsigned long long)a) - 1;
    const uint64_t from_s = ~(a >> s);
    const int e = from_s ? s + __ffsll((unsigned long long)from_s) - 2 : 63;
    const int lo = s > 0 ? s - 1 : 0, hi = e < 63 ? e + 1 : 63;
    const uint64_t win = (hi == 63 ? ~0ull : ((2ull << hi) - 1)) & ~((1ull << lo) - 1);
    uint64_t bw = b & win;
    while (bw) {
      const int t = __ffsll((unsigned long long)bw) - 1;
      const uint64_t zeros_below = t ? (~b & (~0ull >> (64 - t))) : 0ull;
      const int sb = zeros_below ? 64 - __clzll((long long)zeros_below) : 0;
      const uint64_t from_t = ~(b >> t);
      const int eb = from_t ? t + __ffsll((unsigned long long)from_t) - 2 : 63;
      lds_unite(LAB, base_a + s, base_b + sb);
      bw = eb >= 63 ? 0ull : (bw & ~((2ull << eb) - 1));
    }
    a = e >= 63 ? 0ull : (a & ~((2ull << e) - 1));
  }
}

// Chase state of an sdf>0 voxel during pointer jumping: u (local index, 13 bits), steps to u
// (10 bits), status of u: ACT = u's own walk not yet known, FIX = u is a fixed point, EXIT = u's
// successor lies outside the tile, LONG = the walk takes more than 512 steps inside the tile.
constexpr uint32_t CS_ACT = 0, CS_FIX = 1, CS_EXIT = 2, CS_LONG = 3;
__device__ __forceinline__ uint32_t cs_pack(int u, int s, uint32_t st) {
  return (uint32_t)u | ((uint32_t)s << 13) | (st << 23);
}
__device__ __forceinline__ int cs_u(uint32_t w) { return (int)(w & 0x1fffu); }
__device__ __forceinline__ int cs_s(uint32_t w) { return (int)((w >> 13) & 0x3ffu); }
__device__ __forceinline__ uint32_t cs_st(uint32_t w) { return w >> 23; }

// Staged walk that left its tile: its walk record (target, steps; bits 0..49) with the voxel's
// slice and column in its tile row in bits 50..59.
__device__ __forceinline__ uint64_t open_entry(uint64_t rec, int zl, int xl) {
  return rec | ((uint64_t)((zl << 6) | xl) << 50);
}
__device__ __forceinline__ int open_slice(uint64_t e) { return (int)((e >> 56) & 15u); }
__device__ __forceinline__ int open_column(uint64_t e) { return (int)((e >> 50) & 63u); }

// direction code d = (dz+1) << 4 | (dy+1) << 2 | (dx+1)
__device__ __forceinline__ void dir_delta(int d, int& dx, int& dy, int& dz) {
  dz = (d >> 4) - 1;
  dy = ((d >> 2) & 3) - 1;
  dx = (d & 3) - 1;
}

// Three 512-thread blocks per CU need 6 waves per SIMD, i.e. at most 80 VGPRs.
// M24: nz * ny and ny * nx are below 2^24, so every product in a flat-index computation has
// 24-bit factors and uses the full-rate 24-bit multiply (a 32-bit multiply issues at a quarter
// of the VALU rate).
template <bool M24>
__global__ __launch_bounds__(64 * TILE_Y) __attribute__((amdgpu_waves_per_eu(6)))
void k_tile(const float* __restrict__ sdf, int nx, int ny, int nz, float thr, TileOut o) {
  // keys of the haloed slice (T_LY x T_LX), then the x-triple maxima of its rows (T_LY x 64)
  __shared__ uint64_t KB[T_LY * T_LX + T_LY * 64];
  __shared__ uint8_t XD[T_LY][64];  // the x offset (0..2) of each row maximum
  uint64_t* const K = KB;
  uint64_t* const XM = KB + T_LY * T_LX;
  __shared__ int LAB[T_VOX];
  __shared__ uint8_t DIR[T_VOX];
  __shared__ uint64_t MASK[TILE_Z * TILE_Y];  // marker mask of each tile row
  const int lane = threadIdx.x, wave = threadIdx.y;
  const int tid = wave * 64 + lane;
  const int x0 = blockIdx.x * 64, y0 = blockIdx.y * TILE_Y, z0 = blockIdx.z * TILE_Z;
  const int x = x0 + lane, y = y0 + wave;
  const int nzt = min(TILE_Z, nz - z0);
  const int xlast = min(x0 + 63, nx - 1), ylast = min(y0 + TILE_Y - 1, ny - 1);
  const bool col_in = (x < nx) && (y < ny);
  const int64_t seg_id =
      (((int64_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x) * TILE_Y + wave;
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  const int plane_n = ny * nx;
  unsigned nop = 0;
  auto mul = [](int a, int b) -> int {
    return M24 ? (int)__umul24((unsigned)a, (unsigned)b) : a * b;
  };
  const int tbase = (z0 * ny + y0) * nx + x0;  // flat index of the tile's first voxel
  // flat index of this wave's row in slice 0 (wave-uniform: a scalar register)
  const int srow = __builtin_amdgcn_readfirstlane(tbase + wave * nx);
  auto gidx = [&](int l) {
    return tbase + mul(l >> 9, plane_n) + mul((l >> 6) & 7, nx) + (l & 63);
  };
  auto flat = [&](int gz, int gy, int gx) { return mul(mul(gz, ny) + gy, nx) + gx; };

  // This thread's two halo-slice elements (e = tid, tid + 512 of the 10 x 66 slice): in-plane
  // offsets of their wrapped (x, y), fixed for the whole march.
  constexpr int kSliceElems = T_LX * T_LY;
  const bool has1 = tid + 64 * TILE_Y < kSliceElems;
  int off0, off1 = 0;
  {
    // coordinates lie in [-1, 2N) when the volume is at least a haloed slice wide and deep
    const bool near = nx >= T_LX && ny >= T_LY;
    auto wrap = [&](int v, int n) { return near ? wrap_near(v, n) : wrap_any(v, n); };
    const int ly = tid / T_LX, lx = tid - ly * T_LX;
    off0 = mul(wrap(y0 - 1 + ly, ny), nx) + wrap(x0 - 1 + lx, nx);
    if (has1) {
      const int e = tid + 64 * TILE_Y;
      const int ly1 = e / T_LX, lx1 = e - ly1 * T_LX;
      off1 = mul(wrap(y0 - 1 + ly1, ny), nx) + wrap(x0 - 1 + lx1, nx);
    }
  }
  // Loads are unconditional (a thread without a second element re-reads its first; slices past
  // the tile re-read its last, an L2 hit), so no branch makes the compiler drain the loads in
  // flight.
  if (!has1) off1 = off0;
  auto load_slice = [&](int zg, float& a, float& b) {
    const float* __restrict__ plane =
        sdf + (int64_t)wrap_near(min(zg, z0 + nzt), nz) * plane_n;
    a = plane[off0];
    b = plane[off1];
  };
  // 3x3 (x, y) key maximum of this thread's column at slice zg (values a, b already loaded),
  // with the winner's (dy, dx) as dy << 2 | dx, and the centre key. Separable: the x-triple
  // maximum of each of the slice's 10 rows is computed once (1.25 per thread) and shared
  // through LDS, then each thread takes the maximum of three of them: 4.5 compare-selects per
  // voxel instead of 8, at the price of a second barrier. The first position wins among equal keys (a voxel
  // met twice through a periodic wrap), in (dy, dx) row-major order.
  auto slice_max = [&](int zg, float a, float b, uint64_t& centre, int& code) -> uint64_t {
    const uint32_t zbase = (uint32_t)wrap_near(zg, nz) * (uint32_t)plane_n;
    K[tid] = lex_key(a, zbase + (uint32_t)off0);
    if (has1) K[tid + 64 * TILE_Y] = lex_key(b, zbase + (uint32_t)off1);
    __syncthreads();
    // rows wave and (waves 0, 1) TILE_Y + wave: the second element is a wave-uniform branch
    static_assert(T_LY * 64 <= 2 * 64 * TILE_Y, "two x-triples per thread at most");
    auto xtriple = [&](int ly) {
      const uint64_t k0 = K[ly * T_LX + lane], k1 = K[ly * T_LX + lane + 1],
                     k2 = K[ly * T_LX + lane + 2];
      const bool g1 = k1 > k0;
      const uint64_t m1 = g1 ? k1 : k0;
      const bool g2 = k2 > m1;
      XM[ly * 64 + lane] = g2 ? k2 : m1;
      XD[ly][lane] = (uint8_t)(g2 ? 2 : (g1 ? 1 : 0));
    };
    xtriple(wave);
    if (wave < T_LY - TILE_Y) xtriple(TILE_Y + wave);
    centre = K[(wave + 1) * T_LX + lane + 1];
    __syncthreads();
    const uint64_t m0 = XM[wave * 64 + lane], m1 = XM[(wave + 1) * 64 + lane],
                   m2 = XM[(wave + 2) * 64 + lane];
    const bool g1 = m1 > m0;
    const uint64_t n1 = g1 ? m1 : m0;
    const bool g2 = m2 > n1;
    const int dy = g2 ? 2 : (g1 ? 1 : 0);
    code = (dy << 2) | XD[wave + dy][lane];
    return g2 ? m2 : n1;
  };

  // ---- march: slices z0-1 .. z0+nzt, three slices of loads in flight ----
  // Slice k is z0 - 1 + k. The buffers A, B, C are used in rotation without copies, so each
  // load stays in flight until the slice that needs it.
  const int last_k = nzt + 1;
  float A0, A1 = 0.f, B0, B1 = 0.f, C0 = 0.f, C1 = 0.f;
  load_slice(z0 - 1, A0, A1);
  load_slice(z0, B0, B1);
  load_slice(z0 + 1, C0, C1);
  uint64_t c_prev, c_cur, c_next;
  int d_prev, d_cur, d_next;
  uint64_t m_prev = slice_max(z0 - 1, A0, A1, c_prev, d_prev);
  load_slice(z0 + 2, A0, A1);
  uint64_t m_cur = slice_max(z0, B0, B1, c_cur, d_cur);
  load_slice(z0 + 3, B0, B1);
  // this thread's sdf>0 slices, peak slices and sdf == 0 slices
  uint32_t pore_mask = 0, peak_mask = 0, zero_mask = 0;
  // one slice of the march: needs slice k = zl + 2 (values va, vb)
  auto march = [&](int zl, float va, float vb) {
    const int z = z0 + zl;
    const uint64_t m_next = slice_max(z + 1, va, vb, c_next, d_next);
    const int l = (zl * TILE_Y + wave) * 64 + lane;
    const int i = srow + lane + mul(zl, plane_n);
    const float cv = key_value(c_cur);
    bool is_peak = false;
    uint8_t dir = kNoDir;
    if (col_in && !(cv <= 0.0f)) {
      if (cv != cv) {
        is_peak = true;  // nothing beats a NaN centre
      } else {
        uint64_t m = m_prev;
        int dz = 0, cd = d_prev;
        if (m_cur > m) { m = m_cur; dz = 1; cd = d_cur; }
        if (m_next > m) { m = m_next; dz = 2; cd = d_next; }
        is_peak = (int)(uint32_t)m == i;
        if (cv > 0.0f) {
          dir = is_peak ? (uint8_t)kSelf : (uint8_t)((dz << 4) | cd);
          pore_mask |= 1u << zl;
        }
      }
    }
    if (col_in && cv == 0.0f) zero_mask |= 1u << zl;
    const bool mk = col_in && (cv < thr);
    const uint64_t mm = __ballot(mk);
    const uint64_t starts = mm & ~(mm << 1);
    const uint64_t smask = starts & (below | (1ull << lane));
    const int st = smask ? 63 - __clzll((long long)smask) : lane;
    LAB[l] = mk ? l - (lane - st) : -1;
    if (lane == 0) MASK[zl * TILE_Y + wave] = mm;
    DIR[l] = dir;
    if (is_peak) peak_mask |= 1u << zl;
    (void)i;
    m_prev = m_cur;
    m_cur = m_next;
    d_prev = d_cur;
    d_cur = d_next;
    c_cur = c_next;
  };
  for (int zl = 0; zl < nzt; zl += 3) {  // slices zl+2 in C, zl+3 in A, zl+4 in B
    march(zl, C0, C1);
    load_slice(z0 + zl + 4, C0, C1);
    if (zl + 1 >= nzt) break;
    march(zl + 1, A0, A1);
    load_slice(z0 + zl + 5, A0, A1);
    if (zl + 2 >= nzt) break;
    march(zl + 2, B0, B1);
    load_slice(z0 + zl + 6, B0, B1);
  }
  (void)last_k;
  __syncthreads();
  // peak masks: one word per (row, 64-column chunk), rows in flat order, so the set bits of the
  // whole map in word order are the peaks in ascending voxel order
  {
    // word (z, y, chunk): 32-bit index math (the map has fewer than 2^31 words), uniform per wave
    const int pstep = ny * (int)gridDim.x;
    int pw = (z0 * ny + y) * (int)gridDim.x + (int)blockIdx.x;
#pragma unroll
    for (int zl = 0; zl < TILE_Z; ++zl, pw += pstep) {
      const uint64_t pm = __ballot((peak_mask >> zl) & 1u);
      if (lane == 0 && zl < nzt && y < ny) o.peak_map[pw] = pm;
    }
  }

  // ---- CCL inside the tile: forward rows (y+1,z), (y-1,z+1), (y,z+1), (y+1,z+1) ----
  // Lane j of wave y takes slice j/4 and forward row j%4 of its row: all 64 row pairs of a wave
  // run side by side.
  {
    const int zl = lane >> 2, r = lane & 3;
    const bool zp_ok = zl + 1 < nzt, yp_ok = wave + 1 < TILE_Y && y + 1 < ny,
               ym_ok = wave >= 1;
    const int yb = r == 0 ? wave + 1 : (r == 1 ? wave - 1 : (r == 2 ? wave : wave + 1));
    const int zb = r == 0 ? zl : zl + 1;
    const bool ok = y < ny && zl < nzt &&
                    (r == 0 ? yp_ok : (r == 1 ? ym_ok && zp_ok : (r == 2 ? zp_ok : yp_ok && zp_ok)));
    if (ok) {
      const uint64_t ma = MASK[zl * TILE_Y + wave];
      const uint64_t mb = MASK[zb * TILE_Y + yb];
      if (ma && mb) unite_row_runs(LAB, ma, mb, (zl * TILE_Y + wave) * 64, (zb * TILE_Y + yb) * 64);
    }
  }
  __syncthreads();
  // flatten: each thread chases the roots of its 16 voxels side by side (LDS reads of all 16
  // chains in flight together), reading only: the forest stays as the unions left it, so no
  // store can race with a chase (a find with path halving would write ancestors into other
  // voxels' entries, and such a write landing after that voxel's owner stored its root would
  // leave an intermediate node as its label). The results go straight out: the label, the
  // solid first-appearance seed of a component root (a component's root is its smallest
  // marker, one of the tile-local roots, seeded before any flood can lower it), the staged
  // tile-local roots (k_lroot_flatten resolves them after the seams) and the x-face labels of
  // the row chunk (first and last column; entry (chunk, z, y), compact for the seam pass).
  {
    int r[TILE_Z], p[TILE_Z];
#pragma unroll
    for (int zl = 0; zl < TILE_Z; ++zl) r[zl] = zl < nzt ? LAB[(zl * TILE_Y + wave) * 64 + lane] : -1;
#pragma unroll
    for (int zl = 0; zl < TILE_Z; ++zl) p[zl] = LAB[r[zl] >= 0 ? r[zl] : 0];
    bool more = true;
    while (more) {
      more = false;
#pragma unroll
      for (int zl = 0; zl < TILE_Z; ++zl) {
        const bool go = r[zl] >= 0 && p[zl] != r[zl];
        r[zl] = go ? p[zl] : r[zl];
        more |= go;
      }
      if (!more) break;
#pragma unroll
      for (int zl = 0; zl < TILE_Z; ++zl) p[zl] = LAB[r[zl] >= 0 ? r[zl] : 0];
    }
    unsigned nr = 0;
    uint32_t* const lstage = o.lroot_stage + seg_id * T_SEG;
    // x-face entry (chunk, z, y): 32-bit index math (fewer than 2^31 entries)
    int xw = ((int)blockIdx.x * nz + z0) * ny + y;
#pragma unroll
    for (int zl = 0; zl < TILE_Z; ++zl, xw += ny) {
      if (zl >= nzt) break;
      const int l = (zl * TILE_Y + wave) * 64 + lane;
      const int i = srow + lane + mul(zl, plane_n);  // = gidx(l)
      if (col_in) {
        // a voxel that is no marker: -1 if it is a flood candidate (!(sdf >= 0)), else -2
        // sdf >= 0: an sdf>0 slice or an sdf == 0 slice
        o.labels[i] = r[zl] != -1 ? gidx(r[zl]) : ((((pore_mask | zero_mask) >> zl) & 1u) ? -2 : -1);
        if (r[zl] == l) o.min_solid[i] = i;
        if ((zero_mask >> zl) & 1u) o.segsrc[i] = -1;  // never a pore, never flooded
      }
      const bool lroot = col_in && r[zl] == l;
      const uint64_t rm = __ballot(lroot);
      if (lroot) lstage[nr + lanes_below(rm)] = (uint32_t)i;
      nr += (unsigned)__popcll(rm);
      // the chunk's first and last column (uniform lanes: read into scalar registers)
      const int a0 = __builtin_amdgcn_readlane(r[zl], 0);
      const int aL = __builtin_amdgcn_readlane(r[zl], xlast - x0);
      if (lane == 0 && y < ny) {
        int2 v;
        v.x = a0 == -1 ? -1 : gidx(a0);
        v.y = aL == -1 ? -1 : gidx(aL);
        o.xface[xw] = v;
      }
    }
    if (lane == 0) o.lroot_cnt[seg_id] = nr;
  }

  // ---- ascent walks inside the tile: pointer jumping on LAB (now the chase state) ----
  // The chase state goes into the LAB words of sdf>0 voxels, which no marker chase above reads
  // (marker labels point at markers only), so no barrier is needed before it.
  auto in_tile = [&](int ax, int ay, int az) {
    return ax >= 0 && ax < 64 && x0 + ax < nx && ay >= 0 && ay < TILE_Y && y0 + ay < ny &&
           az >= 0 && az < nzt;
  };
  // Chase state of sdf>0 voxel l (slice zl, column xl of this wave's row): its successor inside
  // the tile (ACT), a fixed point (FIX), or a successor outside the tile (EXIT).
  auto walk_init = [&](int l, int zl, int xl) -> bool {
    const int d = DIR[l] & kDirMask;
    uint32_t w;
    bool active = false;
    if (d == kSelf) {
      w = cs_pack(l, 0, CS_FIX);
    } else {
      int dx, dy, dz;
      dir_delta(d, dx, dy, dz);
      const int ax = xl + dx, ay = wave + dy, az = zl + dz;
      if (in_tile(ax, ay, az)) {
        w = cs_pack((az * TILE_Y + ay) * 64 + ax, 1, CS_ACT);
        active = true;
      } else {
        w = cs_pack(l, 0, CS_EXIT);
      }
    }
    LAB[l] = (int)w;
    return active;
  };
  // The wave's sdf>0 voxels as a list in the key buffers (free after the march), in (slice, lane)
  // order: the walk initialisation takes ceil(count / 64) turns instead of 16 slices, and keeps
  // in place only the active walks, which the pointer jumping then compacts after every round
  // (a round costs ceil(active / 64) turns instead of the busiest lane's count of active
  // slices). A wave with more sdf>0 voxels than its share of the buffers keeps per-lane loops
  // over its bit masks.
  constexpr int ALIST_W = (int)(sizeof(KB) / sizeof(uint16_t)) / TILE_Y;
  uint16_t* alist = reinterpret_cast<uint16_t*>(KB) + wave * ALIST_W;
  unsigned na = 0;
#pragma unroll
  for (int zl = 0; zl < TILE_Z; ++zl) {
    const bool pb = (pore_mask >> zl) & 1u;
    const uint64_t pm = __ballot(pb);
    const unsigned pos = na + lanes_below(pm);
    if (pb && pos < (unsigned)ALIST_W) alist[pos] = (uint16_t)((zl * TILE_Y + wave) * 64 + lane);
    na += (unsigned)__popcll(pm);
  }
  const bool listed = na <= (unsigned)ALIST_W;  // wave-uniform
  uint32_t act = 0;
  if (listed) {
    unsigned nk = 0;
    for (unsigned j0 = 0; j0 < na; j0 += 64) {
      // every lane reads its entry before any active walk is written back below j0 + 64
      const unsigned j = j0 + lane;
      const int l = j < na ? (int)alist[j] : 0;
      const bool active = j < na && walk_init(l, l >> 9, l & 63);
      const uint64_t am = __ballot(active);
      if (active) alist[nk + lanes_below(am)] = (uint16_t)l;
      nk += (unsigned)__popcll(am);
    }
    na = nk;
  } else {
    for (uint32_t m = pore_mask; m; m &= m - 1) {
      const int zl = __ffs(m) - 1;
      if (walk_init((zl * TILE_Y + wave) * 64 + lane, zl, lane)) act |= 1u << zl;
    }
  }
  // One jump of active walk l (only its owner writes LAB[l]); returns whether it stays active.
  auto jump = [&](int l) -> bool {
    const uint32_t w = (uint32_t)LAB[l];
    const uint32_t wv = (uint32_t)LAB[cs_u(w)];
    const int s2 = cs_s(w) + cs_s(wv);
    const uint32_t sv = cs_st(wv);
    if (sv == CS_LONG || s2 > kMaxSteps) {
      LAB[l] = (int)cs_pack(cs_u(w), cs_s(w), CS_LONG);
      return false;
    }
    LAB[l] = (int)cs_pack(cs_u(wv), s2, sv);
    return sv == CS_ACT;
  };
  // "any walk still active" per round with one barrier: the waves' flags of round r go to
  // s_more[r % 3], and thread 0 clears the slot of round r + 1, last read in round r - 2, before
  // the barrier of round r (every wave has passed that of round r - 1, so those reads are done)
  __shared__ unsigned s_more[3];
  if (tid < 3) s_more[tid] = 0u;
  __syncthreads();
  for (int round = 0;; ++round) {
    bool any;
    if (listed) {
      unsigned nk = 0;
      for (unsigned j0 = 0; j0 < na; j0 += 64) {
        // every lane reads its entry before any survivor is written back below j0 + 64
        const unsigned j = j0 + lane;
        const int l = j < na ? (int)alist[j] : 0;
        const bool still = j < na && jump(l);
        const uint64_t sm = __ballot(still);
        if (still) alist[nk + lanes_below(sm)] = (uint16_t)l;
        nk += (unsigned)__popcll(sm);
      }
      na = nk;
      any = na != 0;
    } else {
      for (uint32_t m = act; m; m &= m - 1) {
        const int zl = __ffs(m) - 1;
        if (!jump((zl * TILE_Y + wave) * 64 + lane)) act &= ~(1u << zl);
      }
      any = act != 0;
    }
    const int slot = round % 3;
    if (__any(any) && lane == 0) s_more[slot] = 1u;
    if (tid == 0) s_more[slot == 2 ? 0 : slot + 1] = 0u;
    __syncthreads();
    if (s_more[slot] == 0u) break;
  }

  // ---- walk outputs; LAB at each fixed point collects the smallest local index whose walk
  //      ends there (every other LAB word read below is read by the one lane handling it) ----
  bool capped_any = false;
  // Output of the walk of sdf>0 voxel l (slice zl, column xl of this wave's row); returns
  // whether the walk is still open.
  auto walk_out = [&](int l, int zl, int xl, uint64_t& rec) -> bool {
    const int xg = x0 + xl;
    const int i = srow + xl + mul(zl, plane_n);
    int target, steps, fin_u = -1;
    bool fin = false, capped = false;
    const bool root_here = (DIR[l] & kDirMask) == kSelf;
    const uint32_t w = root_here ? cs_pack(l, 0, CS_FIX) : (uint32_t)LAB[l];
    const int u = cs_u(w);
    steps = cs_s(w);
    if (cs_st(w) == CS_FIX) {
      fin = true;
      fin_u = u;
      target = gidx(u);
    } else if (cs_st(w) == CS_EXIT) {
      if (steps == kMaxSteps) {  // the 512th step lands on u, whose successor is outside
        fin = true;
        capped = true;
        target = gidx(u);
      } else {
        int dx, dy, dz;
        dir_delta(DIR[u] & kDirMask, dx, dy, dz);
        const int gx = wrap_near(x0 + (u & 63) + dx, nx);
        const int gy = wrap_near(y0 + ((u >> 6) & 7) + dy, ny);
        const int gz = wrap_near(z0 + (u >> 9) + dz, nz);
        target = flat(gz, gy, gx);
        steps += 1;
      }
    } else {  // CS_LONG: walk the 512 steps one by one
      int cx = xl, cy = wave, cz = zl, cl = l;
      steps = 0;
      while (true) {
        const int d = DIR[cl] & kDirMask;
        if (d == kSelf) {
          fin = true;
          fin_u = cl;
          target = gidx(cl);
          break;
        }
        int dx, dy, dz;
        dir_delta(d, dx, dy, dz);
        const int ax = cx + dx, ay = cy + dy, az = cz + dz;
        ++steps;
        if (!in_tile(ax, ay, az)) {
          target = flat(wrap_near(z0 + az, nz), wrap_near(y0 + ay, ny), wrap_near(x0 + ax, nx));
          break;
        }
        cx = ax;
        cy = ay;
        cz = az;
        cl = (az * TILE_Y + ay) * 64 + ax;
        if (steps == kMaxSteps) {
          fin = true;
          capped = (DIR[cl] & kDirMask) != kSelf;
          if (!capped) fin_u = cl;
          target = gidx(cl);
          break;
        }
      }
    }
    if (fin) {
      o.roots[i] = (int)((uint32_t)target | o.root_flag);
      if (!capped) atomicMin(&LAB[fin_u], l);
      capped_any |= capped;
    }
    const bool boundary = xg == x0 || xg == xlast || y == y0 || y == ylast || zl == 0 ||
                          zl == nzt - 1;
    rec = walk_record(target, steps, fin, capped);
    if (boundary) o.walk[i] = rec;
    return !fin;
  };
  // The wave's sdf>0 voxels as a compact list in (slice, lane) order, in the key buffers (free
  // after the march): the wave then takes ceil(count / 64) turns instead of its busiest lane's
  // count of slices. A wave with more voxels than its share of the buffers uses the per-lane
  // loop. Open walks are staged with their walk record (open_entry), so the resolver needs no
  // record load of its own and no record is written for an open walk off the tile boundary.
  uint64_t* const ostage = o.open_stage + seg_id * T_SEG;
  constexpr int PLIST_W = (int)(sizeof(KB) / sizeof(uint16_t)) / TILE_Y;
  uint16_t* plist = reinterpret_cast<uint16_t*>(KB) + wave * PLIST_W;
  unsigned np = 0;
#pragma unroll
  for (int zl = 0; zl < TILE_Z; ++zl) {
    const bool pb = (pore_mask >> zl) & 1u;
    const uint64_t pm = __ballot(pb);
    const unsigned pos = np + lanes_below(pm);
    if (pb && pos < (unsigned)PLIST_W) plist[pos] = (uint16_t)((zl * TILE_Y + wave) * 64 + lane);
    np += (unsigned)__popcll(pm);
  }
  if (np <= (unsigned)PLIST_W) {
    for (unsigned j0 = 0; j0 < np; j0 += 64) {
      const unsigned j = j0 + lane;
      const int li = j < np ? (int)plist[j] : 0;
      uint64_t rec = 0;
      const bool open = j < np && walk_out(li, li >> 9, li & 63, rec);
      const uint64_t om = __ballot(open);
      if (open)
        ostage[nop + lanes_below(om)] =
            open_entry(rec, li >> 9, li & 63);
      nop += (unsigned)__popcll(om);
    }
  } else {
    // wave-uniform trip count (the busiest lane's), so that nop stays uniform
    uint32_t m = pore_mask;
    while (__any(m != 0)) {
      const bool has = m != 0;
      const int zl = has ? __ffs(m) - 1 : 0;
      uint64_t rec = 0;
      bool open = false;
      if (has) open = walk_out((zl * TILE_Y + wave) * 64 + lane, zl, lane, rec);
      const uint64_t om = __ballot(open);
      if (open) ostage[nop + lanes_below(om)] = open_entry(rec, zl, lane);
      nop += (unsigned)__popcll(om);
      if (has) m &= m - 1;
    }
  }
  if (__any(capped_any) && lane == 0) atomicOr(o.cap_flag, 1ull);
  __syncthreads();
  // roots of this thread's column (sdf>0 peaks): min_pore = the smallest walker's voxel
  for (uint32_t m = pore_mask & peak_mask; m; m &= m - 1) {
    const int zl = __ffs(m) - 1;
    const int l = (zl * TILE_Y + wave) * 64 + lane;
    o.min_pore[gidx(l)] = gidx(LAB[l]);
  }
  if (lane == 0) o.open_cnt[seg_id] = nop;
}

// ---- cross-tile and wrapping marker relations ----
// After k_tile, the relations left are those between voxels of different tiles or across the
// periodic wrap. For a row chunk (64 voxels of one (y, z) row) and its forward rows (y+1, z),
// (y-1, z+1), (y, z+1), (y+1, z+1): a forward row in another tile (or reached by wrapping)
// needs the full run-union rule; one in the same tile only the two contacts the tile skipped
// (the chunk's first column with x-1, its last column with x+1); and the +x relation at the
// chunk start is always cross-tile. Only rows at a tile boundary in y or z have crossing
// forward rows: y % TILE_Y in {0, TILE_Y-1}, y = ny-1, z % TILE_Z = TILE_Z-1, z = nz-1.

// Forward rows of row (y, z): (y+1, z), (y-1, z+1), (y, z+1), (y+1, z+1) as (y', z'), and
// whether each lies in the same tile as the row.
__device__ __forceinline__ void seam_rows(int y, int z, int ny, int nz, int fy[4], int fz[4],
                                          bool same[4]) {
  const int yp = wp1(y, ny), ym = wm1(y, ny), zp = wp1(z, nz);
  const bool ty_p = (y + 1 < ny) && ((y + 1) / TILE_Y == y / TILE_Y);
  const bool ty_m = (y >= 1) && ((y - 1) / TILE_Y == y / TILE_Y);
  const bool tz_p = (z + 1 < nz) && ((z + 1) / TILE_Z == z / TILE_Z);
  fy[0] = yp;
  fz[0] = z;
  fy[1] = ym;
  fz[1] = zp;
  fy[2] = y;
  fz[2] = zp;
  fy[3] = yp;
  fz[3] = zp;
  same[0] = ty_p;
  same[1] = ty_m && tz_p;
  same[2] = tz_p;
  same[3] = ty_p && tz_p;
}
__device__ __forceinline__ int z_boundary_slice(int j, int nz) {  // j < ceil(nz / TILE_Z)
  const int z = j * TILE_Z + TILE_Z - 1;
  return z < nz ? z : nz - 1;
}

// Per-block set of the tile-local root pairs a seam pass has to unite: one seam crosses the
// same pair of grain cores in many rows, so each distinct pair is united once, after the block
// has collected them. A pair that finds no slot is united at once.
constexpr int SEAM_TABLE = 512;
constexpr int SEAM_PROBES = 16;
constexpr unsigned long long kNoPair = ~0ull;
__device__ __forceinline__ void seam_pair(unsigned long long* tab, int* L, int a, int b) {
  if (a == b) return;
  const uint32_t lo = (uint32_t)min(a, b), hi = (uint32_t)max(a, b);
  const unsigned long long key = ((unsigned long long)lo << 32) | hi;
  uint32_t h = (lo * 0x9E3779B1u ^ hi * 0x85EBCA77u) >> 23;  // 9 bits: SEAM_TABLE slots
  for (int p = 0; p < SEAM_PROBES; ++p) {
    const unsigned long long cur = tab[h];
    if (cur == key) return;
    if (cur == kNoPair) {
      const unsigned long long prv = atomicCAS(&tab[h], kNoPair, key);
      if (prv == kNoPair || prv == key) return;
    }
    h = (h + 1) & (SEAM_TABLE - 1);
  }
  uf_unite(L, (int)lo, (int)hi);
}
// Union of a and b (a < b) whose parents pa = L[a], pb = L[b] were read beforehand: when both
// were roots the union is one atomic, else the general union.
__device__ __forceinline__ void seam_unite_from(int* L, int a, int b, int pa, int pb, int old) {
  if (pa == a && pb == b) {
    if (old != b) uf_unite(L, a, old);  // b was hooked meanwhile: a joins what b joined
    return;
  }
  uf_unite(L, a, b);
}
// The block's collected pairs (two slots per thread of the 256), united with all first-parent
// reads and then all root hooks of a thread in flight together.
__device__ __forceinline__ void seam_flush(unsigned long long* tab, int* L) {
  static_assert(SEAM_TABLE == 2 * 256, "two table slots per thread");
  __syncthreads();
  const int t = threadIdx.x;
  const unsigned long long k0 = tab[t], k1 = tab[t + 256];
  const bool v0 = k0 != kNoPair, v1 = k1 != kNoPair;
  const int a0 = v0 ? (int)(k0 >> 32) : 0, b0 = v0 ? (int)(uint32_t)k0 : 0;
  const int a1 = v1 ? (int)(k1 >> 32) : 0, b1 = v1 ? (int)(uint32_t)k1 : 0;
  const int pa0 = L[a0], pb0 = L[b0], pa1 = L[a1], pb1 = L[b1];
  const bool r0 = v0 && pa0 == a0 && pb0 == b0, r1 = v1 && pa1 == a1 && pb1 == b1;
  int o0 = b0, o1 = b1;
  if (r0) o0 = atomicMin(L + b0, a0);
  if (r1) o1 = atomicMin(L + b1, a1);
  if (v0) seam_unite_from(L, a0, b0, pa0, pb0, o0);
  if (v1) seam_unite_from(L, a1, b1, pa1, pb1, o1);
}

// A wave's contacts, listed by ballot in LDS and put through the pair set 64 at a time, so that
// the set's probing runs on full waves and has one call site.
constexpr int SEAM_LIST = 256;  // per-wave list of contacts
struct SeamList {
  int2* pl;
  unsigned np;
  int lane;
  __device__ __forceinline__ void add(bool v, int a, int b) {
    const uint64_t vm = __ballot(v);
    if (v) pl[np + lanes_below(vm)] = make_int2(a, b);
    np += (unsigned)__popcll(vm);
  }
  __device__ __forceinline__ void flush(unsigned long long* tab, int* L) {
    __builtin_amdgcn_wave_barrier();
    for (unsigned j = lane; j < np; j += 64) {
      const int2 q = pl[j];
      seam_pair(tab, L, q.x, q.y);
    }
    __builtin_amdgcn_wave_barrier();
    np = 0;
  }
};
static_assert(SEAM_LIST >= 4 * 64, "room for one row pair's contacts after the flush test");

// Contacts of row P with an adjacent row Q (the next or previous row in y, z or both) in one row
// chunk: each P marker with the Q voxel of its column, and where that is no marker a run's first
// and last P voxel with the Q voxels just before and after the run (interior voxels of a run
// reach those through their own same-column contacts; a run has one label in its chunk). my / c:
// this lane's P / Q label word (a marker's label >= 0, a negative code otherwise); m: P's marker
// ballot; qe: Q's label just outside the chunk (lane 0: column x0-1, the chunk's last column:
// xlast+1); qrow: Q's flat row offset (a one-column chunk loads Q's xlast+1 on lane 0 from it).
__device__ __forceinline__ void seam_contacts(SeamList& sl, unsigned long long* tab, int* L,
                                              int my, uint64_t m, int c, int qe, int qrow, int x,
                                              int nx, bool last_col, int lane) {
  const bool mk = my >= 0;
  const bool run_start = mk && (lane == 0 || !((m >> (lane - 1)) & 1ull));
  const bool run_end = mk && (last_col || !((m >> (lane + 1)) & 1ull));
  const bool cm = c >= 0;
  const int prev_c = __shfl_up(c, 1, 64);
  const int next_c = __shfl_down(c, 1, 64);
  const bool dup = lane > 0 && ((m >> (lane - 1)) & 1ull) && prev_c == c;
  sl.add(mk && cm && !dup, my, c);
  int a = -1, b = -1;
  if (mk && !cm && run_start) a = lane == 0 ? qe : prev_c;
  if (mk && !cm && run_end) b = !last_col ? next_c : (lane != 0 ? qe : L[qrow + wp1(x, nx)]);
  sl.add(a >= 0, my, a);
  sl.add(b >= 0, my, b);
  if (sl.np > (unsigned)(SEAM_LIST - 3 * 64)) sl.flush(tab, L);  // wave-uniform
}

// y seams. Seam s joins the last row ya of tile row s-1 and the first row yb of tile row s (s = 0:
// the volume's last row and row 0, across the periodic wrap). With A = (ya, z) and B = (yb, z),
// the relations A(z)-B(z), A(z)-B(z+1) and B(z)-A(z+1) are all those that cross a y tile
// boundary. A wave walks SEAM_ZR / 4 consecutive slices of the seam with the next slice's two rows
// loaded ahead, so that every row is loaded once (with its outside columns) for the three
// relations it takes part in.
constexpr int SEAM_ZR = 64;  // slices per block
__device__ __forceinline__ void seam_y_block(unsigned long long* tab, int2* list, int* L, int nx,
                                             int ny, int nz, int bx, int s, int bz, int lane,
                                             int wave) {
  const int x0 = bx * 64, x = x0 + lane;
  const bool in = x < nx;
  const int xlast = min(x0 + 63, nx - 1);
  const bool last_col = x == xlast;
  // the column outside the chunk this lane may need: lane 0 the one before the chunk, the last
  // column the one after it (a one-column chunk loads the latter when it needs it)
  const bool edge = in && (lane == 0 || last_col);
  const int xe = lane == 0 ? wm1(x, nx) : wp1(x, nx);
  const int ya = s == 0 ? ny - 1 : s * TILE_Y - 1, yb = s == 0 ? 0 : s * TILE_Y;
  constexpr int ZW = SEAM_ZR / 4;
  const int zs = bz * SEAM_ZR + wave * ZW, ze = min(zs + ZW, nz);
  SeamList sl{list + wave * SEAM_LIST, 0u, lane};
  // a row's label word and outside-column label; loads unconditional (a lane without a column,
  // or without an outside column, re-reads its row's first word)
  auto load = [&](int y, int z, int& v, int& e) {
    const int rb = (z * ny + y) * nx;
    const int lv = L[in ? rb + x : rb], le = L[edge ? rb + xe : rb];
    v = in ? lv : -1;
    e = edge ? le : -1;
  };
  int va = -1, ea = -1, vb = -1, eb = -1;
  if (zs < ze) {  // wave-uniform
    load(ya, zs, va, ea);
    load(yb, zs, vb, eb);
  }
  for (int z = zs; z < ze; ++z) {
    const int z1 = wp1(z, nz);
    int va1, ea1, vb1, eb1;
    load(ya, z1, va1, ea1);
    load(yb, z1, vb1, eb1);
    const uint64_t ma = __ballot(va >= 0), mb = __ballot(vb >= 0);
    if (ma) {  // wave-uniform
      seam_contacts(sl, tab, L, va, ma, vb, eb, (z * ny + yb) * nx, x, nx, last_col, lane);
      seam_contacts(sl, tab, L, va, ma, vb1, eb1, (z1 * ny + yb) * nx, x, nx, last_col, lane);
    }
    if (mb) seam_contacts(sl, tab, L, vb, mb, va1, ea1, (z1 * ny + ya) * nx, x, nx, last_col, lane);
    va = va1;
    ea = ea1;
    vb = vb1;
    eb = eb1;
  }
  sl.flush(tab, L);
}

// z seams: every row (y, z) of a slice z at a z tile boundary (z % TILE_Z = TILE_Z-1, or the last
// slice) with its forward rows in the next slice (y-1, z+1), (y, z+1), (y+1, z+1), except those
// in another y tile row (the y seams take them). A block takes the rows of SEAM_YT y tiles, one
// row per wave at a time, two rows in flight.
constexpr int SEAM_YT = 4;  // y tiles per block
__device__ __forceinline__ void seam_z_block(unsigned long long* tab, int2* list, int* L, int nx,
                                             int ny, int nz, int bx, int by, int bz, int lane,
                                             int wave) {
  const int x0 = bx * 64, x = x0 + lane;
  const bool in = x < nx;
  const int xlast = min(x0 + 63, nx - 1);
  const bool last_col = x == xlast;
  const bool edge = in && (lane == 0 || last_col);
  const int xe = lane == 0 ? wm1(x, nx) : wp1(x, nx);
  const int z = z_boundary_slice(bz, nz), zp = wp1(z, nz);
  constexpr int NROW = SEAM_YT * TILE_Y;
  struct Row {
    int fy[3], fwd[3], fe[3], my;
    bool take[3], ok;
  };
  auto load = [&](int k, Row& R) {
    int y = by * NROW + k;
    const bool ok = k < NROW && y < ny;  // wave-uniform
    if (!ok) y = 0;
    const bool ty_p = (y + 1 < ny) && ((y + 1) / TILE_Y == y / TILE_Y);
    const bool ty_m = (y >= 1) && ((y - 1) / TILE_Y == y / TILE_Y);
    R.fy[0] = wm1(y, ny);
    R.fy[1] = y;
    R.fy[2] = wp1(y, ny);
    R.take[0] = ok && ty_m;
    R.take[1] = ok;
    R.take[2] = ok && ty_p;
    R.ok = ok;
    const int i = ok && in ? (z * ny + y) * nx + x : 0;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const int rb = (zp * ny + R.fy[r]) * nx;
      const bool skip = !R.take[r] || !in;
      R.fwd[r] = L[skip ? i : rb + x];
      R.fe[r] = L[(skip || !edge) ? i : rb + xe];
    }
    R.my = ok && in ? L[i] : -1;
  };
  SeamList sl{list + wave * SEAM_LIST, 0u, lane};
  auto pairs = [&](const Row& R) {
    if (!R.ok) return;
    const uint64_t m = __ballot(R.my >= 0);
    if (m == 0) return;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      if (!R.take[r]) continue;  // wave-uniform
      seam_contacts(sl, tab, L, R.my, m, in ? R.fwd[r] : -1, R.fe[r], (zp * ny + R.fy[r]) * nx, x,
                    nx, last_col, lane);
    }
  };
  // two rows of the wave at a time: the second row's loads are in flight while the first's
  // pairs are collected
  for (int k = wave; k < NROW; k += 8) {
    Row A, B;
    load(k, A);
    load(k + 4, B);
    pairs(A);
    pairs(B);
  }
  sl.flush(tab, L);
}

// One thread per row chunk, from the x-face labels k_tile wrote (entry (chunk, z, y): labels of
// the chunk's first and last column): the +x relation at the chunk start, and for forward rows
// in the same tile the chunk's first column with column x-1 and its last column with column
// x+1. 256 consecutive rows of one chunk per block share a pair set.
__device__ __forceinline__ void seam_edges_block(unsigned long long* tab, int* L,
                                                 const int2* __restrict__ xf, int ny, int nz,
                                                 int cx, int64_t blk, int tid) {
  const int64_t per_chunk = (int64_t)ny * nz;
  const int64_t t = blk * 256 + tid;
  if (t >= (int64_t)cx * per_chunk) return;
  const int bx = (int)(t / per_chunk);
  const int r = (int)(t - (int64_t)bx * per_chunk);
  const int z = r / ny, y = r - z * ny;
  const int bxm = bx > 0 ? bx - 1 : cx - 1, bxp = bx + 1 < cx ? bx + 1 : 0;
  auto face = [&](int b, int zz, int yy) { return xf[((int64_t)b * nz + zz) * ny + yy]; };
  const int2 me = face(bx, z, y);
  if (me.x == -1 && me.y == -1) return;
  if (me.x != -1) {
    const int left = face(bxm, z, y).y;
    if (left != -1) seam_pair(tab, L, me.x, left);
  }
  int fy[4], fz[4];
  bool same[4];
  seam_rows(y, z, ny, nz, fy, fz, same);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (!same[q]) continue;
    const int2 f = face(bx, fz[q], fy[q]);
    if (me.x != -1 && f.x == -1) {
      const int a = face(bxm, fz[q], fy[q]).y;
      if (a != -1) seam_pair(tab, L, me.x, a);
    }
    if (me.y != -1 && f.y == -1) {
      const int b = face(bxp, fz[q], fy[q]).x;
      if (b != -1) seam_pair(tab, L, me.y, b);
    }
  }
}

// All cross-tile and wrapping marker relations in one launch: blocks [0, n0) take the y seams,
// [n0, n0 + n1) the z seams, the rest the x-face relations.
__global__ __launch_bounds__(256) void k_ccl_seams(int* L, const int2* __restrict__ xf, int nx,
                                                   int ny, int nz, int cx, int ysel, int zsel,
                                                   int64_t n0, int64_t n1) {
  __shared__ unsigned long long tab[SEAM_TABLE];
  __shared__ int2 list[4 * SEAM_LIST];
  const int tid = threadIdx.x;
  for (int s = tid; s < SEAM_TABLE; s += 256) tab[s] = kNoPair;
  __syncthreads();
  const int64_t b = blockIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  if (b < n0) {  // grid (cx, ysel = ceil(ny / TILE_Y) seams, ceil(nz / SEAM_ZR))
    const int bx = (int)(b % cx), rest = (int)(b / cx);
    seam_y_block(tab, list, L, nx, ny, nz, bx, rest % ysel, rest / ysel, lane, wave);
  } else if (b < n0 + n1) {  // grid (cx, ceil(ny / (SEAM_YT * TILE_Y)), zsel)
    const int64_t bb = b - n0;
    const int nyt = (ny + SEAM_YT * TILE_Y - 1) / (SEAM_YT * TILE_Y);
    const int bx = (int)(bb % cx), rest = (int)(bb / cx);
    seam_z_block(tab, list, L, nx, ny, nz, bx, rest % nyt, rest / nyt, lane, wave);
  } else {
    seam_edges_block(tab, L, xf, ny, nz, cx, b - n0 - n1, tid);
  }
  seam_flush(tab, L);
}

// ---- tile-local roots -> component roots ----
// After the seams, every tile-local root points at its component root, so a marker's component
// is L[L[i]]. A wave takes LRF_SEGS consecutive staging segments of k_tile (a tile row has only a
// few dozen local roots) and chases their entries as one list, 64 at a time.
constexpr int LRF_SEGS = 4;
__global__ __launch_bounds__(256) void k_lroot_flatten(int* L, const uint32_t* __restrict__ stage,
                                                       const uint32_t* __restrict__ cnt,
                                                       int64_t nseg) {
  const int64_t seg0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * LRF_SEGS;
  if (seg0 >= nseg) return;  // wave-uniform; the kernel has no block barrier
  const int lane = threadIdx.x & 63;
  unsigned c[LRF_SEGS], total = 0;
#pragma unroll
  for (int s = 0; s < LRF_SEGS; ++s) {
    c[s] = seg0 + s < nseg ? cnt[seg0 + s] : 0u;
    total += c[s];
  }
  for (unsigned j = lane; j < total; j += 64) {
    // entry j of the concatenated segments
    unsigned jj = j;
    int64_t seg = seg0;
#pragma unroll
    for (int s = 0; s < LRF_SEGS - 1; ++s)
      if (seg == seg0 + s && jj >= c[s]) {
        jj -= c[s];
        seg = seg0 + s + 1;
      }
    // read-only chase, one final store: a path-halving store here could overwrite the root
    // another thread has just stored for its own local root with a lesser ancestor
    const int r = (int)stage[seg * T_SEG + jj];
    int x = r, p = L[r];
    while (p != x) {
      x = p;
      p = L[x];
    }
    L[r] = x;
  }
}

void launch_lroot_flatten(Ctx* c, Vol v, int32_t* labels, const uint32_t* stage,
                          const uint32_t* cnt) {
  const int64_t nseg = tile_segments(v);
  PNM_LAUNCH(c, "k_lroot_flatten", k_lroot_flatten, dim3(cdiv(nseg, 4 * LRF_SEGS)), dim3(256), 0,
             labels, stage, cnt, nseg);
}

// ---- walks that left their tile ----
// Hops along the walk records: from the voxel a walk entered a tile through, that voxel's own
// record gives where its walk ends inside that tile (or leaves it again) and after how many
// steps, which is where the original walk goes too. When the remaining step budget ends inside a
// tile, the rest is walked step by step on the SDF. The open walks are read where k_tile staged
// them: one wave per staging segment (a tile row's open walks in (slice, lane) order). Each
// thread advances RESOLVE_WALKS walks in lockstep, so that many independent record loads are in
// flight per thread; the rare step-by-step remainders run after all hops.
constexpr int RESOLVE_WALKS = 2;
__global__ __launch_bounds__(256) void k_roots_resolve(const float* __restrict__ sdf, int nx,
                                                       int ny, int nz, int64_t n,
                                                       const uint64_t* __restrict__ stage,
                                                       const uint32_t* __restrict__ cnt,
                                                       int64_t nseg, int seg_cap,
                                                       const uint64_t* __restrict__ walk,
                                                       int32_t* __restrict__ roots,
                                                       uint32_t root_flag,
                                                       int32_t* __restrict__ min_pore,
                                                       unsigned long long* cap_flag,
                                                       UDiv dgx, UDiv dgy) {
  constexpr int K = RESOLVE_WALKS;
  const int lane = threadIdx.x & 63;
  // wave-uniform segment index (fewer than 2^31 segments), so that its decoding runs on the
  // scalar unit
  const uint32_t seg = (uint32_t)__builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
  if ((int64_t)seg >= nseg) return;  // wave-uniform; the kernel has no block barrier
  const int count = (int)cnt[seg];
  const uint64_t* __restrict__ list = stage + (int64_t)seg * seg_cap;
  // the segment's tile row: seg = ((bz * gy + by) * gx + bx) * TILE_Y + row
  const uint32_t tr = seg / TILE_Y;
  const int row = (int)(seg % TILE_Y);
  uint32_t ux, uy;
  const uint32_t rest = udiv(tr, dgx, ux);
  const uint32_t uz = udiv(rest, dgy, uy);
  const int bx = (int)ux, by = (int)uy, bz = (int)uz;
  const int plane_n = nx * ny;
  const int rowbase = ((bz * TILE_Z) * ny + by * TILE_Y + row) * nx + bx * 64;
  unsigned long long flags = 0;
  for (int base = 0; base < count; base += 64 * K) {
    int vox[K], w[K], t[K], root[K];
    uint32_t live = 0, slow = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int jj = base + k * 64 + lane;
      const uint64_t e = jj < count ? list[jj] : 0ull;  // the walk record and its voxel
      vox[k] = jj < count ? rowbase + open_slice(e) * plane_n + open_column(e) : 0;
      w[k] = walk_target(e);
      t[k] = walk_steps(e);
      root[k] = -1;
      if (jj < count) live |= 1u << k;
    }
    while (live) {
      uint64_t rec[K];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        // every hop target is an sdf>0 voxel on a tile boundary and carries a record; a
        // target or record that breaks this is reported (bit 1) instead of followed
        const bool go = ((live >> k) & 1u) && (uint64_t)(uint32_t)w[k] < (uint64_t)n;
        rec[k] = walk[go ? w[k] : 0];
      }
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if (!((live >> k) & 1u)) continue;
        const uint32_t bit = 1u << k;
        if ((uint64_t)(uint32_t)w[k] >= (uint64_t)n) {
          flags |= 2ull;
          live &= ~bit;
          continue;
        }
        const int sw = walk_steps(rec[k]);
        const bool fin = walk_final(rec[k]);
        if (!fin && (sw == 0 || sw > kMaxSteps)) {
          flags |= 2ull;
          live &= ~bit;
        } else if (t[k] == kMaxSteps) {  // the budget ran out exactly on entering this tile
          root[k] = w[k];
          if (!(fin && sw == 0)) flags |= 1ull;
          live &= ~bit;
        } else if (t[k] + sw <= kMaxSteps) {
          t[k] += sw;
          if (fin) {
            root[k] = walk_target(rec[k]);
            if (walk_capped(rec[k])) flags |= 1ull;
            live &= ~bit;
          } else {
            w[k] = walk_target(rec[k]);
          }
        } else {  // fewer steps left than this tile's walk takes: the remainder, below
          slow |= bit;
          live &= ~bit;
        }
      }
    }
    for (uint32_t m = slow; m; m &= m - 1) {
      const int k = __ffs(m) - 1;
      int ww = 0, tt = 0;
#pragma unroll
      for (int q = 0; q < K; ++q)
        if (q == k) {
          ww = w[q];
          tt = t[q];
        }
      for (; tt < kMaxSteps; ++tt) {
        const int wx = ww % nx, wy = (ww / nx) % ny, wz = ww / (nx * ny);
        const int b = ascent_best(sdf, wx, wy, wz, nx, ny, nz, ww, sdf[ww]);
        if (b == ww) break;
        ww = b;
      }
      {
        const int wx = ww % nx, wy = (ww / nx) % ny, wz = ww / (nx * ny);
        if (ascent_best(sdf, wx, wy, wz, nx, ny, nz, ww, sdf[ww]) != ww) flags |= 1ull;
      }
#pragma unroll
      for (int q = 0; q < K; ++q)
        if (q == k) root[q] = ww;
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (root[k] >= 0) roots[vox[k]] = (int)((uint32_t)root[k] | root_flag);
      // of two neighbouring lanes with one root, the one with the larger voxel index skips
      const int prev = __shfl_up(root[k], 1, 64), prev_i = __shfl_up(vox[k], 1, 64);
      if (root[k] >= 0 && vox[k] < root[k] && (lane == 0 || prev != root[k] || prev_i > vox[k]))
        atomicMin(min_pore + root[k], vox[k]);
    }
  }
  if (flags) atomicOr(cap_flag, flags);
}

static dim3 tile_grid(Vol v) {
  return dim3(cdiv(v.nx, 64), cdiv(v.ny, TILE_Y), cdiv(v.nz, TILE_Z));
}

int64_t tile_segments(Vol v) {
  const dim3 g = tile_grid(v);
  return (int64_t)g.x * g.y * g.z * TILE_Y;
}
int64_t tile_segment_cap() { return T_SEG; }

void launch_tile(Ctx* c, const float* sdf, Vol v, float thr, const TileOut& o) {
  if ((int64_t)v.nz * v.ny < (1 << 24) && (int64_t)v.ny * v.nx < (1 << 24))
    PNM_LAUNCH(c, "k_tile", k_tile<true>, tile_grid(v), dim3(64, TILE_Y), 0, sdf, v.nx, v.ny, v.nz,
               thr, o);
  else
    PNM_LAUNCH(c, "k_tile", k_tile<false>, tile_grid(v), dim3(64, TILE_Y), 0, sdf, v.nx, v.ny,
               v.nz, thr, o);
}

void launch_ccl_seams(Ctx* c, Vol v, int32_t* labels, const int32_t* xface) {
  const int cx = (int)cdiv(v.nx, 64);
  const int ysel = (int)cdiv(v.ny, TILE_Y), zsel = (int)cdiv(v.nz, TILE_Z);  // y seams, z seams
  const int64_t n0 = (int64_t)cx * ysel * cdiv(v.nz, SEAM_ZR);
  const int64_t n1 = (int64_t)cx * cdiv(v.ny, SEAM_YT * TILE_Y) * zsel;
  const int64_t n2 = cdiv((int64_t)cx * v.ny * v.nz, 256);
  PNM_LAUNCH(c, "k_ccl_seams", k_ccl_seams, dim3((unsigned)(n0 + n1 + n2)), dim3(256), 0, labels,
             reinterpret_cast<const int2*>(xface), v.nx, v.ny, v.nz, cx, ysel, zsel, n0, n1);
}

void launch_roots_resolve(Ctx* c, const float* sdf, Vol v, const uint64_t* stage,
                          const uint32_t* cnt, int64_t nseg, int seg_cap, const uint64_t* walk,
                          int32_t* roots, uint32_t root_flag, int32_t* min_pore,
                          uint64_t* cap_flag) {
  if (nseg <= 0) return;
  PNM_LAUNCH(c, "k_roots_resolve", k_roots_resolve, dim3(cdiv(nseg, 4)), dim3(256), 0, sdf, v.nx,
             v.ny, v.nz, v.n, stage, cnt, nseg, seg_cap, walk, roots, root_flag, min_pore,
             reinterpret_cast<unsigned long long*>(cap_flag),
             make_udiv((uint32_t)cdiv(v.nx, 64)), make_udiv((uint32_t)cdiv(v.ny, TILE_Y)));
}

}  // namespace pnm
