// k_extract.hip — one streaming pass over the SDF: 26-neighbour peak test (extract_pores
// predicate), ascent successor, watershed markers and the flood band; plus the FP32 centroid
// refinement of the peaks.
//
// FP32 exactness: contraction is off for this file, so `px += dx*w` and `o + (i+f)*s` round
// exactly as the reference's written expressions (pore_extraction.hpp:107-127).
#pragma clang fp contract(off)

#include "k_common.hpp"

namespace pnm {

// Voxel-kernel launch shape: a wave is one 64-voxel x-row; 4 rows per block; grid.z = z-slice.
static dim3 vox_grid(Vol v) { return dim3(cdiv(v.nx, 64), cdiv(v.ny, 4), (unsigned)v.nz); }
static const dim3 kVoxBlock(64, 4, 1);

// One pass over the SDF, LDS-tiled: a 512-thread block owns a 64(x) x 8(y) column tile and
// marches PRE_ZC slices in z. Per voxel:
//   - the 26-neighbour lexicographic (value, index) maximum (pore_extraction.hpp:93-104 and
//     :280-304): peak of the extract_pores predicate (`cv <= 0` excluded, :90) and the ascent
//     successor;
//   - SEG only: marker label = start of its x-run inside the wave's 64-voxel row chunk
//     (init_markers :177; the run is connected along x, so component minima are unchanged),
//     flood-band membership (:245-248), and the first-appearance minimum of each sdf>0 peak
//     seeded with the peak's own index (basin voxels before it lower it later, in the chase).
//
// The maximum is separable over the ordered (value, index) keys of lex_key: max over z of max
// over y of max over x: a haloed slice of keys goes to LDS, the x-triple
// maxima of its 10 rows to LDS, each thread keeps the 3x3 maximum of its own column for three
// consecutive slices in registers.
// Peaks and band voxels are appended to a per-wave staging segment of 64 * PRE_ZC slots.
constexpr int PRE_TY = 8;
constexpr int PRE_ZC = 16;
constexpr int PRE_CAP = 64 * PRE_ZC;  // staging slots per wave
constexpr int PRE_LX = 66, PRE_LY = PRE_TY + 2;

template <bool SEG>
__global__ __launch_bounds__(64 * PRE_TY) void k_pre(const float* __restrict__ sdf, int nx,
                                                     int ny, int nz, float thr, PreOut o) {
  __shared__ uint64_t K[PRE_LY][PRE_LX];
  __shared__ uint64_t XM[PRE_LY][64];
  const int lane = threadIdx.x, wave = threadIdx.y;
  const int tid = wave * 64 + lane;
  const int x0 = blockIdx.x * 64, y0 = blockIdx.y * PRE_TY, z0 = blockIdx.z * PRE_ZC;
  const int x = x0 + lane, y = y0 + wave;
  const bool col_in = (x < nx) && (y < ny);
  const bool near = nx >= 64 && ny > PRE_TY;  // halo coordinates lie in [-N, 2N)
  const int64_t seg_id =
      (((int64_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x) * PRE_TY + wave;
  uint32_t* __restrict__ pstage = o.peak_stage + seg_id * PRE_CAP;
  uint32_t* __restrict__ bstage = SEG ? o.band_stage + seg_id * PRE_CAP : nullptr;
  unsigned npk = 0, nbd = 0;
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;

  // 3x3 (x, y) key maximum of this thread's column at slice zg; also returns the centre key.
  auto slice_max = [&](int zg, uint64_t& centre) -> uint64_t {
    const int zw = wrap_any(zg, nz);
    const float* __restrict__ plane = sdf + (int64_t)zw * ny * nx;
    for (int e = tid; e < PRE_LX * PRE_LY; e += 64 * PRE_TY) {
      const int ly = e / PRE_LX, lx = e - ly * PRE_LX;
      const int gx = near ? wrap_near(x0 - 1 + lx, nx) : wrap_any(x0 - 1 + lx, nx);
      const int gy = near ? wrap_near(y0 - 1 + ly, ny) : wrap_any(y0 - 1 + ly, ny);
      const int gi = (zw * ny + gy) * nx + gx;
      K[ly][lx] = lex_key(plane[(int64_t)gy * nx + gx], (uint32_t)gi);
    }
    __syncthreads();
    for (int e = tid; e < PRE_LY * 64; e += 64 * PRE_TY) {
      const int ly = e >> 6, lx = e & 63;
      XM[ly][lx] = kmax(kmax(K[ly][lx], K[ly][lx + 1]), K[ly][lx + 2]);
    }
    centre = K[wave + 1][lane + 1];
    __syncthreads();
    return kmax(kmax(XM[wave][lane], XM[wave + 1][lane]), XM[wave + 2][lane]);
  };

  uint64_t c_prev, c_cur, c_next;
  uint64_t m_prev = slice_max(z0 - 1, c_prev);
  uint64_t m_cur = slice_max(z0, c_cur);
  const int zend = min(z0 + PRE_ZC, nz);
  for (int z = z0; z < zend; ++z) {
    const uint64_t m_next = slice_max(z + 1, c_next);
    const int i = (z * ny + y) * nx + x;
    const float cv = key_value(c_cur);
    bool is_peak = false;
    int nxt = i;
    if (col_in && !(cv <= 0.0f)) {
      if (cv != cv) {
        is_peak = true;  // nothing beats a NaN centre
      } else {
        const int best = (int)(uint32_t)kmax(kmax(m_prev, m_cur), m_next);
        is_peak = best == i;
        if (cv > 0.0f) nxt = best;
      }
    }
    const uint64_t pm = __ballot(is_peak);
    if (is_peak) pstage[npk + lanes_below(pm)] = (uint32_t)i;
    npk += (unsigned)__popcll(pm);
    if (SEG) {
      const bool mk = col_in && (cv < thr);
      const uint64_t m = __ballot(mk);
      const uint64_t starts = m & ~(m << 1);
      const uint64_t smask = starts & (below | (1ull << lane));
      const int st = smask ? 63 - __clzll((long long)smask) : lane;
      if (col_in) {
        o.labels[i] = mk ? i - (lane - st) : -1;
        o.next[i] = nxt;
        if (is_peak && cv > 0.0f) o.min_pore[i] = i;
      }
      const bool band = col_in && !(cv >= 0.0f) && !mk;
      const uint64_t bm = __ballot(band);
      if (band) bstage[nbd + lanes_below(bm)] = (uint32_t)i;
      nbd += (unsigned)__popcll(bm);
    }
    m_prev = m_cur;
    m_cur = m_next;
    c_cur = c_next;
  }
  if (lane == 0) {
    o.peak_cnt[seg_id] = npk;
    if (SEG) o.band_cnt[seg_id] = nbd;
  }
}

static dim3 pre_grid(Vol v) {
  return dim3(cdiv(v.nx, 64), cdiv(v.ny, PRE_TY), cdiv(v.nz, PRE_ZC));
}

int64_t pre_blocks(Vol v) {
  const dim3 g = pre_grid(v);
  return (int64_t)g.x * g.y * g.z * PRE_TY;
}
int64_t pre_stage_cap() { return PRE_CAP; }

void launch_pre(Ctx* c, const float* sdf, Vol v, float thr, const PreOut& o) {
  if (o.labels)
    PNM_LAUNCH(c, "k_pre", k_pre<true>, pre_grid(v), dim3(64, PRE_TY), 0, sdf, v.nx, v.ny, v.nz,
               thr, o);
  else
    PNM_LAUNCH(c, "k_pre_peaks", k_pre<false>, pre_grid(v), dim3(64, PRE_TY), 0, sdf, v.nx, v.ny,
               v.nz, thr, o);
}

// Squared-positive-SDF weighted centroid, dz -> dy -> dx order, FP32
// (pore_extraction.hpp:107-127, identical block at :543-562). A wave takes two pores: lanes
// 0..26 and 32..58 load one of a pore's 27 values each (so the peaks' scattered loads are spread
// over many waves), and the pore's first lane sums them from LDS in the reference's order.
constexpr int CENT_PER_BLOCK = 8;  // 4 waves x 2 pores
__global__ __launch_bounds__(256) void k_centroids(const float* __restrict__ sdf, int nx, int ny,
                                                   int nz, float ox, float oy, float oz, float sx,
                                                   float sy, float sz,
                                                   const uint32_t* __restrict__ peaks,
                                                   int64_t count, pnm_pore_t* __restrict__ out) {
  __shared__ float V[CENT_PER_BLOCK][32];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int slot = wave * 2 + (lane >> 5), q = lane & 31;
  const int64_t k = (int64_t)blockIdx.x * CENT_PER_BLOCK + slot;
  const bool valid = k < count;
  const int ci = valid ? (int)peaks[k] : 0;
  const int ix = ci % nx, iy = (ci / nx) % ny, iz = ci / (nx * ny);
  if (valid && q < 27) {
    const int a = q / 9, b = (q / 3) % 3, cc = q % 3;
    const int zz = a == 0 ? wm1(iz, nz) : (a == 1 ? iz : wp1(iz, nz));
    const int yy = b == 0 ? wm1(iy, ny) : (b == 1 ? iy : wp1(iy, ny));
    const int xx = cc == 0 ? wm1(ix, nx) : (cc == 1 ? ix : wp1(ix, nx));
    V[slot][q] = sdf[(zz * ny + yy) * nx + xx];
  }
  __builtin_amdgcn_wave_barrier();
  if (!valid || q != 0) return;
  float sw = 0.0f, px = 0.0f, py = 0.0f, pz = 0.0f;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) {
        const float val = V[slot][(a * 3 + b) * 3 + cc];
        float w = val > 0.0f ? val : 0.0f;
        w = w * w;
        sw += w;
        px += (float)(cc - 1) * w;
        py += (float)(b - 1) * w;
        pz += (float)(a - 1) * w;
      }
  float fx = 0.0f, fy = 0.0f, fz = 0.0f;
  if (sw > 1e-6f) {
    fx = px / sw;
    fy = py / sw;
    fz = pz / sw;
  }
  pnm_pore_t P;
  P.x = ox + ((float)ix + fx) * sx;
  P.y = oy + ((float)iy + fy) * sy;
  P.z = oz + ((float)iz + fz) * sz;
  P.radius = V[slot][13];  // the centre value
  out[k] = P;
}

void launch_centroids(Ctx* c, const float* sdf, Vol v, const float org[3], const float spc[3],
                      const uint32_t* peaks, int64_t count, pnm_pore_t* out) {
  if (count <= 0) return;
  PNM_LAUNCH(c, "k_centroids", k_centroids, dim3(cdiv(count, CENT_PER_BLOCK)), dim3(256), 0, sdf,
             v.nx, v.ny, v.nz, org[0], org[1], org[2], spc[0], spc[1], spc[2], peaks, count, out);
}

// Successor of the gradient walk. Mode kWalkPores: ascent for sdf > 0 voxels, descent for the
// others (gradient_path, pore_extraction.hpp:273-304: the direction is fixed by the starting
// voxel); kWalkAll: ascent for every voxel (nf::flow_basins, :621-652); kWalkSolids: descent
// for !(sdf > 0) voxels, the others keep themselves.
__global__ __launch_bounds__(256) void k_next(const float* __restrict__ sdf, int nx, int ny,
                                              int nz, int32_t* __restrict__ next, int mode) {
  const int x = blockIdx.x * 64 + threadIdx.x;
  const int y = blockIdx.y * 4 + threadIdx.y;
  const int z = blockIdx.z;
  if (x >= nx || y >= ny) return;
  const int i = (z * ny + y) * nx + x;
  const float cv = sdf[i];
  const bool pore = cv > 0.0f;
  int nxt = i;
  if (mode == kWalkAll || (mode == kWalkPores && pore))
    nxt = gradient_best<true>(sdf, x, y, z, nx, ny, nz, i, cv);
  else if (!pore)
    nxt = gradient_best<false>(sdf, x, y, z, nx, ny, nz, i, cv);
  next[i] = nxt;
}

void launch_next(Ctx* c, const float* sdf, Vol v, int32_t* next, int mode) {
  PNM_LAUNCH(c, "k_next", k_next, vox_grid(v), kVoxBlock, 0, sdf, v.nx, v.ny, v.nz, next, mode);
}

// Root = the walker after min(depth, 512) successor steps (pore_extraction.hpp:279-309). In
// mode kWalkSolids only !(sdf > 0) voxels are written (the caller holds the pore roots).
__global__ __launch_bounds__(256) void k_chase(const float* __restrict__ sdf, int64_t n,
                                               const int32_t* __restrict__ next,
                                               int32_t* __restrict__ roots, int mode) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (mode == kWalkSolids && sdf[i] > 0.0f) return;
  int w = (int)i;
  for (int s = 0; s < kMaxSteps; ++s) {
    const int nx = next[w];
    if (nx == w) break;
    w = nx;
  }
  roots[i] = w;
}

void launch_chase(Ctx* c, const float* sdf, Vol v, const int32_t* next, int32_t* roots,
                  int mode) {
  PNM_LAUNCH(c, "k_chase", k_chase, dim3(cdiv(v.n, 256)), dim3(256), 0, sdf, v.n, next, roots,
             mode);
}

void launch_descent_roots(Ctx* c, const float* sdf, Vol v, int32_t* roots) {
  int32_t* next = c->arena.get_n<int32_t>((size_t)v.n);
  launch_next(c, sdf, v, next, kWalkSolids);
  launch_chase(c, sdf, v, next, roots, kWalkSolids);
}

}  // namespace pnm
