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

// SEG = false: peaks only (extract_pores). SEG = true: also markers, successor, band, and the
// first-appearance minimum of every sdf>0 peak initialised to the peak itself (a basin's first
// voxel is never after its peak). Lists are staged per block in wave/lane order.
template <bool SEG>
__global__ __launch_bounds__(256) void k_pre(const float* __restrict__ sdf, int nx, int ny,
                                             int nz, float thr, PreOut o) {
  __shared__ unsigned wc_peak[4], wc_band[4];
  const int x = blockIdx.x * 64 + threadIdx.x;
  const int y = blockIdx.y * 4 + threadIdx.y;
  const int z = blockIdx.z;
  const int lane = threadIdx.x, wave = threadIdx.y;
  const bool in = (x < nx) && (y < ny);
  bool is_peak = false, band = false;
  int i = 0;
  float cv = 0.0f;
  if (in) {
    i = (z * ny + y) * nx + x;
    cv = sdf[i];
    int nxt = i;
    bool pos_peak = false;
    if (!(cv <= 0.0f)) {  // extract_pores: `if (cv <= 0.0f) return;` (:90)
      const int best = ascent_best(sdf, x, y, z, nx, ny, nz, i, cv);
      is_peak = (best == i);
      if (cv > 0.0f) nxt = best;
      pos_peak = is_peak && (cv > 0.0f);
    }
    if (SEG) {
      o.next[i] = nxt;
      if (pos_peak) o.min_pore[i] = i;
    }
  }
  if (SEG) {
    // init_markers (:177) with the label of a marker set to the start of its x-run inside this
    // wave's 64-voxel row chunk (the run is connected along x, so the component minimum is
    // unchanged); the CCL then only unites runs.
    const bool mk = in && (cv < thr);
    const uint64_t m = __ballot(mk);
    const uint64_t starts = m & ~(m << 1);
    const uint64_t upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1);
    const uint64_t sm = starts & upto;
    const int s = sm ? 63 - __clzll((long long)sm) : lane;
    if (in) o.labels[i] = mk ? i - (lane - s) : -1;
    band = in && !(cv >= 0.0f) && !mk;  // flood candidates (:245-248)
  }
  const int64_t bid = ((int64_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
  unsigned ptot;
  const unsigned ppos = block_stage_pos<4>(is_peak, wc_peak, wave, lane, ptot);
  if (is_peak) o.peak_stage[bid * 256 + ppos] = (uint32_t)i;
  if (threadIdx.x == 0 && threadIdx.y == 0) o.peak_cnt[bid] = ptot;
  if (SEG) {
    unsigned btot;
    const unsigned bpos = block_stage_pos<4>(band, wc_band, wave, lane, btot);
    if (band) o.band_stage[bid * 256 + bpos] = (uint32_t)i;
    if (threadIdx.x == 0 && threadIdx.y == 0) o.band_cnt[bid] = btot;
  }
}

int64_t pre_blocks(Vol v) { return (int64_t)cdiv(v.nx, 64) * cdiv(v.ny, 4) * v.nz; }

void launch_pre(Ctx* c, const float* sdf, Vol v, float thr, const PreOut& o) {
  if (o.labels)
    PNM_LAUNCH(c, "k_pre", k_pre<true>, vox_grid(v), kVoxBlock, 0, sdf, v.nx, v.ny, v.nz,
                       thr, o);
  else
    PNM_LAUNCH(c, "k_pre_peaks", k_pre<false>, vox_grid(v), kVoxBlock, 0, sdf, v.nx, v.ny, v.nz,
               thr, o);
}

// Squared-positive-SDF weighted centroid, dz -> dy -> dx order, FP32
// (pore_extraction.hpp:107-127, identical block at :543-562).
__global__ void k_centroids(const float* __restrict__ sdf, int nx, int ny, int nz, float ox,
                            float oy, float oz, float sx, float sy, float sz,
                            const uint32_t* __restrict__ peaks, int64_t count,
                            pnm_pore_t* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const int ci = (int)peaks[k];
  const int ix = ci % nx, iy = (ci / nx) % ny, iz = ci / (nx * ny);
  const int xs[3] = {wm1(ix, nx), ix, wp1(ix, nx)};
  const int ys[3] = {wm1(iy, ny), iy, wp1(iy, ny)};
  const int zs[3] = {wm1(iz, nz), iz, wp1(iz, nz)};
  float sw = 0.0f, px = 0.0f, py = 0.0f, pz = 0.0f;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      for (int cc = 0; cc < 3; ++cc) {
        const float val = sdf[(zs[a] * ny + ys[b]) * nx + xs[cc]];
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
  P.radius = sdf[ci];
  out[k] = P;
}

void launch_centroids(Ctx* c, const float* sdf, Vol v, const float org[3], const float spc[3],
                      const uint32_t* peaks, int64_t count, pnm_pore_t* out) {
  if (count <= 0) return;
  PNM_LAUNCH(c, "k_centroids", k_centroids, dim3(cdiv(count, 256)), dim3(256), 0, sdf, v.nx,
                     v.ny, v.nz, org[0], org[1], org[2], spc[0], spc[1], spc[2], peaks, count,
                     out);
}

// Ascent successor of every voxel (all_voxels) or of sdf > 0 voxels (others: themselves).
__global__ __launch_bounds__(256) void k_next(const float* __restrict__ sdf, int nx, int ny,
                                              int nz, int32_t* __restrict__ next,
                                              bool all_voxels) {
  const int x = blockIdx.x * 64 + threadIdx.x;
  const int y = blockIdx.y * 4 + threadIdx.y;
  const int z = blockIdx.z;
  if (x >= nx || y >= ny) return;
  const int i = (z * ny + y) * nx + x;
  const float cv = sdf[i];
  int nxt = i;
  if (all_voxels || cv > 0.0f) nxt = ascent_best(sdf, x, y, z, nx, ny, nz, i, cv);
  next[i] = nxt;
}

void launch_next(Ctx* c, const float* sdf, Vol v, int32_t* next, bool all_voxels) {
  PNM_LAUNCH(c, "k_next", k_next, vox_grid(v), kVoxBlock, 0, sdf, v.nx, v.ny, v.nz, next,
                     all_voxels);
}

// Root = the walker after min(depth, 512) successor steps (pore_extraction.hpp:279-309).
__global__ __launch_bounds__(256) void k_chase(const float* __restrict__ sdf, int64_t n,
                                               const int32_t* __restrict__ next,
                                               int32_t* __restrict__ roots, bool all_voxels) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int w = (int)i;
  if (all_voxels || sdf[i] > 0.0f) {
    for (int s = 0; s < kMaxSteps; ++s) {
      const int nx = next[w];
      if (nx == w) break;
      w = nx;
    }
  }
  roots[i] = w;
}

void launch_chase(Ctx* c, const float* sdf, Vol v, const int32_t* next, int32_t* roots,
                  bool all_voxels) {
  PNM_LAUNCH(c, "k_chase", k_chase, dim3(cdiv(v.n, 256)), dim3(256), 0, sdf, v.n, next,
                     roots, all_voxels);
}

}  // namespace pnm
