// k_flow.hip — network-flow bookkeeping (extract_network_flow_k, pore_extraction.hpp:493-997)
// on gfx950, FP64 throughout where the reference is FP64.
//
// Stages (one stream):
//   segmentation (pipeline.hip) -> per-label peak voxels -> FP32 centroids -> FP64 trilinear
//   pore pressure -> flow basins (capped ascent from every cell) -> face init (+ compacted core,
//   film and basin-boundary face lists) -> core-tier face CCL (lock-free union-find, int64) ->
//   film attachment (monotone min propagation to a fixpoint) -> leftover-tier CCL -> patch roots
//   sorted by (pair, root face) -> per-face records sorted by (throat or pore, face) -> one
//   thread per throat / pore sums its faces in face order.
//
// The reference accumulates Q, A, C and the residual with unordered FP64 atomics (:898-938).
// Summing each key's contributions sequentially in ascending face order makes the result
// deterministic and identical to a serial loop over the faces (the oracle's order).
#pragma clang fp contract(off)

#include <algorithm>
#include <cstdlib>
#include <new>

#include "k_common.hpp"

namespace pnm {

constexpr int64_t kSent = 0x7ffffffffffffffeLL;  // film sentinel, pore_extraction.hpp:671

static int nbits64(uint64_t v) {
  int b = 0;
  while (b < 64 && (v >> b) != 0) ++b;
  return b;
}

struct Geo {
  int nx, ny, nz;
};

__device__ __forceinline__ void cell_xyz(int c, Geo g, int& x, int& y, int& z) {
  x = c % g.nx;
  y = (c / g.nx) % g.ny;
  z = c / (g.nx * g.ny);
}

__device__ __forceinline__ int nb_cell(int x, int y, int z, int d, Geo g) {
  const int xx = d == 0 ? wp1(x, g.nx) : x;
  const int yy = d == 1 ? wp1(y, g.ny) : y;
  const int zz = d == 2 ? wp1(z, g.nz) : z;
  return (zz * g.ny + yy) * g.nx + xx;
}

__global__ void k_iota_u32(uint32_t* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (uint32_t)i;
}

// ---------------- pore centres and pressures ----------------
// nf::peaks (:528-533): peak[seg[r]-1] = r for every sdf>0 self-root r (= the sdf>0 peaks).
__global__ void k_flow_peaks(const float* __restrict__ sdf, const int32_t* __restrict__ seg,
                             const uint32_t* __restrict__ peaks, int64_t npeaks,
                             int32_t* __restrict__ peak) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= npeaks) return;
  const uint32_t r = peaks[j];
  const int32_t l = seg[r];
  if (sdf[r] > 0.0f && l > 0) peak[l - 1] = (int32_t)r;
}

// Largest label of a given segmentation (its pore count) into *mx (starts at 0).
__global__ __launch_bounds__(256) void k_label_max(const int32_t* __restrict__ seg, int64_t n,
                                                   int* __restrict__ mx) {
  int m = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    m = max(m, seg[i]);
#pragma unroll
  for (int sh = 32; sh > 0; sh >>= 1) m = max(m, __shfl_xor(m, sh, 64));
  if ((threadIdx.x & 63) == 0 && m > 0) atomicMax(mx, m);
}

// nf::pore_pressure (:582-613): grid coordinate in FP32, then FP64 trilinear weights in the
// reference's dz -> dy -> dx order.
__global__ void k_pore_pressure(const pnm_pore_t* __restrict__ pores, int64_t np, Geo g,
                                float ox, float oy, float oz, float sx, float sy, float sz,
                                const double* __restrict__ p, double* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  const pnm_pore_t po = pores[k];
  const float gx = (po.x - ox) / sx, gy = (po.y - oy) / sy, gz = (po.z - oz) / sz;
  const double gg[3] = {(double)gx, (double)gy, (double)gz};
  int b[3];
  double f[3];
  for (int a = 0; a < 3; ++a) {
    const double fl = floor(gg[a]);
    b[a] = (int)fl;
    f[a] = gg[a] - fl;
  }
  const int N[3] = {g.nx, g.ny, g.nz};
  double acc = 0.0;
  for (int dz = 0; dz < 2; ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx) {
        const double wt =
            (dx ? f[0] : 1.0 - f[0]) * (dy ? f[1] : 1.0 - f[1]) * (dz ? f[2] : 1.0 - f[2]);
        int q[3] = {b[0] + dx, b[1] + dy, b[2] + dz};
        for (int a = 0; a < 3; ++a) q[a] = ((q[a] % N[a]) + N[a]) % N[a];
        acc += wt * p[((int64_t)q[2] * g.ny + q[1]) * g.nx + q[0]];
      }
  out[k] = acc;
}

// nf::flow_basins (:621-652): flowLab = seg at the capped ascent root of every cell. Each
// thread chases four cells (i, i + B, i + 2B, i + 3B for a block of B) side by side, so their
// dependent loads overlap.
constexpr int kFlowLabChains = 4;
__global__ __launch_bounds__(256) void k_flow_lab(int64_t n, const int32_t* __restrict__ next,
                                                  const int32_t* __restrict__ seg,
                                                  int32_t* __restrict__ lab) {
  const int64_t base = (int64_t)blockIdx.x * blockDim.x * kFlowLabChains + threadIdx.x;
  int w[kFlowLabChains], st[kFlowLabChains];
  bool go[kFlowLabChains];
#pragma unroll
  for (int c = 0; c < kFlowLabChains; ++c) {
    const int64_t i = base + (int64_t)c * blockDim.x;
    go[c] = i < n;
    w[c] = go[c] ? (int)i : 0;
    st[c] = 0;
  }
  while (true) {
    int t[kFlowLabChains];
#pragma unroll
    for (int c = 0; c < kFlowLabChains; ++c) t[c] = go[c] ? next[w[c]] : 0;
    bool more = false;
#pragma unroll
    for (int c = 0; c < kFlowLabChains; ++c) {
      if (!go[c]) continue;
      if (t[c] == w[c] || st[c] == kMaxSteps) {
        go[c] = false;
      } else {
        w[c] = t[c];
        go[c] = ++st[c] < kMaxSteps;
      }
      more |= go[c];
    }
    if (!more) break;
  }
#pragma unroll
  for (int c = 0; c < kFlowLabChains; ++c) {
    const int64_t i = base + (int64_t)c * blockDim.x;
    if (i < n) lab[i] = seg[w[c]];
  }
}

// ---------------- faces ----------------
// Block-wide exclusive prefix of a per-thread count in 0..7 (three ballots per wave, one LDS
// word per wave and list). Every thread of the block must call it; the caller separates two
// uses of the same `wcnt` words by a barrier.
__device__ __forceinline__ void small_count_ballots(unsigned cnt, int lane, unsigned& below,
                                                    unsigned& wave_total) {
  const uint64_t b0 = __ballot(cnt & 1u), b1 = __ballot(cnt & 2u), b2 = __ballot(cnt & 4u);
  const uint64_t m = lane ? (~0ull >> (64 - lane)) : 0ull;
  below = (unsigned)(__popcll(b0 & m) + 2 * __popcll(b1 & m) + 4 * __popcll(b2 & m));
  wave_total = (unsigned)(__popcll(b0) + 2 * __popcll(b1) + 4 * __popcll(b2));
}

constexpr int kFaceBlock = 256;                  // cells per block of k_face_init
constexpr int kFaceCap = 3 * kFaceBlock;         // staged faces per block and list
constexpr int kResCap = 2 * kFaceCap;            // staged residual records per block

// nf::face_init (:681-696). Also stages, per block and in face order, the core faces, the film
// faces, every interface face and every basin boundary face touching a pore (one residual record
// per pore side, (pore index << 32 | face)); compacting the blocks in block order gives lists in
// ascending face order, so the records need only a stable sort on their key bits.
// kGeo: throat geometry without fields: the openness is the binary one of the project solver
// (a face is open iff both of its cells have sdf > 0), and no residual records are staged.
template <bool kGeo>
__global__ __launch_bounds__(kFaceBlock) void k_face_init(
    const float* __restrict__ sdf, const int32_t* __restrict__ lab, Geo g, int64_t n,
    const double* __restrict__ ox, const double* __restrict__ oy, const double* __restrict__ oz,
    int64_t* __restrict__ fpar, int64_t* __restrict__ fpair, int64_t* __restrict__ par,
    uint32_t* __restrict__ st_core, uint32_t* __restrict__ st_film, uint32_t* __restrict__ st_itf,
    uint32_t* __restrict__ cnt3, uint64_t* __restrict__ st_res, uint32_t* __restrict__ cnt_res,
    uint8_t* __restrict__ itfm) {
  __shared__ unsigned wc[4][4];
  const int64_t c64 = (int64_t)blockIdx.x * kFaceBlock + threadIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bool in = c64 < n;
  const int c = (int)c64;
  int x = 0, y = 0, z = 0;
  if (in) cell_xyz(c, g, x, y, z);
  const int a = in ? lab[c] : 0;
  const float sc = in ? sdf[c] : 0.0f;
  unsigned core_m = 0, film_m = 0;  // bit d: the +d face is a core / film face
  int rk0[3] = {-1, -1, -1}, rk1[3] = {-1, -1, -1};  // pore indices of the residual records
  if (in) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int64_t f = 3 * c64 + d;
      const int nb = nb_cell(x, y, z, d, g);
      const int b = lab[nb];
      const float snb = sdf[nb];
      const double opn = kGeo ? ((sc > 0.0f && snb > 0.0f) ? 1.0 : 0.0)
                              : (ox ? (d == 0 ? ox[nb] : (d == 1 ? oy[nb] : oz[nb])) : 1.0);
      const bool itf = (a != b && a > 0 && b > 0 && opn > 0.0);
      const bool is_core = itf && sc > 0.0f && snb > 0.0f;
      // only interface faces get entries: every reader of par, fpar and fpair takes faces from
      // the interface lists or behind the cell masks (itfm) written below
      if (itf) {
        par[f] = is_core ? f : -1;
        fpar[f] = is_core ? f : kSent;
        fpair[f] = ((int64_t)(a < b ? a : b) << 32) | (int64_t)(a < b ? b : a);
      }
      core_m |= (unsigned)is_core << d;
      film_m |= (unsigned)(itf && !is_core) << d;
      if (!kGeo && a != b) {
        if (a > 0) {
          rk0[d] = a - 1;
          if (b > 0) rk1[d] = b - 1;
        } else if (b > 0) {
          rk0[d] = b - 1;
        }
      }
    }
  }
  if (in) itfm[c] = (uint8_t)(core_m | film_m);
  const unsigned n_core = __popc(core_m), n_film = __popc(film_m), n_itf = __popc(core_m | film_m);
  unsigned n_res = 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) n_res += (rk0[d] >= 0) + (rk1[d] >= 0);
  unsigned below[4], wt[4];
  small_count_ballots(n_core, lane, below[0], wt[0]);
  small_count_ballots(n_film, lane, below[1], wt[1]);
  small_count_ballots(n_itf, lane, below[2], wt[2]);
  small_count_ballots(n_res, lane, below[3], wt[3]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < 4; ++k) wc[k][wave] = wt[k];
  __syncthreads();
  unsigned pos[4], tot[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    unsigned base = 0, t = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      const unsigned cw = wc[k][w];
      if (w < wave) base += cw;
      t += cw;
    }
    pos[k] = base + below[k];
    tot[k] = t;
  }
  const int64_t b32 = (int64_t)blockIdx.x * kFaceCap;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const uint32_t f = (uint32_t)(3 * c64 + d);
    if (core_m >> d & 1u) st_core[b32 + pos[0]++] = f;
    if (film_m >> d & 1u) st_film[b32 + pos[1]++] = f;
    if ((core_m | film_m) >> d & 1u) st_itf[b32 + pos[2]++] = f;
    if (rk0[d] >= 0) st_res[(int64_t)blockIdx.x * kResCap + pos[3]++] = ((uint64_t)rk0[d] << 32) | f;
    if (rk1[d] >= 0) st_res[(int64_t)blockIdx.x * kResCap + pos[3]++] = ((uint64_t)rk1[d] << 32) | f;
  }
  if (threadIdx.x == 0) {
    const int64_t nb = gridDim.x;
    cnt3[blockIdx.x] = tot[0];
    cnt3[nb + blockIdx.x] = tot[1];
    cnt3[2 * nb + blockIdx.x] = tot[2];
    cnt_res[blockIdx.x] = tot[3];
  }
}

__device__ __forceinline__ int64_t uf_find64(const int64_t* P, int64_t x) {
  int64_t p = ld_agent64(P + x);
  while (p != x) {
    x = p;
    p = ld_agent64(P + x);
  }
  return x;
}

__device__ __forceinline__ void uf_unite64(int64_t* P, int64_t a, int64_t b) {
  while (true) {
    a = uf_find64(P, a);
    b = uf_find64(P, b);
    if (a == b) return;
    if (a > b) {
      const int64_t t = a;
      a = b;
      b = t;
    }
    const unsigned long long old = atomicMin(reinterpret_cast<unsigned long long*>(P + b),
                                             (unsigned long long)a);
    if ((int64_t)old == b) return;
    b = (int64_t)old;
  }
}

// Iterates the 81 candidate faces (27 cells x 3 directions) of face f that share its pair key
// and pass the doubled-offset adjacency |2dc + e_d2 - e_d|^2 <= 8 (:711-723, :776-791).
template <class F>
__device__ __forceinline__ void for_adjacent(int64_t f, Geo g, const int64_t* __restrict__ fpair,
                                             const uint8_t* __restrict__ itfm, F&& fn) {
  const int c = (int)(f / 3), d = (int)(f % 3);
  int x, y, z;
  cell_xyz(c, g, x, y, z);
  const int64_t pk = fpair[f];
  const int xs[3] = {wm1(x, g.nx), x, wp1(x, g.nx)};
  const int ys[3] = {wm1(y, g.ny), y, wp1(y, g.ny)};
  const int zs[3] = {wm1(z, g.nz), z, wp1(z, g.nz)};
  // the interface faces of the 27 cells (one mask byte per cell): only those can share the pair
  unsigned m[27];
#pragma unroll
  for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
      for (int dx = -1; dx <= 1; ++dx)
        m[(dz + 1) * 9 + (dy + 1) * 3 + dx + 1] =
            itfm[(zs[dz + 1] * g.ny + ys[dy + 1]) * g.nx + xs[dx + 1]];
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        const unsigned mc = m[(dz + 1) * 9 + (dy + 1) * 3 + dx + 1];
        if (!mc) continue;
        const int c2 = (zs[dz + 1] * g.ny + ys[dy + 1]) * g.nx + xs[dx + 1];
        for (int d2 = 0; d2 < 3; ++d2) {
          if (!((mc >> d2) & 1u)) continue;
          const int64_t f2 = 3 * (int64_t)c2 + d2;
          if (f2 == f) continue;
          const int Dx = 2 * dx + (d2 == 0) - (d == 0);
          const int Dy = 2 * dy + (d2 == 1) - (d == 1);
          const int Dz = 2 * dz + (d2 == 2) - (d == 2);
          if (Dx * Dx + Dy * Dy + Dz * Dz > 8) continue;
          if (fpair[f2] != pk) continue;
          fn(f2);
        }
      }
}

// cclPass face_merge (:703-738): unions restricted to f2 > f (the relation is symmetric).
__global__ __launch_bounds__(256) void k_face_union(const uint32_t* __restrict__ list,
                                                    int64_t cnt, Geo g,
                                                    const int64_t* __restrict__ fpair,
                                                    const uint8_t* __restrict__ itfm,
                                                    int64_t* par) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cnt) return;
  const int64_t f = list[j];
  for_adjacent(f, g, fpair, itfm, [&](int64_t f2) {
    if (f2 > f && ld_agent64(par + f2) >= 0) uf_unite64(par, f, f2);
  });
}

// face_flatten (:740-749) fused with core_label / leftover_label: fpar[f] = root.
__global__ void k_face_flatten(const uint32_t* __restrict__ list, int64_t cnt, int64_t* par,
                               int64_t* __restrict__ fpar) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cnt) return;
  const int64_t f = list[j];
  int64_t l = par[f];
  int64_t p = par[l];
  while (p != l) {
    l = p;
    p = par[l];
  }
  par[f] = l;
  fpar[f] = l;
}

// film_attach (:767-800): in-place monotone minimum over adjacent same-pair faces labelled
// with a core root (or an attached film); the fixpoint does not depend on the order.
__global__ __launch_bounds__(256) void k_film_attach(const uint32_t* __restrict__ list,
                                                     int64_t cnt, Geo g,
                                                     const int64_t* __restrict__ fpair,
                                                     const uint8_t* __restrict__ itfm,
                                                     int64_t* fpar, unsigned long long* changed) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cnt) return;
  const int64_t f = list[j];
  const int64_t cur = fpar[f];
  int64_t best = cur;
  for_adjacent(f, g, fpair, itfm, [&](int64_t f2) {
    const int64_t l2 = fpar[f2];
    if (l2 >= 0 && l2 < kSent && l2 < best) best = l2;
  });
  if (best < cur) {
    fpar[f] = best;
    atomicAdd(changed, 1ull);
  }
}

// leftover_init (:807-809) over the film list: par = f for unattached films (core entries of
// par were reset to -1 beforehand); compacts them.
__global__ void k_leftover_init(const uint32_t* __restrict__ film, int64_t cnt,
                                const int64_t* __restrict__ fpar, int64_t* __restrict__ par,
                                uint32_t* __restrict__ left, unsigned long long* ctr) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  bool keep = false;
  int64_t f = 0;
  if (j < cnt) {
    f = film[j];
    keep = fpar[f] == kSent;
    if (keep) par[f] = f;
  }
  const uint64_t s = wave_append(ctr, keep, lane);
  if (keep) left[s] = (uint32_t)f;
}

__global__ void k_reset_par(const uint32_t* __restrict__ list, int64_t cnt,
                            int64_t* __restrict__ par) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < cnt) par[list[j]] = -1;
}

// face_roots (:826-833): interface faces that are their own patch root.
// Patch roots (faces with fpar[f] == f) of the core list followed by the film list, staged per
// 256-entry block in list order (blocks 0 .. nbc-1 take the core list, the rest the film list);
// compacting the blocks in order gives the roots in list order without a contended counter.
__global__ __launch_bounds__(256) void k_face_roots(const uint32_t* __restrict__ core,
                                                    int64_t ncore, const uint32_t* __restrict__ film,
                                                    int64_t nfilm, int64_t nbc,
                                                    const int64_t* __restrict__ fpar,
                                                    uint64_t* __restrict__ stage,
                                                    uint32_t* __restrict__ counts) {
  __shared__ unsigned wc[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bool is_core = (int64_t)blockIdx.x < nbc;
  const int64_t j = (is_core ? (int64_t)blockIdx.x : (int64_t)blockIdx.x - nbc) * 256 + threadIdx.x;
  const int64_t cnt = is_core ? ncore : nfilm;
  bool is_root = false;
  int64_t f = 0;
  if (j < cnt) {
    f = is_core ? core[j] : film[j];
    is_root = fpar[f] == f;
  }
  const uint64_t m = __ballot(is_root);
  if (lane == 0) wc[wave] = (unsigned)__popcll(m);
  __syncthreads();
  unsigned base = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    if (w < wave) base += wc[w];
    tot += wc[w];
  }
  if (is_root) stage[(int64_t)blockIdx.x * 256 + base + lanes_below(m)] = (uint64_t)f;
  if (threadIdx.x == 0) counts[blockIdx.x] = tot;
}

// key for the (pair, root face) order: compact pair (lo << B | hi)
__global__ void k_root_pairkeys(const uint64_t* __restrict__ root_fid,
                                const uint32_t* __restrict__ order, int64_t nt,
                                const int64_t* __restrict__ fpair, int B,
                                uint64_t* __restrict__ key) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const int64_t pk = fpair[root_fid[order[t]]];
  const uint64_t lo = (uint64_t)(pk >> 32), hi = (uint64_t)(pk & 0x7fffffff);
  key[t] = (lo << B) | hi;
}

// Final throat order: throats[t] = pair; the slot of each root face is stored in par[root].
__global__ void k_throat_slots(const uint64_t* __restrict__ root_fid,
                               const uint32_t* __restrict__ order, int64_t nt,
                               const int64_t* __restrict__ fpair, int64_t* __restrict__ par,
                               int32_t* __restrict__ throats) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const int64_t f = (int64_t)root_fid[order[t]];
  const int64_t pk = fpair[f];
  throats[2 * t] = (int32_t)(pk >> 32);
  throats[2 * t + 1] = (int32_t)(pk & 0x7fffffff);
  par[f] = t;
}

// Throat records for every interface face (core and film lists): (slot << 32 | face).
__global__ void k_throat_records(const uint32_t* __restrict__ list, int64_t cnt,
                                 const int64_t* __restrict__ fpar,
                                 const int64_t* __restrict__ slot_of,
                                 uint64_t* __restrict__ rec) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cnt) return;
  const int64_t f = list[j];
  const int64_t slot = slot_of[fpar[f]];
  rec[j] = ((uint64_t)slot << 32) | (uint64_t)f;
}

// Segment [start, end) of each key in a (key << 32 | face)-sorted record list.
__global__ void k_seg_bounds(const uint64_t* __restrict__ rec, int64_t cnt,
                             int64_t* __restrict__ start, int64_t* __restrict__ end) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  const uint64_t k = rec[i] >> 32;
  if (i == 0 || (rec[i - 1] >> 32) != k) start[k] = i;
  if (i == cnt - 1 || (rec[i + 1] >> 32) != k) end[k] = i + 1;
}

struct FluxArgs {
  Geo g;
  const int32_t* lab;
  const double *u, *v, *w, *ox, *oy, *oz;
  double Ax, Ay, Az;
};

// q of face f (throat_flux :889-897): the shared face is the -d face of the +d neighbour.
// kGeo: no velocity field (q = 0); only the open area w0 is wanted.
template <bool kGeo = false>
__device__ __forceinline__ double face_q(const FluxArgs& A, int64_t f, int& a, int& b,
                                         double& w0, int& x, int& y, int& z, int& d) {
  const int c = (int)(f / 3);
  d = (int)(f % 3);
  cell_xyz(c, A.g, x, y, z);
  const int nb = nb_cell(x, y, z, d, A.g);
  a = A.lab[c];
  b = A.lab[nb];
  const double vel = kGeo ? 0.0 : (d == 0 ? A.u[nb] : (d == 1 ? A.v[nb] : A.w[nb]));
  const double opn = A.ox ? (d == 0 ? A.ox[nb] : (d == 1 ? A.oy[nb] : A.oz[nb])) : 1.0;
  const double area = d == 0 ? A.Ax : (d == 1 ? A.Ay : A.Az);
  w0 = opn * area;
  return opn * vel * area;
}

// Per-record terms, computed in parallel; the sums below add them in record order.
// Residual (:898-901): the signed q of each (pore, face) record.
__global__ __launch_bounds__(256) void k_resid_terms(const uint64_t* __restrict__ rec,
                                                     int64_t cnt, FluxArgs A,
                                                     double* __restrict__ term) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  const uint64_t r = rec[i];
  const int64_t f = (int64_t)(r & 0xffffffffull);
  const int k = (int)(r >> 32);
  int a, b, x, y, z, d;
  double w0;
  const double q = face_q(A, f, a, b, w0, x, y, z, d);
  term[i] = (a - 1 == k) ? q : -q;
}

// Throat (:903-939): the oriented q, the open area, and the area-weighted face centre relative
// to the lower pore's peak-voxel centre, min-imaged per axis. kGeo: no q (tq is not written).
template <bool kGeo>
__global__ __launch_bounds__(256) void k_throat_terms(
    const uint64_t* __restrict__ rec, int64_t cnt, FluxArgs A, const int32_t* __restrict__ peak,
    float oxo, float oyo, float ozo, float sx, float sy, float sz, double* __restrict__ tq,
    double* __restrict__ ta, double* __restrict__ tcx, double* __restrict__ tcy,
    double* __restrict__ tcz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  const int64_t f = (int64_t)(rec[i] & 0xffffffffull);
  int a, b, ixx, iyy, izz, d;
  double w0;
  const double q = face_q<kGeo>(A, f, a, b, w0, ixx, iyy, izz, d);
  const int lo = a < b ? a : b;
  const double Lw[3] = {(double)A.g.nx * sx, (double)A.g.ny * sy, (double)A.g.nz * sz};
  const int pk = peak[lo - 1];
  int pkx, pky, pkz;
  cell_xyz(pk, A.g, pkx, pky, pkz);
  double fp[3] = {oxo + (ixx + (d == 0 ? 0.5 : 0.0)) * (double)sx,
                  oyo + (iyy + (d == 1 ? 0.5 : 0.0)) * (double)sy,
                  ozo + (izz + (d == 2 ? 0.5 : 0.0)) * (double)sz};
  const double pc[3] = {oxo + pkx * (double)sx, oyo + pky * (double)sy, ozo + pkz * (double)sz};
  for (int a2 = 0; a2 < 3; ++a2) {
    double dv = fp[a2] - pc[a2];
    dv -= Lw[a2] * round(dv / Lw[a2]);
    fp[a2] = dv;
  }
  if (!kGeo) tq[i] = a < b ? q : -q;
  ta[i] = w0;
  tcx[i] = w0 * fp[0];
  tcy[i] = w0 * fp[1];
  tcz[i] = w0 * fp[2];
}

template <int K>
struct TermSets {
  const double* in[K];
  double* out[K];
};

// Keyed sums in a fixed order (the reference adds with unordered FP64 atomics, :898-938; the
// oracle sums in this same order, so the results are bit-identical): one wave per key takes the
// key's terms [start, end) in blocks of 2048; lane l adds the block's terms l, l + 64, ...,
// l + 64*31 one after another from +0.0 (each step a coalesced 64-lane load), the 64 lane sums
// are combined by a pairwise tree (xor butterfly, s = 1, 2, ..., 32), and the block totals are
// added one after another. K term sequences run as independent chains.
template <int K>
__global__ __launch_bounds__(256) void k_key_sums(const int64_t* __restrict__ start,
                                                  const int64_t* __restrict__ end, int64_t nkeys,
                                                  TermSets<K> T) {
  const int lane = threadIdx.x & 63;
  const int64_t key = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (key >= nkeys) return;  // whole wave
  const int64_t s = start[key], e = end[key];
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0.0;
  for (int64_t b = s; b < e; b += 2048) {
    double p[K];
#pragma unroll
    for (int k = 0; k < K; ++k) p[k] = 0.0;
    if (b + 2048 <= e) {
      double v[8][K];
      for (int j0 = 0; j0 < 32; j0 += 8) {  // eight loads in flight per chain
#pragma unroll
        for (int j = 0; j < 8; ++j)
#pragma unroll
          for (int k = 0; k < K; ++k) v[j][k] = T.in[k][b + 64 * (j0 + j) + lane];
#pragma unroll
        for (int j = 0; j < 8; ++j)
#pragma unroll
          for (int k = 0; k < K; ++k) p[k] += v[j][k];
      }
    } else {
      for (int j = 0; j < 32; ++j) {
        const int64_t i = b + 64 * j + lane;
        if (i < e)
#pragma unroll
          for (int k = 0; k < K; ++k) p[k] += T.in[k][i];
      }
    }
#pragma unroll
    for (int sh = 1; sh < 64; sh <<= 1)
#pragma unroll
      for (int k = 0; k < K; ++k) p[k] = p[k] + __shfl_xor(p[k], sh, 64);
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] += p[k];
  }
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) T.out[k][key] = acc[k];
}

// throat_dp (host in the reference, :958-995).
__global__ void k_throat_dp(const int32_t* __restrict__ throats, int64_t nt,
                            const pnm_pore_t* __restrict__ pores, const double* __restrict__ pp,
                            const int32_t* __restrict__ peak, const double* __restrict__ Ar,
                            const double* __restrict__ Cx, const double* __restrict__ Cy,
                            const double* __restrict__ Cz, Geo g, float ox, float oy, float oz,
                            float sx, float sy, float sz, double gpx, double gpy, double gpz,
                            double* __restrict__ dp, double* __restrict__ tc) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const int li = throats[2 * t], lj = throats[2 * t + 1];
  const pnm_pore_t pi = pores[li - 1], pj = pores[lj - 1];
  const double aw = Ar[t];
  const int pk = peak[li - 1], pkj = peak[lj - 1];
  const float o[3] = {ox, oy, oz};
  const float s[3] = {sx, sy, sz};
  const int N[3] = {g.nx, g.ny, g.nz};
  const double L[3] = {(double)g.nx * sx, (double)g.ny * sy, (double)g.nz * sz};
  const int pki[3] = {pk % g.nx, (pk / g.nx) % g.ny, pk / (g.nx * g.ny)};
  const int pkj3[3] = {pkj % g.nx, (pkj / g.nx) % g.ny, pkj / (g.nx * g.ny)};
  const double pip[3] = {(double)pi.x, (double)pi.y, (double)pi.z};
  const double pjp[3] = {(double)pj.x, (double)pj.y, (double)pj.z};
  const double gp[3] = {gpx, gpy, gpz};
  double macro = 0.0;
  for (int a = 0; a < 3; ++a) {
    const double anc = o[a] + pki[a] * (double)s[a];
    const double ancj = o[a] + pkj3[a] * (double)s[a];
    const double ct = (Cx[t] * (a == 0) + Cy[t] * (a == 1) + Cz[t] * (a == 2)) / aw;
    tc[3 * t + a] = anc + ct;  // interface centroid: anchor + min-image offset
    if (!pp) continue;  // throat geometry only
    const long long Dc = (long long)round((ancj - anc) / (double)s[a]) -
                         (long long)round(ct / (double)s[a] + 1e-6);
    const long long k = (long long)round((double)Dc / N[a]);
    macro += gp[a] * ((pjp[a] - pip[a]) - L[a] * (double)k);
  }
  if (pp) dp[t] = (pp[li - 1] - pp[lj - 1]) - macro;
}

// Per-throat inputs of the conductance surrogate (the host statement is throat_inputs in
// surrogate.py): the interface centroid in voxel coordinates (X, Y, Z) and the raw scalars
// log area, log radius of each pore, log min-image separation of the pore centres, computed in
// FP64 and rounded to FP32.
__global__ void k_throat_features(const int32_t* __restrict__ throats, int64_t nt,
                                  const pnm_pore_t* __restrict__ pores,
                                  const double* __restrict__ Ar, const double* __restrict__ tc,
                                  Geo g, float ox, float oy, float oz, float sx, float sy,
                                  float sz, float* __restrict__ centers,
                                  float* __restrict__ raw) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const pnm_pore_t pi = pores[throats[2 * t] - 1], pj = pores[throats[2 * t + 1] - 1];
  const double o[3] = {ox, oy, oz}, s[3] = {sx, sy, sz};
  const double L[3] = {(double)g.nx * s[0], (double)g.ny * s[1], (double)g.nz * s[2]};
  const double dpos[3] = {(double)pj.x - (double)pi.x, (double)pj.y - (double)pi.y,
                          (double)pj.z - (double)pi.z};
  double q = 0.0;
  for (int a = 0; a < 3; ++a) {
    const double d = dpos[a] - L[a] * rint(dpos[a] / L[a]);  // numpy round: half to even
    q += d * d;
    centers[3 * t + a] = (float)((tc[3 * t + a] - o[a]) / s[a]);
  }
  raw[4 * t + 0] = (float)log(fmax(Ar[t], 1e-12));
  raw[4 * t + 1] = (float)log(fmax((double)pi.radius, 1e-6));
  raw[4 * t + 2] = (float)log(fmax((double)pj.radius, 1e-6));
  raw[4 * t + 3] = (float)log(fmax(sqrt(q), 1e-6));
}

// ------------------------------------------------------------------------------------------
template <class T>
static T* host_array(Ctx* c, const T* dev, int64_t count) {
  T* h = static_cast<T*>(std::malloc(sizeof(T) * (size_t)std::max<int64_t>(count, 1)));
  if (!h) throw std::bad_alloc();
  if (count > 0)
    PNM_HIP(hipMemcpyAsync(h, dev, sizeof(T) * count, hipMemcpyDeviceToHost, c->stream));
  return h;
}

int64_t face_ccl(Ctx* c, const uint32_t* list, int64_t cnt, Geo g, const int64_t* fpair,
                 const uint8_t* itfm, int64_t* par, int64_t* fpar) {
  if (cnt <= 0) return 0;
  PNM_LAUNCH(c, "k_face_union", k_face_union, dim3(cdiv(cnt, 256)), dim3(256), 0, list, cnt, g,
             fpair, itfm, par);
  PNM_LAUNCH(c, "k_face_flatten", k_face_flatten, dim3(cdiv(cnt, 256)), dim3(256), 0, list, cnt,
             par, fpar);
  return cnt;
}

void flow_device(Ctx* c, const float* sdf, Vol v, const float org[3], const float spc[3],
                 const double* u, const double* vv, const double* w, const double* p,
                 const double* ox, const double* oy, const double* oz, const double grad_p[3],
                 pnm_flow_result_t* out, const GeoMode* geo) {
  const int64_t n = v.n;
  const Geo g{v.nx, v.ny, v.nz};
  const bool kgeo = geo != nullptr;
  if (3 * n > (int64_t)0xffffffffLL)
    fail(PNM_EINVAL, "extract_network_flow: 3*Nx*Ny*Nz must fit in 32 bits (face ids)");
  uint64_t* ctr = c->arena.get_n<uint64_t>(8);
  fill_u64(c, ctr, 0, 8);
  unsigned long long* uctr = reinterpret_cast<unsigned long long*>(ctr);
  const int32_t* seg = nullptr;
  const uint32_t* peaks = nullptr;
  int64_t np = 0, npeaks = 0;
  if (kgeo && geo->seg) {
    // a given segmentation: its largest label and the peak list of the extract_pores predicate
    seg = geo->seg;
    const int64_t nb = pre_blocks(v), scap = pre_stage_cap();
    uint32_t* stage = c->arena.get_n<uint32_t>((size_t)(nb * scap));
    uint32_t* cnt = c->arena.get_n<uint32_t>((size_t)nb);
    uint32_t* pk = c->arena.get_n<uint32_t>((size_t)n);
    PreOut o{};
    o.peak_stage = stage;
    o.peak_cnt = cnt;
    launch_pre(c, sdf, v, 0.0f, o);
    compact_staged(c, stage, cnt, nb, scap, pk, ctr + 1);
    PNM_LAUNCH(c, "k_label_max", k_label_max, dim3((unsigned)std::min<int64_t>(cdiv(n, 256), 4096)),
               dim3(256), 0, seg, n, reinterpret_cast<int*>(ctr));
    uint64_t h2[2];
    readback(c, ctr, 2, h2);
    np = (int64_t)(int32_t)(uint32_t)h2[0];
    peaks = pk;
    npeaks = (int64_t)h2[1];
    fill_u64(c, ctr, 0, 2);
  } else {
    int32_t* sg = c->arena.get_n<int32_t>((size_t)n);
    const SegPeaks sp = segment_with_peaks(c, sdf, v, spc, sg, nullptr);
    seg = sg;
    np = sp.n_pores;
    peaks = sp.peaks;
    npeaks = sp.npeaks;
  }
  if (np <= 0) {  // np == 0 returns empty in the reference; np < 0 cannot happen here
    PNM_HIP(hipStreamSynchronize(c->stream));
    return;
  }

  // peaks and centres
  int32_t* peak = c->arena.get_n<int32_t>((size_t)np);
  fill_u32(c, reinterpret_cast<uint32_t*>(peak), 0u, np);
  PNM_LAUNCH(c, "k_flow_peaks", k_flow_peaks, dim3(cdiv(std::max<int64_t>(npeaks, 1), 256)),
             dim3(256), 0, sdf, seg, peaks, npeaks, peak);
  pnm_pore_t* pores = c->arena.get_n<pnm_pore_t>((size_t)np);
  launch_centroids(c, sdf, v, org, spc, reinterpret_cast<const uint32_t*>(peak), np, pores);
  double* pp = nullptr;
  if (!kgeo) {
    pp = c->arena.get_n<double>((size_t)np);
    PNM_LAUNCH(c, "k_pore_pressure", k_pore_pressure, dim3(cdiv(np, 256)), dim3(256), 0, pores,
               np, g, org[0], org[1], org[2], spc[0], spc[1], spc[2], p, pp);
  }

  // flow basins. Geometry only: every open face joins two sdf>0 cells, and the basin of an
  // sdf>0 cell is its segmentation label (both are the label at the end of its ascent), so the
  // segmentation stands for the basins.
  const int32_t* lab = seg;
  if (!kgeo) {
    int32_t* next_all = c->arena.get_n<int32_t>((size_t)n);
    int32_t* fl = c->arena.get_n<int32_t>((size_t)n);
    launch_next(c, sdf, v, next_all, kWalkAll);
    PNM_LAUNCH(c, "k_flow_lab", k_flow_lab, dim3(cdiv(n, 256 * kFlowLabChains)), dim3(256), 0, n,
               next_all, seg, fl);
    lab = fl;
  }

  // faces: labels, pair keys and the face-ordered lists (core, film, all interface faces,
  // residual records), staged per block and compacted in block order
  const int64_t nf = 3 * n;
  int64_t* fpar = c->arena.get_n<int64_t>((size_t)nf);
  int64_t* fpair = c->arena.get_n<int64_t>((size_t)nf);
  int64_t* par = c->arena.get_n<int64_t>((size_t)nf);
  const int64_t nbf = cdiv(n, kFaceBlock);
  uint32_t* st_core = c->arena.get_n<uint32_t>((size_t)(nbf * kFaceCap));
  uint32_t* st_film = c->arena.get_n<uint32_t>((size_t)(nbf * kFaceCap));
  uint32_t* st_itf = c->arena.get_n<uint32_t>((size_t)(nbf * kFaceCap));
  uint64_t* st_res = kgeo ? nullptr : c->arena.get_n<uint64_t>((size_t)(nbf * kResCap));
  uint32_t* cnt3 = c->arena.get_n<uint32_t>((size_t)(3 * nbf));
  uint32_t* cnt_res = c->arena.get_n<uint32_t>((size_t)nbf);
  uint8_t* itfm = c->arena.get_n<uint8_t>((size_t)n);  // per cell: which +d faces are interfaces
  if (kgeo)
    PNM_LAUNCH(c, "k_face_init", k_face_init<true>, dim3((unsigned)nbf), dim3(kFaceBlock), 0,
               sdf, lab, g, n, nullptr, nullptr, nullptr, fpar, fpair, par, st_core, st_film,
               st_itf, cnt3, st_res, cnt_res, itfm);
  else
    PNM_LAUNCH(c, "k_face_init", k_face_init<false>, dim3((unsigned)nbf), dim3(kFaceBlock), 0,
               sdf, lab, g, n, ox, oy, oz, fpar, fpair, par, st_core, st_film, st_itf, cnt3,
               st_res, cnt_res, itfm);
  uint32_t* core_list = c->arena.get_n<uint32_t>((size_t)nf);
  uint32_t* film_list = c->arena.get_n<uint32_t>((size_t)nf);
  uint32_t* itf_list = c->arena.get_n<uint32_t>((size_t)nf);
  uint64_t* res_rec = kgeo ? nullptr : c->arena.get_n<uint64_t>((size_t)(2 * nf));
  {
    const uint32_t* st[3] = {st_core, st_film, st_itf};
    uint32_t* outs[3] = {core_list, film_list, itf_list};
    uint64_t* tots[3] = {ctr + 0, ctr + 1, ctr + 6};
    compact_staged_multi(c, 3, st, cnt3, nbf, kFaceCap, outs, tots);
    if (!kgeo) compact_staged64(c, st_res, cnt_res, nbf, kResCap, res_rec, ctr + 2);
  }
  uint64_t h[8];
  readback(c, ctr, 7, h);
  const int64_t ncore = (int64_t)h[0], nfilm = (int64_t)h[1], nres = (int64_t)h[2];
  const int64_t nitf = (int64_t)h[6];

  // core tier
  face_ccl(c, core_list, ncore, g, fpair, itfm, par, fpar);
  // film attachment to a fixpoint
  // (two sweeps per read-back: a sweep at the fixpoint changes nothing, and every sweep that
  // changes something lowers at least one film label, so nfilm + 2 sweeps always suffice)
  if (nfilm > 0) {
    int64_t sweeps = 0;
    while (true) {
      fill_u64(c, ctr + 3, 0, 1);
      for (int k = 0; k < 2; ++k)
        PNM_LAUNCH(c, "k_film_attach", k_film_attach, dim3(cdiv(nfilm, 256)), dim3(256), 0,
                   film_list, nfilm, g, fpair, itfm, fpar, uctr + 3);
      uint64_t ch[2];
      readback(c, ctr + 3, 1, ch);
      if (ch[0] == 0) break;
      sweeps += 2;
      if (sweeps > nfilm + 2) fail(PNM_ENOCONVERGE, "film attachment did not converge");
    }
  }
  // leftover tier
  uint32_t* left = c->arena.get_n<uint32_t>((size_t)std::max<int64_t>(nfilm, 1));
  if (ncore > 0)
    PNM_LAUNCH(c, "k_reset_par", k_reset_par, dim3(cdiv(ncore, 256)), dim3(256), 0, core_list,
               ncore, par);
  int64_t nleft = 0;
  if (nfilm > 0) {
    fill_u64(c, ctr + 4, 0, 1);
    PNM_LAUNCH(c, "k_leftover_init", k_leftover_init, dim3(cdiv(nfilm, 256)), dim3(256), 0,
               film_list, nfilm, fpar, par, left, uctr + 4);
    readback(c, ctr + 4, 1, h);
    nleft = (int64_t)h[0];
    face_ccl(c, left, nleft, g, fpair, itfm, par, fpar);
  }

  // patch roots -> throats ordered by (pair, root face)
  uint64_t* root_fid = c->arena.get_n<uint64_t>((size_t)std::max<int64_t>(ncore + nfilm, 1));
  {
    const int64_t nbc = cdiv(ncore, 256), nbm = cdiv(nfilm, 256), nbr = nbc + nbm;
    uint64_t* rstage = c->arena.get_n<uint64_t>((size_t)std::max<int64_t>(nbr * 256, 1));
    uint32_t* rcnt = c->arena.get_n<uint32_t>((size_t)std::max<int64_t>(nbr, 1));
    if (nbr > 0) {
      PNM_LAUNCH(c, "k_face_roots", k_face_roots, dim3((unsigned)nbr), dim3(256), 0, core_list,
                 ncore, film_list, nfilm, nbc, fpar, rstage, rcnt);
      compact_staged64(c, rstage, rcnt, nbr, 256, root_fid, ctr + 5);
    } else {
      fill_u64(c, ctr + 5, 0, 1);
    }
  }
  readback(c, ctr + 5, 1, h);
  const int64_t nt = (int64_t)h[0];
  int32_t* throats = c->arena.get_n<int32_t>((size_t)std::max<int64_t>(2 * nt, 2));
  double* Q = kgeo ? nullptr : c->arena.get_n<double>((size_t)std::max<int64_t>(nt, 1));
  double* Ar = c->arena.get_n<double>((size_t)std::max<int64_t>(nt, 1));
  double* Cx = c->arena.get_n<double>((size_t)std::max<int64_t>(nt, 1));
  double* Cy = c->arena.get_n<double>((size_t)std::max<int64_t>(nt, 1));
  double* Cz = c->arena.get_n<double>((size_t)std::max<int64_t>(nt, 1));
  double* dp = kgeo ? nullptr : c->arena.get_n<double>((size_t)std::max<int64_t>(nt, 1));
  double* tc = c->arena.get_n<double>((size_t)std::max<int64_t>(3 * nt, 1));
  double* resid = kgeo ? nullptr : c->arena.get_n<double>((size_t)np);
  FluxArgs fa{g, lab, u, vv, w, kgeo ? nullptr : ox, kgeo ? nullptr : oy, kgeo ? nullptr : oz,
              (double)spc[1] * spc[2], (double)spc[0] * spc[2], (double)spc[0] * spc[1]};
  if (nt > 0) {
    uint32_t* order = c->arena.get_n<uint32_t>((size_t)nt);
    uint64_t* key = c->arena.get_n<uint64_t>((size_t)nt);
    // order = sort roots by fid, then stably by compact pair key
    PNM_LAUNCH(c, "k_iota_u32", k_iota_u32, dim3(cdiv(nt, 256)), dim3(256), 0, order, nt);
    {
      uint64_t* fk = c->arena.get_n<uint64_t>((size_t)nt);
      PNM_HIP(hipMemcpyAsync(fk, root_fid, sizeof(uint64_t) * nt, hipMemcpyDeviceToDevice,
                             c->stream));
      radix_sort_u64(c, fk, order, nt, 0, nbits64((uint64_t)nf));
    }
    const int B = std::max(1, nbits64((uint64_t)np));
    PNM_LAUNCH(c, "k_root_pairkeys", k_root_pairkeys, dim3(cdiv(nt, 256)), dim3(256), 0, root_fid,
               order, nt, fpair, B, key);
    radix_sort_u64(c, key, order, nt, 0, 2 * B);
    PNM_LAUNCH(c, "k_throat_slots", k_throat_slots, dim3(cdiv(nt, 256)), dim3(256), 0, root_fid,
               order, nt, fpair, par, throats);
    // throat records over every interface face, in face order; a stable sort on the slot
    // keeps face order within each throat
    uint64_t* trec = c->arena.get_n<uint64_t>((size_t)nitf);
    PNM_LAUNCH(c, "k_throat_records", k_throat_records, dim3(cdiv(nitf, 256)), dim3(256), 0,
               itf_list, nitf, fpar, par, trec);
    radix_sort_u64(c, trec, nullptr, nitf, 32, 32 + nbits64((uint64_t)nt));
    int64_t* ts = c->arena.get_n<int64_t>((size_t)nt);
    int64_t* te = c->arena.get_n<int64_t>((size_t)nt);
    fill_u64(c, reinterpret_cast<uint64_t*>(ts), 0, nt);
    fill_u64(c, reinterpret_cast<uint64_t*>(te), 0, nt);
    PNM_LAUNCH(c, "k_seg_bounds", k_seg_bounds, dim3(cdiv(nitf, 256)), dim3(256), 0, trec, nitf,
               ts, te);
    double* terms = c->arena.get_n<double>((size_t)(5 * nitf));
    if (kgeo) {
      PNM_LAUNCH(c, "k_throat_terms", k_throat_terms<true>, dim3(cdiv(nitf, 256)), dim3(256), 0,
                 trec, nitf, fa, peak, org[0], org[1], org[2], spc[0], spc[1], spc[2], nullptr,
                 terms + nitf, terms + 2 * nitf, terms + 3 * nitf, terms + 4 * nitf);
      TermSets<4> T{};
      double* outs[4] = {Ar, Cx, Cy, Cz};
      for (int k = 0; k < 4; ++k) {
        T.in[k] = terms + (k + 1) * nitf;
        T.out[k] = outs[k];
      }
      PNM_LAUNCH(c, "k_key_sums", k_key_sums<4>, dim3(cdiv(nt, 4)), dim3(256), 0, ts, te, nt, T);
    } else {
      PNM_LAUNCH(c, "k_throat_terms", k_throat_terms<false>, dim3(cdiv(nitf, 256)), dim3(256), 0,
                 trec, nitf, fa, peak, org[0], org[1], org[2], spc[0], spc[1], spc[2], terms,
                 terms + nitf, terms + 2 * nitf, terms + 3 * nitf, terms + 4 * nitf);
      TermSets<5> T{};
      double* outs[5] = {Q, Ar, Cx, Cy, Cz};
      for (int k = 0; k < 5; ++k) {
        T.in[k] = terms + k * nitf;
        T.out[k] = outs[k];
      }
      PNM_LAUNCH(c, "k_key_sums", k_key_sums<5>, dim3(cdiv(nt, 4)), dim3(256), 0, ts, te, nt, T);
    }
    PNM_LAUNCH(c, "k_throat_dp", k_throat_dp, dim3(cdiv(nt, 256)), dim3(256), 0, throats, nt,
               pores, pp, peak, Ar, Cx, Cy, Cz, g, org[0], org[1], org[2], spc[0], spc[1],
               spc[2], grad_p[0], grad_p[1], grad_p[2], dp, tc);
    if (kgeo && geo->centers && geo->raw && nt <= geo->cap)
      PNM_LAUNCH(c, "k_throat_features", k_throat_features, dim3(cdiv(nt, 256)), dim3(256), 0,
                 throats, nt, pores, Ar, tc, g, org[0], org[1], org[2], spc[0], spc[1], spc[2],
                 geo->centers, geo->raw);
  }
  if (kgeo) {
    out->np = np;
    out->nt = nt;
    out->pores = host_array(c, pores, np);
    out->throats = host_array(c, throats, 2 * nt);
    out->throat_area = host_array(c, Ar, nt);
    out->throat_center = host_array(c, tc, 3 * nt);
    PNM_HIP(hipStreamSynchronize(c->stream));
    return;
  }
  // residual: records in face order, stably sorted by pore
  {
    int64_t* rs = c->arena.get_n<int64_t>((size_t)np);
    int64_t* re = c->arena.get_n<int64_t>((size_t)np);
    fill_u64(c, reinterpret_cast<uint64_t*>(rs), 0, np);
    fill_u64(c, reinterpret_cast<uint64_t*>(re), 0, np);
    double* rterm = c->arena.get_n<double>((size_t)std::max<int64_t>(nres, 1));
    if (nres > 0) {
      radix_sort_u64(c, res_rec, nullptr, nres, 32, 32 + nbits64((uint64_t)np));
      PNM_LAUNCH(c, "k_seg_bounds", k_seg_bounds, dim3(cdiv(nres, 256)), dim3(256), 0, res_rec,
                 nres, rs, re);
      PNM_LAUNCH(c, "k_resid_terms", k_resid_terms, dim3(cdiv(nres, 256)), dim3(256), 0, res_rec,
                 nres, fa, rterm);
    }
    TermSets<1> T{};
    T.in[0] = rterm;
    T.out[0] = resid;
    PNM_LAUNCH(c, "k_key_sums", k_key_sums<1>, dim3(cdiv(np, 4)), dim3(256), 0, rs, re, np, T);
  }

  out->np = np;
  out->nt = nt;
  out->pores = host_array(c, pores, np);
  out->pore_pressure = host_array(c, pp, np);
  out->pore_residual = host_array(c, resid, np);
  out->throats = host_array(c, throats, 2 * nt);
  out->throat_flow = host_array(c, Q, nt);
  out->throat_area = host_array(c, Ar, nt);
  out->throat_dp = host_array(c, dp, nt);
  out->throat_center = host_array(c, tc, 3 * nt);
  PNM_HIP(hipStreamSynchronize(c->stream));
}

}  // namespace pnm
