// k_surrogate.hip — throat-conductance surrogate: a small 3D CNN trained and run with the
// project's own kernels. No reference counterpart: the plain bookkeeping of
// extract_network_flow (pore_extraction.hpp:463-997) produces its targets g = Q/dp and stays
// its fallback.
//
// Network (parameters FP32, ~44.5k):
//   16^3 SDF patch -> conv3^3 1->16, ReLU, avg-pool 2 -> conv3^3 16->32, ReLU, avg-pool 2
//   -> conv3^3 32->32, ReLU, global average -> [32 features, 4 throat scalars]
//   -> dense 36->64, ReLU -> dense 64->1 = log g.
// Convolutions (zero padding 1) are implicit GEMMs on the MFMA: activations and weights are
// rounded to BF16, products accumulate in FP32 (v_mfma_f32_16x16x16_bf16). Forward
//   z[p][co] = sum_k col[p][k] W[co][k]                 (col = im2col of the input, k = tap*CI+ci)
// and backward
//   dW[co][k] = sum_p dz[p][co] col[p][k]               (split over p, fixed-order reduction)
//   dcol[p][k] = sum_co dz[p][co] W[co][k],  dX = col2im(dcol)  (gather: no atomics)
// all as one "NT" GEMM C[M][N] = A[M][K] . Bt[N][K]^T with K-contiguous BF16 operands.
// Activations are channels-last: [batch][z][y][x][channel]. The dense head, pooling, loss and
// Adam run in FP32.
#include <cmath>

#include "k_common.hpp"

namespace pnm {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Round-to-nearest-even FP32 -> BF16 (bits), NaN kept quiet.
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// ---------------- NT GEMM on the MFMA ----------------
// C_part[s][m][n] = sum_{k in split s} A[m][k] * Bt[n][k]. M, N, K multiples of 16; each wave
// owns one 16x16 tile of one split. Lane l feeds A[m0 + l%16][k + 4*(l/16) .. +3] and
// Bt[n0 + l%16][k + 4*(l/16) .. +3] and receives C[m0 + 4*(l/16) + i][n0 + l%16], i = 0..3.
__global__ __launch_bounds__(256) void k_gemm_nt(const uint16_t* __restrict__ A,
                                                 const uint16_t* __restrict__ Bt,
                                                 float* __restrict__ Cp, int M, int N, int K,
                                                 int kchunk) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int tiles_n = N >> 4;
  const int64_t tile = (int64_t)blockIdx.x * 4 + wv;
  const int64_t ntiles = (int64_t)(M >> 4) * tiles_n;
  if (tile >= ntiles) return;  // whole wave
  const int m0 = (int)(tile / tiles_n) << 4, n0 = (int)(tile % tiles_n) << 4;
  const int s = blockIdx.y;
  const int k_lo = s * kchunk, k_hi = min(K, k_lo + kchunk);
  const int r = lane & 15, q = (lane >> 4) * 4;
  const uint16_t* __restrict__ ap = A + (int64_t)(m0 + r) * K + q;
  const uint16_t* __restrict__ bp = Bt + (int64_t)(n0 + r) * K + q;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k = k_lo; k < k_hi; k += 16) {
    const bf16x4 a = *reinterpret_cast<const bf16x4*>(ap + k);
    const bf16x4 b = *reinterpret_cast<const bf16x4*>(bp + k);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
  float* __restrict__ cp = Cp + (int64_t)s * M * N;
#pragma unroll
  for (int i = 0; i < 4; ++i) cp[(int64_t)(m0 + q + i) * N + n0 + r] = acc[i];
}

// C[m][n] = sum_s Cp[s][m][n] (+ bias[n]) (ReLU), FP32; optional BF16 copy.
__global__ void k_split_reduce(const float* __restrict__ Cp, int splits, int64_t MN, int N,
                               const float* __restrict__ bias, int relu, float* __restrict__ C) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= MN) return;
  float v = 0.f;
  for (int s = 0; s < splits; ++s) v += Cp[(int64_t)s * MN + i];
  if (bias) v += bias[i % N];
  if (relu) v = v > 0.f ? v : 0.f;
  C[i] = v;
}

// im2col with zero padding: col[p][k] (and colT[k][p]) for p = (b, z, y, x) of a D^3 grid,
// k = tap*CI + ci (tap = (dz+1)*9 + (dy+1)*3 + (dx+1)), zero for k >= 27*CI (K padding).
// col[p][k] (and colT[k][p] when wanted) for a tile of 64 positions x 64 columns per block:
// the tile is built in LDS, then written row by row in both layouts, so both stores are
// contiguous across the threads (P is a multiple of 64).
__global__ __launch_bounds__(256) void k_im2col(const float* __restrict__ act, int D, int CI,
                                                int Kp, int64_t P, uint16_t* __restrict__ col,
                                                uint16_t* __restrict__ colT) {
  __shared__ uint16_t T[64][66];
  const int64_t p0 = (int64_t)blockIdx.x * 64;
  const int k0 = blockIdx.y * 64;
  const int D3 = D * D * D;
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int pl = e >> 6, kl = e & 63;
    const int64_t p = p0 + pl;
    const int k = k0 + kl;
    float v = 0.f;
    if (k < 27 * CI) {
      const int tap = k / CI, ci = k - tap * CI;
      const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
      const int64_t b = p / D3;
      const int r = (int)(p - b * D3);
      const int x = r % D + dx, y = (r / D) % D + dy, z = r / (D * D) + dz;
      if (x >= 0 && x < D && y >= 0 && y < D && z >= 0 && z < D)
        v = act[(b * D3 + (int64_t)(z * D + y) * D + x) * CI + ci];
    }
    const uint16_t h = f2bf(v);
    T[pl][kl] = h;
    if (k < Kp) col[p * Kp + k] = h;
  }
  if (!colT) return;  // uniform
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int kl = e >> 6, pl = e & 63;
    const int k = k0 + kl;
    if (k < Kp) colT[(int64_t)k * P + p0 + pl] = T[pl][kl];
  }
}

// dX[b][q][ci] = sum_tap dcol[b][q - off(tap)][tap*CI + ci] over in-range positions.
__global__ void k_col2im(const float* __restrict__ dcol, int D, int CI, int Kp, int64_t P,
                         float* __restrict__ dx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P * CI) return;
  const int64_t q = i / CI;
  const int ci = (int)(i - q * CI);
  const int D3 = D * D * D;
  const int64_t b = q / D3;
  const int r = (int)(q - b * D3);
  const int x = r % D, y = (r / D) % D, z = r / (D * D);
  float acc = 0.f;
  for (int tap = 0; tap < 27; ++tap) {
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, ddx = tap % 3 - 1;
    const int px = x - ddx, py = y - dy, pz = z - dz;
    if (px < 0 || px >= D || py < 0 || py >= D || pz < 0 || pz >= D) continue;
    const int64_t p = b * D3 + (int64_t)(pz * D + py) * D + px;
    acc += dcol[p * Kp + tap * CI + ci];
  }
  dx[i] = acc;
}

// 2x2x2 average pooling, channels-last, D -> D/2.
__global__ void k_pool_fwd(const float* __restrict__ a, int B, int D, int C, float* __restrict__ o) {
  const int h = D / 2;
  const int64_t n = (int64_t)B * h * h * h * C;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = (int)(i % C);
  int64_t t = i / C;
  const int x = (int)(t % h), y = (int)((t / h) % h), z = (int)((t / ((int64_t)h * h)) % h);
  const int64_t b = t / ((int64_t)h * h * h);
  float s = 0.f;
  for (int dz = 0; dz < 2; ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx)
        s += a[((b * D + 2 * z + dz) * D * D + (int64_t)(2 * y + dy) * D + 2 * x + dx) * C + c];
  o[i] = s * 0.125f;
}

// da = pool_bwd(dp) masked by ReLU (a > 0); writes dz in FP32, BF16 [P][C] and BF16 [C][P].
__global__ void k_pool_bwd_relu(const float* __restrict__ dp, const float* __restrict__ a, int B,
                                int D, int C, float* __restrict__ dz, uint16_t* __restrict__ dzb,
                                uint16_t* __restrict__ dzT) {
  const int64_t P = (int64_t)B * D * D * D;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P * C) return;
  const int64_t p = i / C;
  const int c = (int)(i - p * C);
  const int D3 = D * D * D, h = D / 2;
  const int64_t b = p / D3;
  const int r = (int)(p - b * D3);
  const int x = r % D, y = (r / D) % D, z = r / (D * D);
  const float g = dp[((b * h + z / 2) * h * h + (int64_t)(y / 2) * h + x / 2) * C + c] * 0.125f;
  const float v = a[i] > 0.f ? g : 0.f;
  dz[i] = v;
  const uint16_t hb = f2bf(v);
  dzb[i] = hb;
  dzT[(int64_t)c * P + p] = hb;
}

// Global average of a [B][S][C] activation -> feat[B][36] (first C entries); the other 4 are
// the throat scalars.
__global__ void k_gap_concat(const float* __restrict__ a, int B, int S, int C,
                             const float* __restrict__ scal, float* __restrict__ h) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)B * (C + 4)) return;
  const int64_t b = i / (C + 4);
  const int j = (int)(i - b * (C + 4));
  if (j < C) {
    float s = 0.f;
    for (int k = 0; k < S; ++k) s += a[(b * S + k) * C + j];
    h[i] = s / (float)S;
  } else {
    h[i] = scal[b * 4 + (j - C)];
  }
}

// dz3 = (dg / S) masked by ReLU(a3 > 0).
__global__ void k_gap_bwd_relu(const float* __restrict__ dh, const float* __restrict__ a, int B,
                               int S, int C, float* __restrict__ dz, uint16_t* __restrict__ dzb,
                               uint16_t* __restrict__ dzT) {
  const int64_t P = (int64_t)B * S;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P * C) return;
  const int64_t p = i / C;
  const int c = (int)(i - p * C);
  const int64_t b = p / S;
  const float g = dh[b * (C + 4) + c] / (float)S;
  const float v = a[i] > 0.f ? g : 0.f;
  dz[i] = v;
  const uint16_t hb = f2bf(v);
  dzb[i] = hb;
  dzT[(int64_t)c * P + p] = hb;
}

// Dense head forward: r4 = ReLU(W4 h + b4) [B][64]; y = W5 r4 + b5.
__global__ void k_head_fwd(const float* __restrict__ h, const float* __restrict__ W4,
                           const float* __restrict__ b4, const float* __restrict__ W5,
                           const float* __restrict__ b5, int B, float* __restrict__ r4,
                           float* __restrict__ y) {
  __shared__ float sh[64];
  const int b = blockIdx.x, j = threadIdx.x;  // 64 threads
  float s = b4[j];
  for (int k = 0; k < 36; ++k) s += W4[j * 36 + k] * h[(int64_t)b * 36 + k];
  s = s > 0.f ? s : 0.f;
  r4[(int64_t)b * 64 + j] = s;
  sh[j] = W5[j] * s;
  __syncthreads();
  if (j == 0) {
    float t = b5[0];
    for (int k = 0; k < 64; ++k) t += sh[k];
    y[b] = t;
  }
  (void)B;
}

// Loss over the batch, mean over samples, in log g: Huber with threshold delta (delta > 0) or
// squared error (delta <= 0); dy = d(loss)/dy. Written by one block in a fixed order.
__global__ void k_loss(const float* __restrict__ y, const float* __restrict__ t, int B,
                       float delta, float* __restrict__ dy, float* __restrict__ loss) {
  __shared__ float sh[256];
  float acc = 0.f;
  for (int b = threadIdx.x; b < B; b += blockDim.x) {
    const float e = y[b] - t[b];
    float l, g;
    if (delta <= 0.f) {
      l = e * e;
      g = 2.f * e;
    } else if (fabsf(e) <= delta) {
      l = 0.5f * e * e;
      g = e;
    } else {
      l = delta * (fabsf(e) - 0.5f * delta);
      g = e > 0.f ? delta : -delta;
    }
    dy[b] = g / (float)B;
    acc += l;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x >> 1; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0 && loss) *loss = sh[0] / (float)B;
}

// Dense head backward, part 1 (one block per sample): dz4 = (dy W5) masked, dh = W4^T dz4.
__global__ void k_head_bwd_act(const float* __restrict__ dy, const float* __restrict__ r4,
                               const float* __restrict__ W4, const float* __restrict__ W5,
                               float* __restrict__ dz4, float* __restrict__ dh) {
  __shared__ float sz[64];
  const int b = blockIdx.x, j = threadIdx.x;  // 64 threads
  const float g = r4[(int64_t)b * 64 + j] > 0.f ? dy[b] * W5[j] : 0.f;
  dz4[(int64_t)b * 64 + j] = g;
  sz[j] = g;
  __syncthreads();
  if (j < 36) {
    float s = 0.f;
    for (int k = 0; k < 64; ++k) s += W4[k * 36 + j] * sz[k];
    dh[(int64_t)b * 36 + j] = s;
  }
}

// Dense head backward, part 2 (weight gradients, summed over the batch in order):
// threads 0..2303 -> dW4[j][k], 2304..2367 -> db4[j], 2368..2431 -> dW5[j], 2432 -> db5.
__global__ void k_head_bwd_w(const float* __restrict__ dy, const float* __restrict__ r4,
                             const float* __restrict__ h, const float* __restrict__ dz4, int B,
                             float* __restrict__ gW4, float* __restrict__ gb4,
                             float* __restrict__ gW5, float* __restrict__ gb5) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float s = 0.f;
  if (i < 2304) {
    const int j = i / 36, k = i % 36;
    for (int b = 0; b < B; ++b) s += dz4[(int64_t)b * 64 + j] * h[(int64_t)b * 36 + k];
    gW4[i] = s;
  } else if (i < 2368) {
    const int j = i - 2304;
    for (int b = 0; b < B; ++b) s += dz4[(int64_t)b * 64 + j];
    gb4[j] = s;
  } else if (i < 2432) {
    const int j = i - 2368;
    for (int b = 0; b < B; ++b) s += dy[b] * r4[(int64_t)b * 64 + j];
    gW5[j] = s;
  } else if (i == 2432) {
    for (int b = 0; b < B; ++b) s += dy[b];
    gb5[0] = s;
  }
}

// Bias gradient db[c] = sum_p dz[p][c] in a fixed order, in two steps: block j sums the rows
// p of its contiguous range (each thread one channel column of a row slice, coalesced rows),
// then one block per channel adds the block partials in order.
constexpr int kBiasBlocks = 512;
__global__ __launch_bounds__(256) void k_bias_part(const float* __restrict__ dz, int64_t P, int C,
                                                   float* __restrict__ part) {
  __shared__ float sh[256];
  const int rows = 256 / C;  // C divides 256
  const int c = threadIdx.x % C, r = threadIdx.x / C;
  const int64_t per = (P + kBiasBlocks - 1) / kBiasBlocks;
  const int64_t a = (int64_t)blockIdx.x * per, b = min(P, a + per);
  float acc = 0.f;
  for (int64_t p = a + r; p < b; p += rows) acc += dz[p * C + c];
  sh[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x < C) {
    float s = 0.f;
    for (int k = 0; k < rows; ++k) s += sh[k * C + threadIdx.x];
    part[(int64_t)blockIdx.x * C + threadIdx.x] = s;
  }
}
__global__ __launch_bounds__(256) void k_bias_grad(const float* __restrict__ part, int C,
                                                   float* __restrict__ db) {
  __shared__ float sh[256];
  const int c = blockIdx.x;
  float acc = 0.f;
  for (int j = threadIdx.x; j < kBiasBlocks; j += blockDim.x) acc += part[(int64_t)j * C + c];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s2 = blockDim.x >> 1; s2 > 0; s2 >>= 1) {
    if ((int)threadIdx.x < s2) sh[threadIdx.x] += sh[threadIdx.x + s2];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[c] = sh[0];
}

// BF16 copies of a conv weight Wt[co][Kp]: as is (forward Bt) and transposed [Kp][co].
__global__ void k_weight_bf16(const float* __restrict__ W, int CO, int Kp,
                              uint16_t* __restrict__ wb, uint16_t* __restrict__ wT) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= CO * Kp) return;
  const int co = i / Kp, k = i - co * Kp;
  const uint16_t h = f2bf(W[i]);
  wb[i] = h;
  wT[k * CO + co] = h;
}

// Adam (Kingma & Ba 2015) with bias correction, FP32.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// 16^3 SDF patch (periodic, nearest voxel) centred on each throat centre (voxel coordinates,
// xyz), scaled by 1/scale.
__global__ void k_patches(const float* __restrict__ sdf, int nx, int ny, int nz,
                          const float* __restrict__ centers, int B, float scale,
                          float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)B * 4096) return;
  const int64_t b = i >> 12;
  const int r = (int)(i & 4095);
  const int px = r & 15, py = (r >> 4) & 15, pz = r >> 8;
  const int cx = (int)floorf(centers[b * 3 + 0]) - 7, cy = (int)floorf(centers[b * 3 + 1]) - 7,
            cz = (int)floorf(centers[b * 3 + 2]) - 7;
  const int x = wrap_any(cx + px, nx), y = wrap_any(cy + py, ny), z = wrap_any(cz + pz, nz);
  out[i] = sdf[((int64_t)z * ny + y) * nx + x] / scale;
}

// ---------------- host orchestration ----------------
struct ConvSpec {
  int D, CI, CO, Kp;
};
static const ConvSpec kConv[3] = {{16, 1, 16, 32}, {8, 16, 32, 432}, {4, 32, 32, 864}};

// parameter offsets in the flat FP32 vector
struct SurLayout {
  int64_t w[3], b[3], W4, b4, W5, b5, total;
};
static SurLayout sur_layout() {
  SurLayout L{};
  int64_t o = 0;
  for (int l = 0; l < 3; ++l) {
    L.w[l] = o;
    o += (int64_t)kConv[l].CO * kConv[l].Kp;
    L.b[l] = o;
    o += kConv[l].CO;
  }
  L.W4 = o;
  o += 64 * 36;
  L.b4 = o;
  o += 64;
  L.W5 = o;
  o += 64;
  L.b5 = o;
  o += 1;
  L.total = o;
  return L;
}

int64_t surrogate_param_count() { return sur_layout().total; }

// He-normal initial weights (std::mt19937_64 + Box-Muller), zero biases; padded K columns zero.
void surrogate_init(uint64_t seed, float* host) {
  const SurLayout L = sur_layout();
  std::fill(host, host + L.total, 0.0f);
  uint64_t s = seed ? seed : 0x9e3779b97f4a7c15ull;
  auto next = [&]() {  // splitmix64
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  };
  auto normal = [&]() {
    const double u1 = ((next() >> 11) + 1.0) / 9007199254740993.0;
    const double u2 = (next() >> 11) / 9007199254740992.0;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  };
  for (int l = 0; l < 3; ++l) {
    const ConvSpec& c = kConv[l];
    const double sd = std::sqrt(2.0 / (27.0 * c.CI));
    for (int co = 0; co < c.CO; ++co)
      for (int k = 0; k < 27 * c.CI; ++k) host[L.w[l] + (int64_t)co * c.Kp + k] = (float)(sd * normal());
  }
  for (int i = 0; i < 64 * 36; ++i) host[L.W4 + i] = (float)(std::sqrt(2.0 / 36.0) * normal());
  for (int i = 0; i < 64; ++i) host[L.W5 + i] = (float)(std::sqrt(1.0 / 64.0) * normal());
}

// Activations of the last forward pass (kept for backward) live in the context's surrogate
// workspace; their pointers are remembered in SurActs.
struct SurActs {
  int B = 0;
  float* x0 = nullptr;
  uint16_t* col[3] = {};
  uint16_t* colT[3] = {};
  float* a[3] = {};   // post-ReLU conv outputs [P][CO]
  float* p[2] = {};   // pooled
  float* h = nullptr;  // [B][36]
  float* r4 = nullptr;
  float* y = nullptr;
};

static SurActs& acts(Ctx* c) {
  if (!c->sur_acts) c->sur_acts = new SurActs();
  return *static_cast<SurActs*>(c->sur_acts);
}

static void gemm_nt(Ctx* c, const uint16_t* A, const uint16_t* Bt, float* C, int64_t M, int N,
                    int K, const float* bias, bool relu) {
  // split K so that a launch has at least ~2048 waves
  const int64_t tiles = (M / 16) * (N / 16);
  int splits = 1;
  while (tiles * splits < 2048 && K / (16 * splits * 2) >= 4) splits *= 2;
  const int kchunk = ((K / 16 + splits - 1) / splits) * 16;
  splits = (K + kchunk - 1) / kchunk;
  float* Cp = splits == 1 && !bias && !relu ? C : c->arena.get_n<float>((size_t)(splits * M * N));
  PNM_LAUNCH(c, "k_gemm_nt", k_gemm_nt, dim3(cdiv(tiles, 4), splits), dim3(256), 0, A, Bt, Cp,
             (int)M, N, K, kchunk);
  if (Cp != C)
    PNM_LAUNCH(c, "k_split_reduce", k_split_reduce, dim3(cdiv(M * N, 256)), dim3(256), 0, Cp,
               splits, M * N, N, bias, relu ? 1 : 0, C);
}

void surrogate_forward(Ctx* c, int B, const float* patches, const float* scalars,
                       const float* params, float* out) {
  const SurLayout L = sur_layout();
  SurActs& S = acts(c);
  c->sur.reset();
  S.B = B;
  S.x0 = const_cast<float*>(patches);
  const float* in = patches;
  for (int l = 0; l < 3; ++l) {
    const ConvSpec& cs = kConv[l];
    const int64_t P = (int64_t)B * cs.D * cs.D * cs.D;
    S.col[l] = c->sur.get_n<uint16_t>((size_t)(P * cs.Kp));
    S.colT[l] = c->sur.get_n<uint16_t>((size_t)(P * cs.Kp));
    S.a[l] = c->sur.get_n<float>((size_t)(P * cs.CO));
    PNM_LAUNCH(c, "k_im2col", k_im2col, dim3((unsigned)(P / 64), cdiv(cs.Kp, 64)), dim3(256), 0,
               in, cs.D, cs.CI, cs.Kp, P, S.col[l], S.colT[l]);
    uint16_t* wb = c->arena.get_n<uint16_t>((size_t)(cs.CO * cs.Kp));
    uint16_t* wT = c->arena.get_n<uint16_t>((size_t)(cs.CO * cs.Kp));
    PNM_LAUNCH(c, "k_weight_bf16", k_weight_bf16, dim3(cdiv(cs.CO * cs.Kp, 256)), dim3(256), 0,
               params + L.w[l], cs.CO, cs.Kp, wb, wT);
    gemm_nt(c, S.col[l], wb, S.a[l], P, cs.CO, cs.Kp, params + L.b[l], true);
    if (l < 2) {
      const int h = cs.D / 2;
      S.p[l] = c->sur.get_n<float>((size_t)((int64_t)B * h * h * h * cs.CO));
      PNM_LAUNCH(c, "k_pool_fwd", k_pool_fwd, dim3(cdiv((int64_t)B * h * h * h * cs.CO, 256)),
                 dim3(256), 0, S.a[l], B, cs.D, cs.CO, S.p[l]);
      in = S.p[l];
    }
  }
  S.h = c->sur.get_n<float>((size_t)B * 36);
  S.r4 = c->sur.get_n<float>((size_t)B * 64);
  PNM_LAUNCH(c, "k_gap_concat", k_gap_concat, dim3(cdiv((int64_t)B * 36, 256)), dim3(256), 0,
             S.a[2], B, 64, 32, scalars, S.h);
  PNM_LAUNCH(c, "k_head_fwd", k_head_fwd, dim3(B), dim3(64), 0, S.h, params + L.W4,
             params + L.b4, params + L.W5, params + L.b5, B, S.r4, out);
  S.y = out;
}

void surrogate_backward(Ctx* c, int B, const float* dy, const float* params, float* grads) {
  const SurLayout L = sur_layout();
  SurActs& S = acts(c);
  if (S.B != B || !S.h) fail(PNM_EINVAL, "backward needs the forward pass of the same batch");
  float* dz4 = c->arena.get_n<float>((size_t)B * 64);
  float* dh = c->arena.get_n<float>((size_t)B * 36);
  PNM_LAUNCH(c, "k_head_bwd_act", k_head_bwd_act, dim3(B), dim3(64), 0, dy, S.r4, params + L.W4,
             params + L.W5, dz4, dh);
  PNM_LAUNCH(c, "k_head_bwd_w", k_head_bwd_w, dim3(cdiv(2433, 256)), dim3(256), 0, dy, S.r4, S.h,
             dz4, B, grads + L.W4, grads + L.b4, grads + L.W5, grads + L.b5);
  float* dnext = nullptr;  // gradient w.r.t. the pooled input of layer l+1
  for (int l = 2; l >= 0; --l) {
    const ConvSpec& cs = kConv[l];
    const int64_t P = (int64_t)B * cs.D * cs.D * cs.D;
    float* dz = c->arena.get_n<float>((size_t)(P * cs.CO));
    uint16_t* dzb = c->arena.get_n<uint16_t>((size_t)(P * cs.CO));
    uint16_t* dzT = c->arena.get_n<uint16_t>((size_t)(P * cs.CO));
    if (l == 2)
      PNM_LAUNCH(c, "k_gap_bwd_relu", k_gap_bwd_relu, dim3(cdiv(P * cs.CO, 256)), dim3(256), 0, dh,
                 S.a[2], B, 64, cs.CO, dz, dzb, dzT);
    else
      PNM_LAUNCH(c, "k_pool_bwd_relu", k_pool_bwd_relu, dim3(cdiv(P * cs.CO, 256)), dim3(256), 0,
                 dnext, S.a[l], B, cs.D, cs.CO, dz, dzb, dzT);
    {
      float* part = c->arena.get_n<float>((size_t)kBiasBlocks * cs.CO);
      PNM_LAUNCH(c, "k_bias_part", k_bias_part, dim3(kBiasBlocks), dim3(256), 0, dz, P, cs.CO,
                 part);
      PNM_LAUNCH(c, "k_bias_grad", k_bias_grad, dim3(cs.CO), dim3(256), 0, part, cs.CO,
                 grads + L.b[l]);
    }
    // dW[co][k] = sum_p dzT[co][p] colT[k][p]
    gemm_nt(c, dzT, S.colT[l], grads + L.w[l], cs.CO, cs.Kp, (int)P, nullptr, false);
    if (l > 0) {
      uint16_t* wb = c->arena.get_n<uint16_t>((size_t)(cs.CO * cs.Kp));
      uint16_t* wT = c->arena.get_n<uint16_t>((size_t)(cs.CO * cs.Kp));
      PNM_LAUNCH(c, "k_weight_bf16", k_weight_bf16, dim3(cdiv(cs.CO * cs.Kp, 256)), dim3(256), 0,
                 params + L.w[l], cs.CO, cs.Kp, wb, wT);
      float* dcol = c->arena.get_n<float>((size_t)(P * cs.Kp));
      gemm_nt(c, dzb, wT, dcol, P, cs.Kp, cs.CO, nullptr, false);
      float* dx = c->arena.get_n<float>((size_t)(P * cs.CI));
      PNM_LAUNCH(c, "k_col2im", k_col2im, dim3(cdiv(P * cs.CI, 256)), dim3(256), 0, dcol, cs.D,
                 cs.CI, cs.Kp, P, dx);
      dnext = dx;
    }
  }
}

void surrogate_loss(Ctx* c, int B, const float* y, const float* t, float delta, float* dy,
                    float* loss) {
  PNM_LAUNCH(c, "k_loss", k_loss, dim3(1), dim3(256), 0, y, t, B, delta, dy, loss);
}

void surrogate_adam(Ctx* c, float* params, const float* grads, float* m, float* v, int step,
                    float lr, float b1, float b2, float eps) {
  const int64_t n = surrogate_param_count();
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  PNM_LAUNCH(c, "k_adam", k_adam, dim3(cdiv(n, 256)), dim3(256), 0, params, grads, m, v, n, lr,
             b1, b2, eps, c1, c2);
}

void surrogate_patches(Ctx* c, const float* sdf, Vol v, const float* centers, int B, float scale,
                       float* out) {
  PNM_LAUNCH(c, "k_patches", k_patches, dim3(cdiv((int64_t)B * 4096, 256)), dim3(256), 0, sdf,
             v.nx, v.ny, v.nz, centers, B, scale, out);
}

// ---------------- fused inference ----------------
// The whole network for one throat at a time in one workgroup (8 waves), persistent over the
// throats. Every convolution is an implicit GEMM on v_mfma_f32_16x16x32_bf16 (lane l holds
// rows/columns l % 16 and K elements 8 (l / 16) .. + 7) with both operands in LDS:
//   - weights (84 KB, converted once per workgroup) stored in MFMA fragment order, so a lane
//     reads its 16 bytes of a K-step contiguously;
//   - the throat's activations with their zero halo, channels-last in planes of 8 channels
//     (A1: 2 planes of [10^3][8], A2: 4 planes of [6^3][8]), so the 8 lanes a 16-byte read
//     serves at once cover 128 contiguous bytes.
// The 16^3 patch of the next throat is fetched into registers while the current one is
// computed and stored into LDS as soon as layer 0 has read the current one. Bias, ReLU and the
// 2x pooling are applied to the accumulators and the pooled values rounded to BF16 into the
// next layer's input; layer 2's global average is summed from its accumulators (lane shuffles,
// one partial per z-slice) and the dense head runs on one wave. Rounding points are those of
// the training forward (surrogate_forward); the FP32 sums differ from it only in order.
struct SurOffs {
  int w0, b0, w1, b1, w2, b2, W4, b4, W5, b5;
};

constexpr int SI_THREADS = 512;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ bf16x8 lds8(const uint16_t* p) {
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8*>(p));
}
__device__ __forceinline__ bf16x4 lds4(const uint16_t* p) {
  return *reinterpret_cast<const bf16x4*>(p);
}
// Sum over the wave by DPP: row_shr 1, 2, 4, 8 within each 16-lane row, then row_bcast 15 and
// 31 across rows; lane 63 receives the total.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ float dpp_add(float v) {
  return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), kCtrl, kRowMask,
                                                        0xf, false));
}
__device__ __forceinline__ float wave_sum_dpp(float v) {
  v = dpp_add<0x111, 0xf>(v);  // row_shr:1
  v = dpp_add<0x112, 0xf>(v);  // row_shr:2
  v = dpp_add<0x114, 0xf>(v);  // row_shr:4
  v = dpp_add<0x118, 0xf>(v);  // row_shr:8
  v = dpp_add<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v = dpp_add<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  return v;
}
// FP32 -> BF16 bits with the gfx950 conversion instruction (round to nearest even, as f2bf for
// every finite value)
__device__ __forceinline__ uint16_t f2bf_hw(float f) {
  return __builtin_bit_cast(uint16_t, (__bf16)f);
}
__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
// offset of tap t (z-major 3x3x3) in a haloed grid of edge E, in voxels
__host__ __device__ constexpr int tap_off(int t, int E) {
  return ((t / 9) * E + (t / 3) % 3) * E + t % 3;
}

// One 8-channel plane of the haloed 10^3 layer-0 output and of the haloed 6^3 layer-1 output
// (unpadded: 64 bytes of padding between planes measured 25% slower in layer 1).
constexpr int kA1Plane = 1000 * 8;
constexpr int kA2Plane = 216 * 8;
// The patch with its zero halo: 18 slabs of 18 rows of 18 voxels, rows 19 and slabs 418 BF16
// apart, so that the four lane groups' 16-voxel reads of the layer-0 gather (four different
// taps) fall on different LDS banks.
constexpr int kX0Row = 19, kX0Slab = 418;

__global__ __launch_bounds__(SI_THREADS) void k_sur_infer(
    const float* __restrict__ sdf, int nx, int ny, int nz, const float* __restrict__ centers,
    const float* __restrict__ scal, int nt, float scale, float inv_scale,
    const float* __restrict__ P, SurOffs L, float* __restrict__ out) {
  __shared__ __align__(16) uint16_t W0f[64 * 8];            // [lane][8]
  __shared__ __align__(16) uint16_t W1f[2 * 13 * 64 * 8];   // [col tile][K-step][lane][8]
  __shared__ __align__(16) uint16_t W1t[2 * 64 * 4];        // tap 26 (K = 16): [tile][lane][4]
  __shared__ __align__(16) uint16_t W2f[2 * 27 * 64 * 8];   // [col tile][K-step][lane][8]
  __shared__ __align__(16) uint16_t X0[18 * kX0Slab];
  __shared__ __align__(16) uint16_t A1[2 * kA1Plane];
  __shared__ __align__(16) uint16_t A2[4 * kA2Plane];
  __shared__ __align__(16) float W4t[36 * 64];              // dense 36 -> 64, [in][out]
  __shared__ float Bs[208];  // conv biases 0..79, dense bias 80..143, output weights 144..207
  __shared__ float part[4 * 32];  // layer 2 after ReLU: per z-slice channel sums
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int n = lane & 15, g = lane >> 4;

  // weights as B fragments: lane (n, g) of K-step s holds W[tile * 16 + n][32 s + 8 g .. + 7]
  for (int i = tid; i < 64 * 8; i += SI_THREADS) {
    const int l = i >> 3, k = 8 * (l >> 4) + (i & 7);
    W0f[i] = k < 27 ? f2bf(P[L.w0 + (l & 15) * 32 + k]) : (uint16_t)0;
  }
  for (int i = tid; i < 2 * 13 * 512; i += SI_THREADS) {
    const int l = (i >> 3) & 63, s = (i >> 9) % 13, tile = i / (13 * 512);
    W1f[i] = f2bf(P[L.w1 + (16 * tile + (l & 15)) * 432 + 32 * s + 8 * (l >> 4) + (i & 7)]);
  }
  for (int i = tid; i < 2 * 256; i += SI_THREADS) {
    const int l = (i >> 2) & 63, tile = i >> 8;
    W1t[i] = f2bf(P[L.w1 + (16 * tile + (l & 15)) * 432 + 416 + 4 * (l >> 4) + (i & 3)]);
  }
  for (int i = tid; i < 2 * 27 * 512; i += SI_THREADS) {
    const int l = (i >> 3) & 63, s = (i >> 9) % 27, tile = i / (27 * 512);
    W2f[i] = f2bf(P[L.w2 + (16 * tile + (l & 15)) * 864 + 32 * s + 8 * (l >> 4) + (i & 7)]);
  }
  for (int i = tid; i < 36 * 64; i += SI_THREADS) W4t[i] = P[L.W4 + (i & 63) * 36 + (i >> 6)];
  if (tid < 16) Bs[tid] = P[L.b0 + tid];
  else if (tid < 48) Bs[tid] = P[L.b1 + tid - 16];
  else if (tid < 80) Bs[tid] = P[L.b2 + tid - 48];
  else if (tid < 144) Bs[tid] = P[L.b4 + tid - 80];
  else if (tid < 208) Bs[tid] = P[L.W5 + tid - 144];
  for (int i = tid; i < 18 * kX0Slab; i += SI_THREADS) X0[i] = 0;
  for (int i = tid; i < 2 * kA1Plane; i += SI_THREADS) A1[i] = 0;
  for (int i = tid; i < 4 * kA2Plane; i += SI_THREADS) A2[i] = 0;
  const float b5 = P[L.b5];

  // patch gather: this thread's 8 voxels are (x, y) = (fx, fy), z = fz + 2 j
  const int fx = tid & 15, fy = (tid >> 4) & 15, fz = tid >> 8;
  float pre[8];
  // centre of the throat the next fetch gathers, loaded one phase ahead of it
  float cxyz[3] = {0.f, 0.f, 0.f};
  auto load_center = [&](int t) {
    if (t >= nt) return;
#pragma unroll
    for (int k = 0; k < 3; ++k) cxyz[k] = centers[3 * t + k];
  };
  auto fetch = [&](int t) {
    if (t >= nt) return;
    // the patch corner is wrapped once; the thread's offset (< 16) wraps at most once more
    int x = wrap_any((int)floorf(cxyz[0]) - 7, nx) + fx;
    int y = wrap_any((int)floorf(cxyz[1]) - 7, ny) + fy;
    int z = wrap_any((int)floorf(cxyz[2]) - 7, nz) + fz;
    x = nx >= 16 ? (x >= nx ? x - nx : x) : wrap_any(x, nx);
    y = ny >= 16 ? (y >= ny ? y - ny : y) : wrap_any(y, ny);
    z = nz >= 16 ? (z >= nz ? z - nz : z) : wrap_any(z, nz);
    const float* row = sdf + (int64_t)y * nx + x;
    const int64_t slab = (int64_t)nx * ny;
    int zz[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      zz[j] = z;
      z += 2;
      if (nz >= 16) z = z >= nz ? z - nz : z;  // one wrap at most: z < nz + 2
      else z = wrap_any(z, nz);
    }
    // raw values only: the loads stay in flight until the stash one throat later
#pragma unroll
    for (int j = 0; j < 8; ++j) pre[j] = row[(int64_t)zz[j] * slab];
  };
  auto stash = [&](int t) {
    if (t >= nt) return;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const float v = inv_scale != 0.0f ? pre[j] * inv_scale : pre[j] / scale;  // 2^-k: exact
      X0[(fz + 2 * j + 1) * kX0Slab + (fy + 1) * kX0Row + fx + 1] = f2bf_hw(v);
    }
  };

  // layer-0 A operand: taps 8 g + j of the lane's voxel; taps >= 27 (zero weights) read tap 0
  int off0[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int tp = 8 * g + j;
    off0[j] = tp < 27 ? (tp / 9) * kX0Slab + ((tp / 3) % 3) * kX0Row + tp % 3 : 0;
  }
  // layer-1 A operand: K-step s covers taps 2 s and 2 s + 1; lane group g reads tap
  // 2 s + (g >> 1), channel plane g & 1
  const int t1 = g >> 1;
  const uint16_t* a1lane = A1 + (g & 1) * kA1Plane + ((n >> 3) * 10 + (n & 7)) * 8;
  const uint16_t* a1last = A1 + (g >> 1) * kA1Plane + ((n >> 3) * 10 + (n & 7)) * 8 + 4 * (g & 1);

  const int G = (int)gridDim.x;
  int t = blockIdx.x;
  load_center(t);
  fetch(t);
  load_center(t + G);
  __syncthreads();  // halos are zero before any interior is written
  stash(t);
  fetch(t + G);
  for (; t < nt; t += G) {
    __syncthreads();  // X0 holds throat t; A1, A2 and part are free
    load_center(t + 2 * G);  // used by this iteration's fetch
    // the throat scalars for this iteration's dense head (uniform loads)
    const float sc4[4] = {scal[4 * t], scal[4 * t + 1], scal[4 * t + 2], scal[4 * t + 3]};

    // ---- layer 0: 1 -> 16 channels on 16^3, K = 27 taps (+5 zero); a tile is one x-row.
    //      Wave w computes the 8 pooled rows of pooled slice w. ----
    {
      const bf16x8 b = lds8(W0f + lane * 8);
      const float bias = Bs[n];
      const f32x4 bias4 = {bias, bias, bias, bias};  // the MFMA adds the bias
      const int pz = wave;
#pragma unroll
      for (int py = 0; py < 8; ++py) {
        float v[2][2][4];
#pragma unroll
        for (int dz = 0; dz < 2; ++dz)
#pragma unroll
          for (int dy = 0; dy < 2; ++dy) {
            const uint16_t* base = X0 + (2 * pz + dz) * kX0Slab + (2 * py + dy) * kX0Row + n;
            u16x8 a;
#pragma unroll
            for (int j = 0; j < 8; ++j) a[j] = base[off0[j]];
            const f32x4 acc = mfma32(__builtin_bit_cast(bf16x8, a), b, bias4);
#pragma unroll
            for (int i = 0; i < 4; ++i)  // row 4 g + i = x, column n = channel
              v[dz][dy][i] = acc[i] > 0.f ? acc[i] : 0.f;
          }
#pragma unroll
        for (int h = 0; h < 2; ++h) {  // pooled x = 2 g + h
          float s = 0.f;
#pragma unroll
          for (int dz = 0; dz < 2; ++dz)
#pragma unroll
            for (int dy = 0; dy < 2; ++dy) {
              s += v[dz][dy][2 * h];
              s += v[dz][dy][2 * h + 1];
            }
          A1[(n >> 3) * kA1Plane + (((pz + 1) * 10 + py + 1) * 10 + 2 * g + h + 1) * 8 + (n & 7)] =
              f2bf_hw(s * 0.125f);
        }
      }
    }
    __syncthreads();  // A1 complete; X0 free
    stash(t + G);
    fetch(t + 2 * G);

    // ---- layer 1: 16 -> 32 channels on 8^3, K = 27 taps x 16 channels. A tile is 16
    //      positions (two 8-voxel rows y, y + 1 of one slice); this wave owns the tiles
    //      (z, rows 2 yt, 2 yt + 1) for z in {2 pz, 2 pz + 1} and yt in {py0, py0 + 1}, both
    //      column tiles. ----
    {
      const int pz = wave >> 1, py0 = 2 * (wave & 1);
      const int wbase = (2 * pz * 10 + 2 * py0) * 10 * 8;
      f32x4 acc[2][2][2];
#pragma unroll
      for (int a0 = 0; a0 < 2; ++a0)
#pragma unroll
        for (int a1 = 0; a1 < 2; ++a1)
#pragma unroll
          for (int a2 = 0; a2 < 2; ++a2) acc[a0][a1][a2] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < 13; ++s) {
        const int toff = (t1 ? tap_off(2 * s + 1, 10) : tap_off(2 * s, 10)) * 8;
        const uint16_t* ap = a1lane + wbase + toff;
        const bf16x8 bf0 = lds8(W1f + (0 * 13 + s) * 512 + lane * 8);
        const bf16x8 bf1 = lds8(W1f + (1 * 13 + s) * 512 + lane * 8);
#pragma unroll
        for (int jy = 0; jy < 2; ++jy)
#pragma unroll
          for (int dz = 0; dz < 2; ++dz) {
            const bf16x8 a = lds8(ap + (dz * 100 + jy * 20) * 8);
            acc[jy][dz][0] = mfma32(a, bf0, acc[jy][dz][0]);
            acc[jy][dz][1] = mfma32(a, bf1, acc[jy][dz][1]);
          }
      }
      {  // tap 26 alone: K = 16, lane group g reads channels 4 g .. 4 g + 3
        const uint16_t* ap = a1last + wbase + tap_off(26, 10) * 8;
        const bf16x4 bf0 = lds4(W1t + lane * 4);
        const bf16x4 bf1 = lds4(W1t + 256 + lane * 4);
#pragma unroll
        for (int jy = 0; jy < 2; ++jy)
#pragma unroll
          for (int dz = 0; dz < 2; ++dz) {
            const bf16x4 a = lds4(ap + (dz * 100 + jy * 20) * 8);
            acc[jy][dz][0] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bf0, acc[jy][dz][0],
                                                                      0, 0, 0);
            acc[jy][dz][1] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bf1, acc[jy][dz][1],
                                                                      0, 0, 0);
          }
      }
      // C row 4 g + i: y offset g >> 1, x = 4 (g & 1) + i; the y + 1 row sits 32 lanes up
#pragma unroll
      for (int jy = 0; jy < 2; ++jy)
#pragma unroll
        for (int nt2 = 0; nt2 < 2; ++nt2) {
          const float bias = Bs[16 + 16 * nt2 + n];
          float v[2][4], w[2][4];
#pragma unroll
          for (int dz = 0; dz < 2; ++dz)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const float c = acc[jy][dz][nt2][i] + bias;
              v[dz][i] = c > 0.f ? c : 0.f;
              w[dz][i] = __shfl_xor(v[dz][i], 32, 64);
            }
          if (g < 2) {
            const int py = py0 + jy, c = 16 * nt2 + n;
#pragma unroll
            for (int h = 0; h < 2; ++h) {  // pooled x = 2 g + h
              float s = 0.f;
#pragma unroll
              for (int dz = 0; dz < 2; ++dz) {
                s += v[dz][2 * h];
                s += v[dz][2 * h + 1];
                s += w[dz][2 * h];
                s += w[dz][2 * h + 1];
              }
              A2[(c >> 3) * kA2Plane + (((pz + 1) * 6 + py + 1) * 6 + 2 * g + h + 1) * 8 +
                 (c & 7)] = f2bf_hw(s * 0.125f);
            }
          }
        }
    }
    __syncthreads();  // A2 complete

    // ---- layer 2: 32 -> 32 channels on 4^3, K = 27 taps x 32 channels (one tap per K-step);
    //      one tile per wave: slice z, column tile nt2; then its channel sums over the slice ----
    {
      // A row n is the voxel (y, x) = (n & 3, n >> 2) of slice z: the 8 lanes of a 16-byte
      // read then cover 8 different 16-byte bank groups
      const int z = wave >> 1, nt2 = wave & 1;
      const uint16_t* ap = A2 + g * kA2Plane + ((z * 6 + (n & 3)) * 6 + (n >> 2)) * 8;
      const uint16_t* bp = W2f + nt2 * 27 * 512 + lane * 8;
      const float bias = Bs[48 + 16 * nt2 + n];
      f32x4 acc = {bias, bias, bias, bias};
#pragma unroll
      for (int s = 0; s < 27; ++s)
        acc = mfma32(lds8(ap + tap_off(s, 6) * 8), lds8(bp + s * 512), acc);
      float sum = 0.f;
#pragma unroll
      for (int i = 0; i < 4; ++i) sum += acc[i] > 0.f ? acc[i] : 0.f;  // 4 voxels of the slice
      sum += __shfl_xor(sum, 16, 64);
      sum += __shfl_xor(sum, 32, 64);
      if (g == 0) part[z * 32 + 16 * nt2 + n] = sum;
    }
    __syncthreads();  // part complete

    // ---- global average, scalars, dense head (wave 0) ----
    // every lane forms the 36 inputs itself (broadcast reads of the slice sums); lane j
    // computes hidden unit j, and the output is the wave's sum
    if (wave == 0) {
      float q[4] = {Bs[80 + lane], 0.f, 0.f, 0.f};  // four independent chains
#pragma unroll
      for (int k = 0; k < 32; ++k) {
        const float h = (((part[k] + part[32 + k]) + part[64 + k]) + part[96 + k]) / 64.0f;
        q[k & 3] += W4t[k * 64 + lane] * h;
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) q[k] += W4t[(32 + k) * 64 + lane] * sc4[k];
      float s = (q[0] + q[1]) + (q[2] + q[3]);
      s = s > 0.f ? s : 0.f;
      const float y = wave_sum_dpp(Bs[144 + lane] * s);
      if (lane == 63) out[t] = b5 + y;
    }
  }
}

void surrogate_infer(Ctx* c, const float* sdf, Vol v, const float* centers, const float* scalars,
                     int B, float scale, const float* params, float* out) {
  const SurLayout L = sur_layout();
  const SurOffs o{(int)L.w[0], (int)L.b[0], (int)L.w[1], (int)L.b[1], (int)L.w[2], (int)L.b[2],
                  (int)L.W4, (int)L.b4, (int)L.W5, (int)L.b5};
  int cus = 0;
  PNM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
  const int groups = std::min(B, std::max(cus, 1));
  // a power-of-two scale divides exactly as a multiplication by its inverse
  int ex = 0;
  const float inv = std::frexp(scale, &ex) == 0.5f ? std::ldexp(1.0f, 1 - ex) : 0.0f;
  PNM_LAUNCH(c, "k_sur_infer", k_sur_infer, dim3((unsigned)groups), dim3(SI_THREADS), 0, sdf,
             v.nx, v.ny, v.nz, centers, scalars, B, scale, inv, params, o, out);
}

void surrogate_release(Ctx* c) {
  delete static_cast<SurActs*>(c->sur_acts);
  c->sur_acts = nullptr;
}

}  // namespace pnm
