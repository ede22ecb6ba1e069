// k_stokes.hip — the project's MAC Stokes solver (data generation for extract_network_flow).
//
// Periodic MAC grid, flow's layout (pore_extraction.hpp:476-477): u(i,j,k) on the -x face of
// cell (i,j,k), v on the -y face, w on the -z face, p at the cell centre. A cell is fluid iff
// sdf > 0; a face is open iff both cells it separates are fluid (binary openness); closed faces
// carry zero velocity. With viscosity mu, body force f and spacing h, the discrete system
//      mu * (-Lap) u + G p = f      on open faces
//                     G^T u = 0      in fluid cells     (G^T u = -div u)
// with G p = (p[c] - p[c - e_d]) / h_d on the -d face of c, is symmetric: [[A, G], [G^T, 0]].
// It is solved by preconditioned MINRES (Paige & Saunders 1975) in FP64, preconditioner
// diag(2 mu sum 1/h_d^2 on faces, 1/mu on cells). The scalar recurrences run on the device, so
// an iteration is five kernels and no host synchronisation; the host reads the residual
// estimate every 32 iterations.
#include "k_common.hpp"

namespace pnm {

struct V4 {
  double* c[4];  // u, v, w, p
};
struct CV4 {
  const double* c[4];
};
static CV4 cv(const V4& a) { return CV4{{a.c[0], a.c[1], a.c[2], a.c[3]}}; }

struct Grid {
  int nx, ny, nz;
  double ih2[3];  // 1/h_d^2
  double ih[3];   // 1/h_d
  double mu;
};

constexpr int kRedBlocks = 1024;  // partial sums per reduction (fixed: deterministic order)

// Masks: bit d = the -d face of the cell is open, bit 3 = the cell is fluid.
__global__ void k_st_masks(const float* __restrict__ sdf, int nx, int ny, int nz, int64_t n,
                           uint8_t* __restrict__ m) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int x = (int)(c % nx), y = (int)((c / nx) % ny), z = (int)(c / ((int64_t)nx * ny));
  const bool f = sdf[c] > 0.0f;
  const bool fx = sdf[(int64_t)(z * ny + y) * nx + wm1(x, nx)] > 0.0f;
  const bool fy = sdf[(int64_t)(z * ny + wm1(y, ny)) * nx + x] > 0.0f;
  const bool fz = sdf[(int64_t)(wm1(z, nz) * ny + y) * nx + x] > 0.0f;
  m[c] = (uint8_t)((f && fx ? 1 : 0) | (f && fy ? 2 : 0) | (f && fz ? 4 : 0) | (f ? 8 : 0));
}

__global__ void k_st_openness(const uint8_t* __restrict__ m, int64_t n, double* ox, double* oy,
                              double* oz) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const uint8_t b = m[c];
  if (ox) ox[c] = (b & 1) ? 1.0 : 0.0;
  if (oy) oy[c] = (b & 2) ? 1.0 : 0.0;
  if (oz) oz[c] = (b & 4) ? 1.0 : 0.0;
}

// (M x)[c] for the four unknowns of cell c, scaled by `scale`.
__device__ __forceinline__ void st_apply(const CV4& x, const uint8_t* __restrict__ m,
                                         const Grid& g, int64_t c, double scale, double out[4]) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int xi = (int)(c % nx), yi = (int)((c / nx) % ny), zi = (int)(c / ((int64_t)nx * ny));
  const int64_t row = (int64_t)nx * ny;
  const int64_t nb[3][2] = {
      {(int64_t)(zi * ny + yi) * nx + wm1(xi, nx), (int64_t)(zi * ny + yi) * nx + wp1(xi, nx)},
      {(int64_t)(zi * ny + wm1(yi, ny)) * nx + xi, (int64_t)(zi * ny + wp1(yi, ny)) * nx + xi},
      {(int64_t)wm1(zi, nz) * row + (int64_t)yi * nx + xi,
       (int64_t)wp1(zi, nz) * row + (int64_t)yi * nx + xi}};
  const uint8_t b = m[c];
  const double pc = x.c[3][c];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (!((b >> d) & 1)) {
      out[d] = 0.0;
      continue;
    }
    const double* __restrict__ xd = x.c[d];
    const double xc = xd[c];
    double lap = 0.0;
#pragma unroll
    for (int e = 0; e < 3; ++e) lap += (2.0 * xc - xd[nb[e][0]] - xd[nb[e][1]]) * g.ih2[e];
    out[d] = scale * (g.mu * lap + (pc - x.c[3][nb[d][0]]) * g.ih[d]);
  }
  if (b & 8) {
    double gt = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) gt += (x.c[d][c] - x.c[d][nb[d][1]]) * g.ih[d];
    out[3] = scale * gt;
  } else {
    out[3] = 0.0;
  }
}

// Device-side MINRES scalars.
struct MinresState {
  double beta, oldb, alfa, dbar, epsln, oldeps, phibar, cs, sn, delta, phi, denom, beta1;
  int itn;
  int breakdown;
};

__device__ __forceinline__ double block_sum(double v, double* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = blockDim.x >> 1; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  return sh[0];
}

// y = M(Y)/beta - (beta/oldb) R1 (itn >= 2) into R1's storage; V = Y/beta; partial sums of V.y.
__global__ __launch_bounds__(256) void k_ms_a(CV4 Y, V4 R1, V4 V, const uint8_t* __restrict__ m,
                                              Grid g, int64_t n, const MinresState* st,
                                              double* part) {
  __shared__ double sh[256];
  const double inv = 1.0 / st->beta;
  const double c1 = st->itn >= 1 ? st->beta / st->oldb : 0.0;
  double acc = 0.0;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (int64_t)gridDim.x * blockDim.x) {
    double t[4];
    st_apply(Y, m, g, c, inv, t);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double vk = Y.c[k][c] * inv;
      const double yk = t[k] - c1 * R1.c[k][c];
      V.c[k][c] = vk;
      R1.c[k][c] = yk;
      acc += vk * yk;
    }
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_ms_alfa(MinresState* st, const double* part) {
  __shared__ double sh[kRedBlocks];
  const double s = block_sum(part[threadIdx.x], sh);
  if (threadIdx.x == 0) st->alfa = s;
}

// y -= (alfa/beta) R2 (y in R1's storage, now r2); Z = P^{-1} y; partial sums of y.Z.
__global__ __launch_bounds__(256) void k_ms_b(V4 Yn, CV4 R2, V4 Z, const uint8_t* __restrict__ m,
                                              Grid g, int64_t n, const MinresState* st,
                                              double* part) {
  __shared__ double sh[256];
  const double c2 = st->alfa / st->beta;
  const double pu = 1.0 / (2.0 * g.mu * (g.ih2[0] + g.ih2[1] + g.ih2[2])), pp = g.mu;
  double acc = 0.0;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double y = Yn.c[k][c] - c2 * R2.c[k][c];
      const double z = y * (k < 3 ? pu : pp);
      Yn.c[k][c] = y;
      Z.c[k][c] = z;
      acc += y * z;
    }
  }
  (void)m;
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Givens step (Paige & Saunders): new beta, rotation, phi and the w-update coefficients.
__global__ void k_ms_rot(MinresState* st, const double* part) {
  __shared__ double sh[kRedBlocks];
  const double s = block_sum(part[threadIdx.x], sh);
  if (threadIdx.x != 0) return;
  st->oldb = st->beta;
  st->beta = sqrt(s > 0.0 ? s : 0.0);
  st->oldeps = st->epsln;
  const double alfa = st->alfa;
  st->delta = st->cs * st->dbar + st->sn * alfa;
  const double gbar = st->sn * st->dbar - st->cs * alfa;
  st->epsln = st->sn * st->beta;
  st->dbar = -st->cs * st->beta;
  double gamma = sqrt(gbar * gbar + st->beta * st->beta);
  if (gamma < 1e-300) gamma = 1e-300;
  st->cs = gbar / gamma;
  st->sn = st->beta / gamma;
  st->phi = st->cs * st->phibar;
  st->phibar = st->sn * st->phibar;
  st->denom = 1.0 / gamma;
  st->itn += 1;
  if (!(st->beta > 0.0)) st->breakdown = 1;
}

// w = (V - oldeps W1 - delta W2) * denom into W1's storage; X += phi w.
__global__ __launch_bounds__(256) void k_ms_c(CV4 V, V4 W1, CV4 W2, V4 X, int64_t n,
                                              const MinresState* st) {
  const double oe = st->oldeps, de = st->delta, dn = st->denom, ph = st->phi;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double w = (V.c[k][c] - oe * W1.c[k][c] - de * W2.c[k][c]) * dn;
      W1.c[k][c] = w;
      X.c[k][c] += ph * w;
    }
  }
}

// r = b - M x (b = f on open faces), Z = P^{-1} r, partial sums of r.Z and b.b.
__global__ __launch_bounds__(256) void k_ms_init(CV4 X, V4 R, V4 Z, const uint8_t* __restrict__ m,
                                                 Grid g, int64_t n, double f0, double f1,
                                                 double f2, double* part, double* part_b) {
  __shared__ double sh[256];
  const double pu = 1.0 / (2.0 * g.mu * (g.ih2[0] + g.ih2[1] + g.ih2[2])), pp = g.mu;
  const double f[3] = {f0, f1, f2};
  double acc = 0.0, accb = 0.0;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (int64_t)gridDim.x * blockDim.x) {
    double t[4];
    st_apply(X, m, g, c, 1.0, t);
    const uint8_t b = m[c];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double bk = (k < 3 && ((b >> k) & 1)) ? f[k] : 0.0;
      const double r = bk - t[k];
      const double z = r * (k < 3 ? pu : pp);
      R.c[k][c] = r;
      Z.c[k][c] = z;
      acc += r * z;
      accb += bk * bk;
    }
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
  __syncthreads();
  const double sb = block_sum(accb, sh);
  if (threadIdx.x == 0) part_b[blockIdx.x] = sb;
}

__global__ void k_ms_start(MinresState* st, const double* part, const double* part_b,
                           double* bnorm2) {
  __shared__ double sh[kRedBlocks];
  const double s = block_sum(part[threadIdx.x], sh);
  __syncthreads();
  const double sb = block_sum(part_b[threadIdx.x], sh);
  if (threadIdx.x != 0) return;
  const double beta1 = sqrt(s > 0.0 ? s : 0.0);
  st->beta1 = beta1;
  st->beta = beta1;
  st->oldb = 0.0;
  st->alfa = 0.0;
  st->dbar = 0.0;
  st->epsln = 0.0;
  st->oldeps = 0.0;
  st->phibar = beta1;
  st->cs = -1.0;
  st->sn = 0.0;
  st->delta = 0.0;
  st->phi = 0.0;
  st->denom = 0.0;
  st->itn = 0;
  st->breakdown = !(beta1 > 0.0);
  *bnorm2 = sb;
}

// ||b - M x||^2 partial sums (true residual, unpreconditioned).
__global__ __launch_bounds__(256) void k_ms_resid(CV4 X, const uint8_t* __restrict__ m, Grid g,
                                                  int64_t n, double f0, double f1, double f2,
                                                  double* part) {
  __shared__ double sh[256];
  const double f[3] = {f0, f1, f2};
  double acc = 0.0;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (int64_t)gridDim.x * blockDim.x) {
    double t[4];
    st_apply(X, m, g, c, 1.0, t);
    const uint8_t b = m[c];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double bk = (k < 3 && ((b >> k) & 1)) ? f[k] : 0.0;
      const double r = bk - t[k];
      acc += r * r;
    }
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_ms_sum(const double* part, double* out) {
  __shared__ double sh[kRedBlocks];
  const double s = block_sum(part[threadIdx.x], sh);
  if (threadIdx.x == 0) *out = s;
}

static V4 v4_arena(Ctx* c, int64_t n) {
  double* b = c->arena.get_n<double>((size_t)(4 * n));
  return V4{{b, b + n, b + 2 * n, b + 3 * n}};
}

StokesInfo stokes_device(Ctx* c, const float* sdf, Vol v, const float spc[3],
                         const double force[3], double mu, double rtol, int max_iter, double* u,
                         double* vv, double* w, double* p, double* ox, double* oy, double* oz) {
  const int64_t n = v.n;
  Grid g{v.nx, v.ny, v.nz, {}, {}, mu};
  for (int d = 0; d < 3; ++d) {
    const double h = (double)spc[d];
    g.ih[d] = 1.0 / h;
    g.ih2[d] = 1.0 / (h * h);
  }
  uint8_t* m = c->arena.get_n<uint8_t>((size_t)n);
  PNM_LAUNCH(c, "k_st_masks", k_st_masks, dim3(cdiv(n, 256)), dim3(256), 0, sdf, v.nx, v.ny,
             v.nz, n, m);
  if (ox || oy || oz)
    PNM_LAUNCH(c, "k_st_openness", k_st_openness, dim3(cdiv(n, 256)), dim3(256), 0, m, n, ox, oy,
               oz);
  V4 X{{u, vv, w, p}};
  for (int k = 0; k < 4; ++k) PNM_HIP(hipMemsetAsync(X.c[k], 0, sizeof(double) * n, c->stream));
  V4 R1 = v4_arena(c, n), R2 = v4_arena(c, n), Y = v4_arena(c, n), V = v4_arena(c, n);
  V4 W = v4_arena(c, n), W1 = v4_arena(c, n), W2 = v4_arena(c, n);
  for (int k = 0; k < 4; ++k) {
    PNM_HIP(hipMemsetAsync(W1.c[k], 0, sizeof(double) * n, c->stream));
    PNM_HIP(hipMemsetAsync(W2.c[k], 0, sizeof(double) * n, c->stream));
  }
  (void)W;
  double* part = c->arena.get_n<double>(kRedBlocks);
  double* part_b = c->arena.get_n<double>(kRedBlocks);
  double* scal = c->arena.get_n<double>(4);  // [0] |b|^2 [1] |b - Mx|^2
  MinresState* st = reinterpret_cast<MinresState*>(c->arena.get_n<uint8_t>(sizeof(MinresState)));
  const dim3 grid(kRedBlocks), blk(256);
  PNM_LAUNCH(c, "k_ms_init", k_ms_init, grid, blk, 0, cv(X), R1, Y, m, g, n, force[0], force[1],
             force[2], part, part_b);
  PNM_LAUNCH(c, "k_ms_start", k_ms_start, dim3(1), dim3(kRedBlocks), 0, st, part, part_b, scal);
  // r2 = r1 (both the initial residual)
  for (int k = 0; k < 4; ++k)
    PNM_HIP(hipMemcpyAsync(R2.c[k], R1.c[k], sizeof(double) * n, hipMemcpyDeviceToDevice,
                           c->stream));
  StokesInfo info{};
  MinresState hs{};
  PNM_HIP(hipMemcpyAsync(&hs, st, sizeof(MinresState), hipMemcpyDeviceToHost, c->stream));
  PNM_HIP(hipStreamSynchronize(c->stream));
  int it = 0;
  if (hs.beta1 > 0.0 && !hs.breakdown) {
    const double target = rtol * hs.beta1;
    while (it < max_iter) {
      const int chunk = std::min(32, max_iter - it);
      for (int k = 0; k < chunk; ++k) {
        // Y holds P^{-1} r2; R1 becomes y; rotate (r1, r2, y) and (w1, w2)
        PNM_LAUNCH(c, "k_ms_a", k_ms_a, grid, blk, 0, cv(Y), R1, V, m, g, n, st, part);
        PNM_LAUNCH(c, "k_ms_alfa", k_ms_alfa, dim3(1), dim3(kRedBlocks), 0, st, part);
        PNM_LAUNCH(c, "k_ms_b", k_ms_b, grid, blk, 0, R1, cv(R2), Y, m, g, n, st, part);
        PNM_LAUNCH(c, "k_ms_rot", k_ms_rot, dim3(1), dim3(kRedBlocks), 0, st, part);
        PNM_LAUNCH(c, "k_ms_c", k_ms_c, grid, blk, 0, cv(V), W1, cv(W2), X, n, st);
        std::swap(R1, R2);  // r1 <- r2, r2 <- y (held in old r1)
        std::swap(W1, W2);  // w2 <- w (in old w1), w1 <- old w2
      }
      it += chunk;
      PNM_HIP(hipMemcpyAsync(&hs, st, sizeof(MinresState), hipMemcpyDeviceToHost, c->stream));
      PNM_HIP(hipStreamSynchronize(c->stream));
      if (hs.breakdown || hs.phibar <= target) break;
    }
  }
  PNM_LAUNCH(c, "k_ms_resid", k_ms_resid, grid, blk, 0, cv(X), m, g, n, force[0], force[1],
             force[2], part);
  PNM_LAUNCH(c, "k_ms_sum", k_ms_sum, dim3(1), dim3(kRedBlocks), 0, part, scal + 1);
  double hsc[2];
  PNM_HIP(hipMemcpyAsync(hsc, scal, sizeof(hsc), hipMemcpyDeviceToHost, c->stream));
  PNM_HIP(hipStreamSynchronize(c->stream));
  info.iterations = hs.itn;
  info.rel_residual = hsc[0] > 0.0 ? sqrt(hsc[1] / hsc[0]) : 0.0;
  return info;
}

}  // namespace pnm
