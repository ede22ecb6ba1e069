/*
 * synth.c — seeded synthetic SDF generators (host C, OpenMP), loaded through ctypes by the
 * benchmark and the tests. Not part of the extraction path.
 *
 * Every generator is deterministic for a given seed (xoshiro256** seeded by splitmix64; serial
 * loops where order matters). Geometry is in voxel units with voxel centres at integer
 * coordinates and periodic (min-image) distances.
 *
 *   pnms_pack      dense non-overlapping periodic sphere packing: random placement, then
 *                  overlap relaxation while the radii are inflated from 60% to 100%
 *   pnms_boolean   Boolean model: Poisson number of overlapping spheres (lognormal radii) with
 *                  intensity chosen so that exp(-lambda E[V]) = target porosity
 *   pnms_sdf_spheres  SDF = min_i(|x - c_i| - r_i): negative in the solid spheres
 *   pnms_sdf_chambers pore SDF of spherical chambers joined by capsule tubes:
 *                  max over shapes of (r - distance to the shape's core); positive in the pore
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct {
  uint64_t s[4];
} rng_t;

static uint64_t splitmix64(uint64_t* x) {
  uint64_t z = (*x += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
static void rng_seed(rng_t* r, uint64_t seed) {
  for (int i = 0; i < 4; ++i) r->s[i] = splitmix64(&seed);
}
static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static uint64_t rng_next(rng_t* r) {
  uint64_t* s = r->s;
  const uint64_t result = rotl(s[1] * 5, 7) * 9;
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl(s[3], 45);
  return result;
}
static double rng_u01(rng_t* r) { return (double)(rng_next(r) >> 11) * 0x1.0p-53; }
static double rng_normal(rng_t* r) {
  double u1 = rng_u01(r), u2 = rng_u01(r);
  if (u1 < 1e-300) u1 = 1e-300;
  return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

static double draw_radius(rng_t* r, int dist, double a, double b) {
  if (dist == 0) return a + (b - a) * rng_u01(r);           /* uniform [a, b] */
  return a * exp(b * rng_normal(r));                          /* lognormal median a, sigma b */
}

static inline double mimg(double d, double L) { return d - L * nearbyint(d / L); }

/* ---------------- dense packing ---------------- */

typedef struct {
  int nc[3];
  double cs[3];
  int* head;
  int* next;
} cells_t;

static void cells_build(cells_t* g, const double L[3], double min_cell, const double* x,
                        const double* y, const double* z, int64_t n) {
  for (int a = 0; a < 3; ++a) {
    int k = (int)floor(L[a] / min_cell);
    if (k < 1) k = 1;
    g->nc[a] = k;
    g->cs[a] = L[a] / k;
  }
  const int64_t tot = (int64_t)g->nc[0] * g->nc[1] * g->nc[2];
  g->head = (int*)malloc(sizeof(int) * (size_t)tot);
  g->next = (int*)malloc(sizeof(int) * (size_t)(n ? n : 1));
  for (int64_t c = 0; c < tot; ++c) g->head[c] = -1;
  for (int64_t i = 0; i < n; ++i) {
    const double p[3] = {x[i], y[i], z[i]};
    int ci[3];
    for (int a = 0; a < 3; ++a) {
      double q = p[a] - L[a] * floor(p[a] / L[a]);
      int k = (int)(q / g->cs[a]);
      if (k >= g->nc[a]) k = g->nc[a] - 1;
      if (k < 0) k = 0;
      ci[a] = k;
    }
    const int64_t c = ((int64_t)ci[2] * g->nc[1] + ci[1]) * g->nc[0] + ci[0];
    g->next[i] = g->head[c];
    g->head[c] = (int)i;
  }
}
static void cells_free(cells_t* g) {
  free(g->head);
  free(g->next);
}

/* Relaxation sweep at radius scale s: pairs closer than s(ri+rj) are pushed apart
 * symmetrically along the min-image separation. Returns the largest overlap seen. */
static double relax_pair(const double L[3], double s, double* x, double* y, double* z,
                         const double* r, int64_t i, int64_t j) {
  double d[3] = {mimg(x[j] - x[i], L[0]), mimg(y[j] - y[i], L[1]), mimg(z[j] - z[i], L[2])};
  const double dist = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  const double ov = s * (r[i] + r[j]) - dist;
  if (ov <= 0.0) return 0.0;
  double u[3] = {1.0, 0.0, 0.0};
  if (dist > 1e-12) {
    u[0] = d[0] / dist;
    u[1] = d[1] / dist;
    u[2] = d[2] / dist;
  }
  const double push = 0.5 * ov * 1.0001;
  x[i] -= u[0] * push;
  y[i] -= u[1] * push;
  z[i] -= u[2] * push;
  x[j] += u[0] * push;
  y[j] += u[1] * push;
  z[j] += u[2] * push;
  return ov;
}

static double relax_sweep(const double L[3], double s, double* x, double* y, double* z,
                          const double* r, int64_t n, double rmax) {
  cells_t g;
  cells_build(&g, L, 2.0 * rmax * s + 1e-9, x, y, z, n);
  double worst = 0.0;
  if (g.nc[0] < 3 || g.nc[1] < 3 || g.nc[2] < 3) {
    for (int64_t i = 0; i < n; ++i)
      for (int64_t j = i + 1; j < n; ++j) {
        const double ov = relax_pair(L, s, x, y, z, r, i, j);
        if (ov > worst) worst = ov;
      }
    cells_free(&g);
    return worst;
  }
  for (int64_t i = 0; i < n; ++i) {
    int ci[3];
    const double p[3] = {x[i], y[i], z[i]};
    for (int a = 0; a < 3; ++a) {
      double q = p[a] - L[a] * floor(p[a] / L[a]);
      int k = (int)(q / g.cs[a]);
      if (k >= g.nc[a]) k = g.nc[a] - 1;
      ci[a] = k;
    }
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int cc[3] = {(ci[0] + dx + g.nc[0]) % g.nc[0], (ci[1] + dy + g.nc[1]) % g.nc[1],
                             (ci[2] + dz + g.nc[2]) % g.nc[2]};
          const int64_t c = ((int64_t)cc[2] * g.nc[1] + cc[1]) * g.nc[0] + cc[0];
          for (int j = g.head[c]; j >= 0; j = g.next[j]) {
            if (j <= i) continue;
            const double ov = relax_pair(L, s, x, y, z, r, i, j);
            if (ov > worst) worst = ov;
          }
        }
  }
  cells_free(&g);
  return worst;
}

/* Returns the number of spheres (<= cap). *solid_out = achieved solid fraction. Radii in
 * r[] are final (already scaled when the packing jammed below full size). */
int64_t pnms_pack(const int32_t dims[3], int dist, double a, double b, double target_solid,
                  uint64_t seed, int max_sweeps, double* cx, double* cy, double* cz, double* r,
                  int64_t cap, double* solid_out) {
  rng_t g;
  rng_seed(&g, seed);
  const double L[3] = {dims[0], dims[1], dims[2]};
  const double V = L[0] * L[1] * L[2];
  double vol = 0.0, rmax = 0.0;
  int64_t n = 0;
  while (vol < target_solid * V && n < cap) {
    const double rr = draw_radius(&g, dist, a, b);
    r[n] = rr;
    if (rr > rmax) rmax = rr;
    vol += 4.0 / 3.0 * M_PI * rr * rr * rr;
    ++n;
  }
  /* place largest first is not needed for relaxation; positions uniform */
  for (int64_t i = 0; i < n; ++i) {
    cx[i] = rng_u01(&g) * L[0];
    cy[i] = rng_u01(&g) * L[1];
    cz[i] = rng_u01(&g) * L[2];
  }
  double s = 0.6;
  int sweeps = 0;
  double s_ok = 0.0;
  while (sweeps < max_sweeps) {
    const double worst = relax_sweep(L, s, cx, cy, cz, r, n, rmax);
    ++sweeps;
    if (worst < 1e-3 * s) {
      s_ok = s;
      if (s >= 1.0) break;
      s = s + 0.01 > 1.0 ? 1.0 : s + 0.01;
    }
  }
  if (s_ok < s) {
    /* jammed before reaching full size: settle at the last overlap-free scale */
    if (s_ok <= 0.0) s_ok = 0.6;
    for (int k = 0; k < 200; ++k)
      if (relax_sweep(L, s_ok, cx, cy, cz, r, n, rmax) < 1e-6) break;
    for (int64_t i = 0; i < n; ++i) r[i] *= s_ok;
  }
  for (int64_t i = 0; i < n; ++i) {
    cx[i] -= L[0] * floor(cx[i] / L[0]);
    cy[i] -= L[1] * floor(cy[i] / L[1]);
    cz[i] -= L[2] * floor(cz[i] / L[2]);
  }
  double sv = 0.0;
  for (int64_t i = 0; i < n; ++i) sv += 4.0 / 3.0 * M_PI * r[i] * r[i] * r[i];
  if (solid_out) *solid_out = sv / V;
  return n;
}

/* Boolean model: N ~ Poisson(lambda V) with lambda = -ln(porosity) / E[V_sphere];
 * E[r^3] for lognormal(median m, sigma s) = m^3 exp(4.5 s^2). */
int64_t pnms_boolean(const int32_t dims[3], double median, double sigma, double porosity,
                     uint64_t seed, double* cx, double* cy, double* cz, double* r, int64_t cap) {
  rng_t g;
  rng_seed(&g, seed);
  const double V = (double)dims[0] * dims[1] * dims[2];
  const double er3 = median * median * median * exp(4.5 * sigma * sigma);
  const double lam = -log(porosity) / (4.0 / 3.0 * M_PI * er3);
  const double mean = lam * V;
  /* Poisson via normal approximation for large means (mean >> 100), exact otherwise */
  int64_t n;
  if (mean > 500.0) {
    n = (int64_t)llround(mean + sqrt(mean) * rng_normal(&g));
  } else {
    const double Lm = exp(-mean);
    double p = 1.0;
    n = -1;
    do {
      ++n;
      p *= rng_u01(&g);
    } while (p > Lm);
  }
  if (n > cap) n = cap;
  for (int64_t i = 0; i < n; ++i) {
    cx[i] = rng_u01(&g) * dims[0];
    cy[i] = rng_u01(&g) * dims[1];
    cz[i] = rng_u01(&g) * dims[2];
    r[i] = draw_radius(&g, 1, median, sigma);
  }
  return n;
}

/* SDF = min_i(|x - c_i|_minimage - r_i), evaluated exactly: every sphere updates the voxels of
 * its bounding box widened by `margin`; voxels left above margin (or any voxel when the box
 * covers the whole axis) are resolved by brute force over all spheres. */
int pnms_sdf_spheres(const int32_t dims[3], const double* cx, const double* cy,
                     const double* cz, const double* r, int64_t ns, double margin, float* sdf) {
  const int Nx = dims[0], Ny = dims[1], Nz = dims[2];
  const int64_t n = (int64_t)Nx * Ny * Nz;
  double* acc = (double*)malloc(sizeof(double) * (size_t)n);
  if (!acc) return -1;
  for (int64_t i = 0; i < n; ++i) acc[i] = INFINITY;
  /* z-slab parallel: each thread owns whole z slices, spheres touching them update */
#pragma omp parallel for schedule(dynamic, 1)
  for (int z = 0; z < Nz; ++z) {
    for (int64_t k = 0; k < ns; ++k) {
      const double R = r[k] + margin;
      double dz = mimg((double)z - cz[k], Nz);
      if (fabs(dz) > R) continue;
      const int y0 = (int)floor(cy[k] - R), y1 = (int)ceil(cy[k] + R);
      const int x0 = (int)floor(cx[k] - R), x1 = (int)ceil(cx[k] + R);
      const int ylo = (y1 - y0 + 1 >= Ny) ? 0 : y0, yhi = (y1 - y0 + 1 >= Ny) ? Ny - 1 : y1;
      const int xlo = (x1 - x0 + 1 >= Nx) ? 0 : x0, xhi = (x1 - x0 + 1 >= Nx) ? Nx - 1 : x1;
      for (int yy = ylo; yy <= yhi; ++yy) {
        const int y = ((yy % Ny) + Ny) % Ny;
        const double dy = mimg((double)y - cy[k], Ny);
        const double dyz2 = dy * dy + dz * dz;
        if (dyz2 > R * R) continue;
        double* row = acc + ((int64_t)z * Ny + y) * Nx;
        for (int xx = xlo; xx <= xhi; ++xx) {
          const int x = ((xx % Nx) + Nx) % Nx;
          const double dx = mimg((double)x - cx[k], Nx);
          const double v = sqrt(dx * dx + dyz2) - r[k];
          if (v < row[x]) row[x] = v;
        }
      }
    }
  }
  int64_t nbrute = 0;
#pragma omp parallel for reduction(+ : nbrute) schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    if (acc[i] > margin) {
      ++nbrute;
      const int x = (int)(i % Nx), y = (int)((i / Nx) % Ny), z = (int)(i / ((int64_t)Nx * Ny));
      double best = INFINITY;
      for (int64_t k = 0; k < ns; ++k) {
        const double dx = mimg((double)x - cx[k], Nx), dy = mimg((double)y - cy[k], Ny),
                     dz = mimg((double)z - cz[k], Nz);
        const double v = sqrt(dx * dx + dy * dy + dz * dz) - r[k];
        if (v < best) best = v;
      }
      acc[i] = best;
    }
    sdf[i] = (float)acc[i];
  }
  free(acc);
  return (int)(nbrute > 0x7fffffff ? 0x7fffffff : nbrute);
}

/* Pore SDF of spherical chambers at lattice points joined along +x/+y/+z by capsule tubes:
 * sdf = max(max_c (Rc - |x - c|), max_t (rt - dist(x, segment_t))); positive in the pore. */
int pnms_sdf_chambers(const int32_t dims[3], int m, double chamber_r, double tube_rmin,
                      double tube_rmax, uint64_t seed, float* sdf, double* tube_r_out) {
  rng_t g;
  rng_seed(&g, seed);
  const int Nx = dims[0], Ny = dims[1], Nz = dims[2];
  const double L[3] = {Nx, Ny, Nz};
  const double pitch[3] = {L[0] / m, L[1] / m, L[2] / m};
  const int nch = m * m * m;
  double* tr = (double*)malloc(sizeof(double) * (size_t)nch * 3);
  for (int k = 0; k < nch * 3; ++k) tr[k] = tube_rmin + (tube_rmax - tube_rmin) * rng_u01(&g);
  if (tube_r_out) memcpy(tube_r_out, tr, sizeof(double) * (size_t)nch * 3);
  const int64_t n = (int64_t)Nx * Ny * Nz;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const int x = (int)(i % Nx), y = (int)((i / Nx) % Ny), z = (int)(i / ((int64_t)Nx * Ny));
    const double p[3] = {x, y, z};
    /* nearest lattice cell */
    int base[3];
    for (int a = 0; a < 3; ++a) base[a] = (int)floor(p[a] / pitch[a] - 0.5);
    double best = -INFINITY;
    for (int oz = 0; oz <= 1; ++oz)
      for (int oy = 0; oy <= 1; ++oy)
        for (int ox = 0; ox <= 1; ++ox) {
          const int cidx[3] = {base[0] + ox, base[1] + oy, base[2] + oz};
          int cw[3];
          double c[3];
          for (int a = 0; a < 3; ++a) {
            cw[a] = ((cidx[a] % m) + m) % m;
            c[a] = (cidx[a] + 0.5) * pitch[a];
          }
          double d[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
          const double dc = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
          if (chamber_r - dc > best) best = chamber_r - dc;
          for (int ax = 0; ax < 3; ++ax) {
            /* tube from c along +ax to the next centre */
            const double t = d[ax];
            if (t < 0.0 || t > pitch[ax]) continue;
            double q2 = 0.0;
            for (int a = 0; a < 3; ++a)
              if (a != ax) q2 += d[a] * d[a];
            const int ci = (cw[2] * m + cw[1]) * m + cw[0];
            const double v = tr[ci * 3 + ax] - sqrt(q2);
            if (v > best) best = v;
          }
        }
    sdf[i] = (float)best;
  }
  free(tr);
  return 0;
}

/* Random sequential addition, largest radii first: each sphere tries up to `attempts` uniform
 * positions and is kept at the first one that overlaps no kept sphere (periodic min-image);
 * spheres that find no place are dropped. Stops when the kept volume reaches target_solid.
 * Returns the number kept. */
static int cmp_desc(const void* a, const void* b) {
  const double x = *(const double*)a, y = *(const double*)b;
  return (x < y) - (x > y);
}

int64_t pnms_rsa(const int32_t dims[3], int dist, double a, double b, double target_solid,
                 uint64_t seed, int attempts, double* cx, double* cy, double* cz, double* r,
                 int64_t cap, double* solid_out) {
  rng_t g;
  rng_seed(&g, seed);
  const double L[3] = {dims[0], dims[1], dims[2]};
  const double V = L[0] * L[1] * L[2];
  /* draw a candidate list with 1.5x the target volume, sort descending */
  int64_t m = 0;
  double vol = 0.0;
  double* rad = (double*)malloc(sizeof(double) * (size_t)cap);
  while (vol < 1.5 * target_solid * V && m < cap) {
    const double rr = draw_radius(&g, dist, a, b);
    rad[m++] = rr;
    vol += 4.0 / 3.0 * M_PI * rr * rr * rr;
  }
  qsort(rad, (size_t)m, sizeof(double), cmp_desc);
  const double rmax = m ? rad[0] : 1.0;
  int nc[3];
  double cs[3];
  for (int k = 0; k < 3; ++k) {
    nc[k] = (int)floor(L[k] / (2.0 * rmax));
    if (nc[k] < 1) nc[k] = 1;
    cs[k] = L[k] / nc[k];
  }
  const int64_t ncell = (int64_t)nc[0] * nc[1] * nc[2];
  int* head = (int*)malloc(sizeof(int) * (size_t)ncell);
  int* next = (int*)malloc(sizeof(int) * (size_t)cap);
  for (int64_t c = 0; c < ncell; ++c) head[c] = -1;
  const int small = nc[0] < 3 || nc[1] < 3 || nc[2] < 3;
  int64_t kept = 0;
  double kv = 0.0;
  for (int64_t s = 0; s < m && kv < target_solid * V; ++s) {
    const double rr = rad[s];
    for (int t = 0; t < attempts; ++t) {
      const double p[3] = {rng_u01(&g) * L[0], rng_u01(&g) * L[1], rng_u01(&g) * L[2]};
      int ci[3];
      for (int k = 0; k < 3; ++k) {
        ci[k] = (int)(p[k] / cs[k]);
        if (ci[k] >= nc[k]) ci[k] = nc[k] - 1;
      }
      int ok = 1;
      if (small) {
        for (int64_t j = 0; j < kept && ok; ++j) {
          const double dx = mimg(p[0] - cx[j], L[0]), dy = mimg(p[1] - cy[j], L[1]),
                       dz = mimg(p[2] - cz[j], L[2]);
          if (dx * dx + dy * dy + dz * dz < (rr + r[j]) * (rr + r[j])) ok = 0;
        }
      } else {
        for (int dz = -1; dz <= 1 && ok; ++dz)
          for (int dy = -1; dy <= 1 && ok; ++dy)
            for (int dx = -1; dx <= 1 && ok; ++dx) {
              const int cc[3] = {(ci[0] + dx + nc[0]) % nc[0], (ci[1] + dy + nc[1]) % nc[1],
                                 (ci[2] + dz + nc[2]) % nc[2]};
              const int64_t c = ((int64_t)cc[2] * nc[1] + cc[1]) * nc[0] + cc[0];
              for (int j = head[c]; j >= 0 && ok; j = next[j]) {
                const double ex = mimg(p[0] - cx[j], L[0]), ey = mimg(p[1] - cy[j], L[1]),
                             ez = mimg(p[2] - cz[j], L[2]);
                if (ex * ex + ey * ey + ez * ez < (rr + r[j]) * (rr + r[j])) ok = 0;
              }
            }
      }
      if (!ok) continue;
      cx[kept] = p[0];
      cy[kept] = p[1];
      cz[kept] = p[2];
      r[kept] = rr;
      const int64_t c = ((int64_t)ci[2] * nc[1] + ci[1]) * nc[0] + ci[0];
      next[kept] = head[c];
      head[c] = (int)kept;
      kv += 4.0 / 3.0 * M_PI * rr * rr * rr;
      ++kept;
      break;
    }
  }
  free(rad);
  free(head);
  free(next);
  if (solid_out) *solid_out = kv / V;
  return kept;
}
