/*
 * pnm_oracle.c — serial CPU oracle (TEST INFRASTRUCTURE ONLY; the product never links it).
 *
 * A plain-C restatement of the single-rank algorithm in src/pore_extraction.hpp of
 * computational-chemical-engineering/peclet-pnm. Each function cites the reference lines it
 * follows. Parallel reference kernels become serial loops; every fixpoint the reference reaches
 * with device atomics has an order-free characterisation (component minimum, BFS-layer minimum,
 * min reachable core root), so the serial loop reaches the same fixpoint.
 *
 * Build: gcc -O2 -std=c11 -ffp-contract=off -fno-fast-math -fPIC -shared (see oracle/Makefile).
 */
#include "pnm_oracle.h"

#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---- indexing: pore_extraction.hpp:59-64 (get_idx, periodic wrap on every access) ---- */
static inline int wrapm(int x, int N) { return ((x % N) + N) % N; }
/* Same result as wrapm for x in [-1, N] (every stencil offset here is -1, 0 or +1). */
static inline int wrap1(int x, int N) { return x < 0 ? x + N : (x >= N ? x - N : x); }
static inline int gidx1(int x, int y, int z, int Nx, int Ny, int Nz) {
  return wrap1(z, Nz) * Ny * Nx + wrap1(y, Ny) * Nx + wrap1(x, Nx);
}
static inline int gidxm(int x, int y, int z, int Nx, int Ny, int Nz) {
  return wrapm(z, Nz) * Ny * Nx + wrapm(y, Ny) * Nx + wrapm(x, Nx);
}

void pnmo_free(void* p) { free(p); }

/* Squared-positive-SDF weighted 3x3x3 centroid, summed dz -> dy -> dx in FP32:
 * pore_extraction.hpp:107-127 (and the identical nf::centers block :543-562). */
static pnmo_pore centroid_pore(const float* sdf, int ix, int iy, int iz, int Nx, int Ny, int Nz,
                               const float o[3], const float s[3], float radius) {
  float sw = 0.0f, px = 0.0f, py = 0.0f, pz = 0.0f;
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        const float v = sdf[gidx1(ix + dx, iy + dy, iz + dz, Nx, Ny, Nz)];
        float w = v > 0.0f ? v : 0.0f;
        w = w * w;
        sw += w;
        px += (float)dx * w;
        py += (float)dy * w;
        pz += (float)dz * w;
      }
  float fx = 0.0f, fy = 0.0f, fz = 0.0f;
  if (sw > 1e-6f) {
    fx = px / sw;
    fy = py / sw;
    fz = pz / sw;
  }
  pnmo_pore P;
  P.x = o[0] + ((float)ix + fx) * s[0];
  P.y = o[1] + ((float)iy + fy) * s[1];
  P.z = o[2] + ((float)iz + fz) * s[2];
  P.radius = radius;
  return P;
}

/* ---- pores: extractPoresView, pore_extraction.hpp:69-140 ----
 * Peak iff cv is not <= 0 and no 26-neighbour has (nv > cv) or (nv == cv and ni > ci).
 * The reference emits in atomic-slot order (a race); the oracle emits in ascending index. */
int64_t pnmo_extract_pores(const float* sdf, const int32_t dims[3], const float origin[3],
                           const float spacing[3], pnmo_pore* out, int64_t cap) {
  const int Nx = dims[0], Ny = dims[1], Nz = dims[2];
  int64_t count = 0;
  for (int iz = 0; iz < Nz; ++iz)
    for (int iy = 0; iy < Ny; ++iy)
      for (int ix = 0; ix < Nx; ++ix) {
        const int ci = (iz * Ny + iy) * Nx + ix;
        const float cv = sdf[ci];
        if (cv <= 0.0f) continue;
        int peak = 1;
        for (int dz = -1; dz <= 1 && peak; ++dz)
          for (int dy = -1; dy <= 1 && peak; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
              if (dx == 0 && dy == 0 && dz == 0) continue;
              const int ni = gidx1(ix + dx, iy + dy, iz + dz, Nx, Ny, Nz);
              const float nv = sdf[ni];
              if (nv > cv || (nv == cv && ni > ci)) {
                peak = 0;
                break;
              }
            }
        if (!peak) continue;
        if (count < cap)
          out[count] = centroid_pore(sdf, ix, iy, iz, Nx, Ny, Nz, origin, spacing, cv);
        ++count;
      }
  return count;
}

/* ---- step 1, init_markers: pore_extraction.hpp:163-179 ----
 * thr = -1.5f * min(spacing); labels[i] = i if sdf < thr else -1. */
void pnmo_init_markers(const float* sdf, const int32_t dims[3], const float spacing[3],
                       int32_t* labels) {
  const int64_t n = (int64_t)dims[0] * dims[1] * dims[2];
  float mn = spacing[1] < spacing[2] ? spacing[1] : spacing[2];
  mn = spacing[0] < mn ? spacing[0] : mn;
  const float thr = -1.5f * mn;
  for (int64_t i = 0; i < n; ++i) labels[i] = (sdf[i] < thr) ? (int32_t)i : -1;
}

/* ---- step 2, union-find CCL over markers: pore_extraction.hpp:183-230 ----
 * merge over the 13 forward periodic neighbours (:204-206), atomic_min hooking of the larger
 * root onto the smaller, full flatten after every sweep, until a sweep changes nothing.
 * Fixpoint: each marker's label = min flat index of its periodic 26-connected component. */
int pnmo_ccl(const int32_t dims[3], int32_t* L) {
  static const int dz_l[13] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0};
  static const int dy_l[13] = {-1, -1, -1, 0, 0, 0, 1, 1, 1, 1, 1, 1, 0};
  static const int dx_l[13] = {-1, 0, 1, -1, 0, 1, -1, 0, 1, -1, 0, 1, 1};
  const int Nx = dims[0], Ny = dims[1], Nz = dims[2];
  const int64_t n = (int64_t)Nx * Ny * Nz;
  int sweeps = 0, changed = 1;
  while (changed) {
    changed = 0;
    ++sweeps;
    for (int iz = 0; iz < Nz; ++iz)
      for (int iy = 0; iy < Ny; ++iy)
        for (int ix = 0; ix < Nx; ++ix) {
          const int idx = (iz * Ny + iy) * Nx + ix;
          const int my = L[idx];
          if (my == -1) continue;
          for (int k = 0; k < 13; ++k) {
            const int ni = gidx1(ix + dx_l[k], iy + dy_l[k], iz + dz_l[k], Nx, Ny, Nz);
            const int nl = L[ni];
            if (nl != -1 && my != nl) {
              int rm = my;
              while (rm != L[rm]) rm = L[rm];
              int rn = nl;
              while (rn != L[rn]) rn = L[rn];
              if (rm != rn) {
                const int sml = rm < rn ? rm : rn, lrg = rm < rn ? rn : rm;
                if (sml < L[lrg]) L[lrg] = sml; /* atomic_min */
                changed = 1;
              }
            }
          }
        }
    /* flatten :183-194 */
    for (int64_t i = 0; i < n; ++i) {
      int l = L[i];
      if (l != -1) {
        while (l != L[l]) l = L[l];
        L[i] = l;
      }
    }
  }
  return sweeps;
}

/* ---- step 3, Jacobi flood of the shallow solid: pore_extraction.hpp:237-269 ----
 * Voxels with !(sdf >= 0) and label -1 take the minimum label (!= -1) among their 26
 * neighbours in the PREVIOUS sweep; repeat until a sweep changes nothing. */
int pnmo_flood(const float* sdf, const int32_t dims[3], int32_t* labels) {
  const int Nx = dims[0], Ny = dims[1], Nz = dims[2];
  const int64_t n = (int64_t)Nx * Ny * Nz;
  int32_t* prev = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
  int sweeps = 0, changed = 1;
  while (changed) {
    changed = 0;
    ++sweeps;
    memcpy(prev, labels, sizeof(int32_t) * (size_t)n);
    for (int iz = 0; iz < Nz; ++iz)
      for (int iy = 0; iy < Ny; ++iy)
        for (int ix = 0; ix < Nx; ++ix) {
          const int idx = (iz * Ny + iy) * Nx + ix;
          if (sdf[idx] >= 0.0f) continue;
          if (prev[idx] != -1) continue;
          int best = -1;
          for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
              for (int dx = -1; dx <= 1; ++dx) {
                if (dx == 0 && dy == 0 && dz == 0) continue;
                const int nl = prev[gidx1(ix + dx, iy + dy, iz + dz, Nx, Ny, Nz)];
                if (nl != -1 && (best == -1 || nl < best)) best = nl;
              }
          if (best != -1) {
            labels[idx] = best;
            changed = 1;
          }
        }
  }
  free(prev);
  return sweeps;
}

/* One ascent step: lexicographic maximum over (sdf, flat index) of the walker and its 26
 * neighbours (pore_extraction.hpp:280-304 ascent branch; :629-644 for the flow basins). */
/* One step of the gradient walk (pore_extraction.hpp:280-304): over the 26 periodic
 * neighbours in dz -> dy -> dx order, ascent takes nv > bv, descent nv < bv, and either takes an
 * equal value with a larger index than the current best (NaN never compares). */
static inline int gradient_step(const float* sdf, int walker, int Nx, int Ny, int Nz, int up) {
  int best = walker;
  float bv = sdf[walker];
  const int wx = walker % Nx, wy = (walker / Nx) % Ny, wz = walker / (Nx * Ny);
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        if (dx == 0 && dy == 0 && dz == 0) continue;
        const int ni = gidx1(wx + dx, wy + dy, wz + dz, Nx, Ny, Nz);
        const float nv = sdf[ni];
        if (up ? (nv > bv) : (nv < bv)) {
          bv = nv;
          best = ni;
        } else if (nv == bv && ni > best)
          best = ni;
      }
  return best;
}

/* ---- step 4, gradient_path: pore_extraction.hpp:271-311 (MAX_STEPS = 512, :278) ----
 * The root is the walker after min(depth, 512) steps: ascent from sdf > 0 voxels, descent from
 * the others (the direction is fixed by the starting voxel, :276). Only sdf > 0 roots feed the
 * segmentation (:333-334, :342, :380, :530); the descent roots are the rest of rootsOut
 * (:383-384). ascent_all = 1 restates nf::flow_basins (:621-652) instead: ascent from every
 * voxel. */
void pnmo_gradient_roots(const float* sdf, const int32_t dims[3], int ascent_all,
                         int32_t* roots) {
  const int Nx = dims[0], Ny = dims[1], Nz = dims[2];
  const int64_t n = (int64_t)Nx * Ny * Nz;
  for (int64_t i = 0; i < n; ++i) {
    const int up = ascent_all || sdf[i] > 0.0f;
    int walker = (int)i;
    for (int s = 0; s < 512; ++s) {
      const int best = gradient_step(sdf, walker, Nx, Ny, Nz, up);
      if (best == walker) break;
      walker = best;
    }
    roots[i] = walker;
  }
}

/* ---- step 5, renumbering by first appearance: pore_extraction.hpp:313-382 ----
 * Pores: id = 1 + rank of the basin's minimum voxel index; solids: id = -(1 + rank) of the
 * label's minimum voxel index among !(sdf>0) labelled voxels, flagged only where that voxel has
 * sdf <= 0 (:343-345); unlabelled solid -> 0. */
void pnmo_relabel(const float* sdf, const int32_t* roots, const int32_t* labels, int64_t n,
                  int32_t* seg) {
  const int kBig = 0x7fffffff;
  int32_t* minPore = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int32_t* minSolid = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int32_t* poreId = (int32_t*)calloc((size_t)n, sizeof(int32_t));
  int32_t* solidId = (int32_t*)calloc((size_t)n, sizeof(int32_t));
  for (int64_t i = 0; i < n; ++i) minPore[i] = minSolid[i] = kBig;
  /* (a) relabel_min :331-337 */
  for (int64_t i = 0; i < n; ++i) {
    if (sdf[i] > 0.0f) {
      if ((int32_t)i < minPore[roots[i]]) minPore[roots[i]] = (int32_t)i;
    } else if (labels[i] != -1) {
      if ((int32_t)i < minSolid[labels[i]]) minSolid[labels[i]] = (int32_t)i;
    }
  }
  /* (b)-(d) relabel_first + exclusive scans + assign :340-374 (serial scan = running count) */
  int32_t pr = 0, sr = 0;
  for (int64_t i = 0; i < n; ++i) {
    const int pf = (sdf[i] > 0.0f && minPore[roots[i]] == (int32_t)i);
    const int sf = (sdf[i] <= 0.0f && labels[i] != -1 && minSolid[labels[i]] == (int32_t)i);
    if (pf) poreId[roots[i]] = pr + 1;
    if (sf) solidId[labels[i]] = -(sr + 1);
    pr += pf;
    sr += sf;
  }
  /* (e) relabel_seg :378-381 */
  for (int64_t i = 0; i < n; ++i)
    seg[i] = (sdf[i] > 0.0f) ? poreId[roots[i]] : (labels[i] == -1 ? 0 : solidId[labels[i]]);
  free(minPore);
  free(minSolid);
  free(poreId);
  free(solidId);
}

/* ---- segmentVolumeView: pore_extraction.hpp:157-386 ---- */
int pnmo_segment_volume(const float* sdf, const int32_t dims[3], const float spacing[3],
                        int32_t* seg, int32_t* roots_out, int32_t* labels_out) {
  const int64_t n = (int64_t)dims[0] * dims[1] * dims[2];
  if (n == 0) return 0;
  int32_t* labels = labels_out ? labels_out : (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int32_t* roots = roots_out ? roots_out : (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  pnmo_init_markers(sdf, dims, spacing, labels);
  pnmo_ccl(dims, labels);
  pnmo_flood(sdf, dims, labels);
  pnmo_gradient_roots(sdf, dims, 0, roots);
  pnmo_relabel(sdf, roots, labels, n, seg);
  if (!labels_out) free(labels);
  if (!roots_out) free(roots);
  return 0;
}

/* Fixed summation order of the per-pore and per-throat sums (shared with the device's
 * reduction, k_key_sums in k_flow.hip): the terms t[0..m) of one key are taken in blocks of 2048;
 * in a block starting at b, "lane" l (0..63) adds t[b + l], t[b + 64 + l], ..., t[b + 64*31 + l]
 * (those below m) one after another starting from +0.0; the 64 lane sums are combined by a
 * pairwise tree, s = 1, 2, 4, ..., 32: p[l] = p[l] + p[l + s] for l a multiple of 2s; the block
 * totals p[0] are added one after another starting from +0.0. */
static double blocked_sum(const double* t, int64_t stride, int64_t m) {
  double acc = 0.0;
  for (int64_t b = 0; b < m; b += 2048) {
    double p[64];
    for (int l = 0; l < 64; ++l) {
      p[l] = 0.0;
      for (int j = 0; j < 32; ++j) {
        const int64_t i = b + 64 * (int64_t)j + l;
        if (i < m) p[l] += t[i * stride];
      }
    }
    for (int sh = 1; sh < 64; sh <<= 1)
      for (int l = 0; l < 64; l += 2 * sh) p[l] = p[l] + p[l + sh];
    acc += p[0];
  }
  return acc;
}

/* out[k][key] = blocked sum of component k of the records with that key, taken in record
 * order (a stable counting sort by key, then blocked_sum per key). vals: K doubles per record. */
static void keyed_blocked_sums(const int64_t* key, const double* vals, int K, int64_t cnt,
                               int64_t nkeys, double** out) {
  int64_t* start = (int64_t*)calloc((size_t)nkeys + 1, sizeof(int64_t));
  for (int64_t i = 0; i < cnt; ++i) ++start[key[i] + 1];
  for (int64_t k = 0; k < nkeys; ++k) start[k + 1] += start[k];
  int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nkeys + 1));
  memcpy(fill, start, sizeof(int64_t) * (size_t)(nkeys + 1));
  double* sorted = (double*)malloc(sizeof(double) * (size_t)(K * (cnt ? cnt : 1)));
  for (int64_t i = 0; i < cnt; ++i) {
    const int64_t pos = fill[key[i]]++;
    for (int k = 0; k < K; ++k) sorted[(int64_t)K * pos + k] = vals[(int64_t)K * i + k];
  }
  for (int64_t k = 0; k < nkeys; ++k)
    for (int c = 0; c < K; ++c)
      out[c][k] = blocked_sum(sorted + (int64_t)K * start[k] + c, K, start[k + 1] - start[k]);
  free(sorted);
  free(fill);
  free(start);
}

/* ---- topology: extractTopologyView, pore_extraction.hpp:401-451 ----
 * Every +x/+y/+z face (wrapped) with differing labels gives (min,max); host sort + unique in
 * signed lexicographic pair order (:448-449). */
static int cmp_pair(const void* a, const void* b) {
  const int32_t* p = (const int32_t*)a;
  const int32_t* q = (const int32_t*)b;
  if (p[0] != q[0]) return p[0] < q[0] ? -1 : 1;
  if (p[1] != q[1]) return p[1] < q[1] ? -1 : 1;
  return 0;
}

int64_t pnmo_extract_topology(const int32_t* seg, const int32_t dims[3], int32_t** pairs) {
  const int Nx = dims[0], Ny = dims[1], Nz = dims[2];
  const int64_t n = (int64_t)Nx * Ny * Nz;
  *pairs = NULL;
  if (n == 0) return 0;
  int64_t cap = 3 * n, cnt = 0;
  int32_t* buf = (int32_t*)malloc(sizeof(int32_t) * (size_t)(2 * cap));
  for (int iz = 0; iz < Nz; ++iz)
    for (int iy = 0; iy < Ny; ++iy)
      for (int ix = 0; ix < Nx; ++ix) {
        const int idx = (iz * Ny + iy) * Nx + ix;
        const int my = seg[idx];
        for (int k = 0; k < 3; ++k) {
          const int nl = seg[gidx1(ix + (k == 0), iy + (k == 1), iz + (k == 2), Nx, Ny, Nz)];
          if (my != nl) {
            buf[2 * cnt] = my < nl ? my : nl;
            buf[2 * cnt + 1] = my < nl ? nl : my;
            ++cnt;
          }
        }
      }
  qsort(buf, (size_t)cnt, 2 * sizeof(int32_t), cmp_pair);
  int64_t u = 0;
  for (int64_t i = 0; i < cnt; ++i) {
    if (u > 0 && buf[2 * (u - 1)] == buf[2 * i] && buf[2 * (u - 1) + 1] == buf[2 * i + 1])
      continue;
    buf[2 * u] = buf[2 * i];
    buf[2 * u + 1] = buf[2 * i + 1];
    ++u;
  }
  *pairs = buf;
  return u;
}

/* ====================== network flow: pore_extraction.hpp:493-997 ====================== */

#define KSENT ((int64_t)0x7ffffffffffffffeLL) /* :671 */

typedef struct {
  int64_t key, fid;
} root_rec;

static int cmp_root(const void* a, const void* b) {
  const root_rec* p = (const root_rec*)a;
  const root_rec* q = (const root_rec*)b;
  if (p->key != q->key) return p->key < q->key ? -1 : 1;
  if (p->fid != q->fid) return p->fid < q->fid ? -1 : 1;
  return 0;
}

/* cclPass :699-755 — union-find over faces with par >= 0, same pair key and doubled-offset
 * adjacency |2dc + e_d2 - e_d|^2 <= 8; atomic_min hooking, flatten, until no change. */
static void face_ccl(int Nx, int Ny, int Nz, const int64_t* fpair, int64_t* par) {
  const int64_t nf = 3 * (int64_t)Nx * Ny * Nz;
  int changed = 1;
  while (changed) {
    changed = 0;
    for (int izz = 0; izz < Nz; ++izz)
      for (int iyy = 0; iyy < Ny; ++iyy)
        for (int ixx = 0; ixx < Nx; ++ixx) {
          const int c = (izz * Ny + iyy) * Nx + ixx;
          for (int d = 0; d < 3; ++d) {
            const int64_t f = 3 * (int64_t)c + d;
            if (par[f] < 0) continue;
            const int64_t pk = fpair[f];
            for (int dz = -1; dz <= 1; ++dz)
              for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                  const int c2 = gidx1(ixx + dx, iyy + dy, izz + dz, Nx, Ny, Nz);
                  for (int d2 = 0; d2 < 3; ++d2) {
                    const int64_t f2 = 3 * (int64_t)c2 + d2;
                    if (f2 == f || par[f2] < 0 || fpair[f2] != pk) continue;
                    const int Dx = 2 * dx + (d2 == 0) - (d == 0);
                    const int Dy = 2 * dy + (d2 == 1) - (d == 1);
                    const int Dz = 2 * dz + (d2 == 2) - (d == 2);
                    if (Dx * Dx + Dy * Dy + Dz * Dz > 8) continue;
                    int64_t rm = f;
                    while (rm != par[rm]) rm = par[rm];
                    int64_t rn = f2;
                    while (rn != par[rn]) rn = par[rn];
                    if (rm != rn) {
                      const int64_t sml = rm < rn ? rm : rn, lrg = rm < rn ? rn : rm;
                      if (sml < par[lrg]) par[lrg] = sml;
                      changed = 1;
                    }
                  }
                }
          }
        }
    for (int64_t f = 0; f < nf; ++f) {
      int64_t l = par[f];
      if (l >= 0) {
        while (l != par[l]) l = par[l];
        par[f] = l;
      }
    }
  }
}

int pnmo_extract_network_flow(const float* sdf, const int32_t dims[3], const float origin[3],
                              const float spacing[3], const double* u, const double* v,
                              const double* w, const double* p, const double* ox,
                              const double* oy, const double* oz, const double grad_p[3],
                              int want_debug, pnmo_flow_result* out) {
  memset(out, 0, sizeof(*out));
  const int Nx = dims[0], Ny = dims[1], Nz = dims[2];
  const int64_t n = (int64_t)Nx * Ny * Nz;
  if (n == 0) return 0;
  const int ho = (ox != NULL);

  /* segmentation with roots :508-511 */
  int32_t* seg = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int32_t* roots = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  pnmo_segment_volume(sdf, dims, spacing, seg, roots, NULL);

  /* np = max(seg) :514-524; the reference returns empty for np == 0 and has undefined
   * behaviour for np < 0 (all-solid); both return empty here. */
  int np = INT_MIN;
  for (int64_t i = 0; i < n; ++i)
    if (seg[i] > np) np = seg[i];
  if (np <= 0) {
    free(seg);
    free(roots);
    return 0;
  }

  /* peaks :527-533 (zero-initialised View: a label without a self-root points at voxel 0) */
  int32_t* peak = (int32_t*)calloc((size_t)np, sizeof(int32_t));
  for (int64_t i = 0; i < n; ++i)
    if (sdf[i] > 0.0f && roots[i] == (int32_t)i) peak[seg[i] - 1] = (int32_t)i;

  /* centers :534-566 */
  pnmo_pore* pores = (pnmo_pore*)malloc(sizeof(pnmo_pore) * (size_t)np);
  for (int k = 0; k < np; ++k) {
    const int ci = peak[k];
    const int ix = ci % Nx, iy = (ci / Nx) % Ny, iz = ci / (Nx * Ny);
    pores[k] = centroid_pore(sdf, ix, iy, iz, Nx, Ny, Nz, origin, spacing, sdf[ci]);
  }

  /* pore_pressure :582-613 — grid coordinate in FP32, widened; trilinear in FP64 */
  double* ppres = (double*)malloc(sizeof(double) * (size_t)np);
  for (int k = 0; k < np; ++k) {
    const pnmo_pore po = pores[k];
    const float gx = (po.x - origin[0]) / spacing[0];
    const float gy = (po.y - origin[1]) / spacing[1];
    const float gz = (po.z - origin[2]) / spacing[2];
    const double g[3] = {(double)gx, (double)gy, (double)gz};
    int b[3];
    double f[3];
    for (int a = 0; a < 3; ++a) {
      const double fl = floor(g[a]);
      b[a] = (int)fl;
      f[a] = g[a] - fl;
    }
    double acc = 0.0;
    for (int dz = 0; dz < 2; ++dz)
      for (int dy = 0; dy < 2; ++dy)
        for (int dx = 0; dx < 2; ++dx) {
          const double wt =
              (dx ? f[0] : 1.0 - f[0]) * (dy ? f[1] : 1.0 - f[1]) * (dz ? f[2] : 1.0 - f[2]);
          acc += wt * p[gidxm(b[0] + dx, b[1] + dy, b[2] + dz, Nx, Ny, Nz)];
        }
    ppres[k] = acc;
  }

  /* flow basins :621-652 — ascent from every cell; flowLab = seg at the walker */
  int32_t* flowLab = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  {
    int32_t* fr = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    pnmo_gradient_roots(sdf, dims, 1, fr);
    for (int64_t i = 0; i < n; ++i) flowLab[i] = seg[fr[i]];
    free(fr);
  }

  /* face init :670-696 */
  const int64_t nf = 3 * n;
  int64_t* fpar = (int64_t*)malloc(sizeof(int64_t) * (size_t)nf);
  int64_t* fpair = (int64_t*)malloc(sizeof(int64_t) * (size_t)nf);
  int64_t* par = (int64_t*)malloc(sizeof(int64_t) * (size_t)nf);
  char* core = (char*)malloc((size_t)nf);
  for (int izz = 0; izz < Nz; ++izz)
    for (int iyy = 0; iyy < Ny; ++iyy)
      for (int ixx = 0; ixx < Nx; ++ixx) {
        const int c = (izz * Ny + iyy) * Nx + ixx;
        for (int d = 0; d < 3; ++d) {
          const int64_t f = 3 * (int64_t)c + d;
          const int nb = gidx1(ixx + (d == 0), iyy + (d == 1), izz + (d == 2), Nx, Ny, Nz);
          const int a = flowLab[c], b = flowLab[nb];
          const double opn = ho ? (d == 0 ? ox[nb] : (d == 1 ? oy[nb] : oz[nb])) : 1.0;
          const int itf = (a != b && a > 0 && b > 0 && opn > 0.0);
          core[f] = (itf && sdf[c] > 0.0f && sdf[nb] > 0.0f) ? 1 : 0;
          par[f] = core[f] ? f : -1;
          fpar[f] = itf ? (core[f] ? f : KSENT) : -1;
          fpair[f] = itf ? (((int64_t)(a < b ? a : b) << 32) | (int64_t)(a < b ? b : a)) : -1;
        }
      }
  /* core tier :756-762 */
  face_ccl(Nx, Ny, Nz, fpair, par);
  for (int64_t f = 0; f < nf; ++f)
    if (core[f]) fpar[f] = par[f];
  /* film attachment :763-805 (in place; its fixpoint is the min reachable core root) */
  {
    int changed = 1;
    while (changed) {
      changed = 0;
      for (int izz = 0; izz < Nz; ++izz)
        for (int iyy = 0; iyy < Ny; ++iyy)
          for (int ixx = 0; ixx < Nx; ++ixx) {
            const int c = (izz * Ny + iyy) * Nx + ixx;
            for (int d = 0; d < 3; ++d) {
              const int64_t f = 3 * (int64_t)c + d;
              if (fpar[f] < 0 || core[f]) continue;
              int64_t best = fpar[f];
              const int64_t pk = fpair[f];
              for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                  for (int dx = -1; dx <= 1; ++dx) {
                    const int c2 = gidx1(ixx + dx, iyy + dy, izz + dz, Nx, Ny, Nz);
                    for (int d2 = 0; d2 < 3; ++d2) {
                      const int64_t f2 = 3 * (int64_t)c2 + d2;
                      if (f2 == f || fpair[f2] != pk) continue;
                      const int64_t l2 = fpar[f2];
                      if (l2 < 0 || l2 >= KSENT || l2 >= best) continue;
                      const int Dx = 2 * dx + (d2 == 0) - (d == 0);
                      const int Dy = 2 * dy + (d2 == 1) - (d == 1);
                      const int Dz = 2 * dz + (d2 == 2) - (d == 2);
                      if (Dx * Dx + Dy * Dy + Dz * Dz > 8) continue;
                      best = l2;
                    }
                  }
              if (best < fpar[f]) {
                fpar[f] = best;
                changed = 1;
              }
            }
          }
    }
  }
  /* leftover tier :806-818 */
  for (int64_t f = 0; f < nf; ++f) par[f] = (fpar[f] == KSENT) ? f : -1;
  face_ccl(Nx, Ny, Nz, fpair, par);
  for (int64_t f = 0; f < nf; ++f)
    if (fpar[f] == KSENT) fpar[f] = par[f];

  /* patch roots -> throat slots sorted by (pair, root fid) :820-862 */
  int64_t nt = 0;
  for (int64_t f = 0; f < nf; ++f)
    if (fpar[f] == f) ++nt;
  root_rec* rr = (root_rec*)malloc(sizeof(root_rec) * (size_t)(nt ? nt : 1));
  {
    int64_t k = 0;
    for (int64_t f = 0; f < nf; ++f)
      if (fpar[f] == f) {
        rr[k].key = fpair[f];
        rr[k].fid = f;
        ++k;
      }
  }
  qsort(rr, (size_t)nt, sizeof(root_rec), cmp_root);
  int32_t* slotOfFace = (int32_t*)malloc(sizeof(int32_t) * (size_t)nf);
  for (int64_t f = 0; f < nf; ++f) slotOfFace[f] = -1;
  int32_t* throats = (int32_t*)malloc(sizeof(int32_t) * (size_t)(2 * (nt ? nt : 1)));
  for (int64_t t = 0; t < nt; ++t) {
    throats[2 * t] = (int32_t)(rr[t].key >> 32);
    throats[2 * t + 1] = (int32_t)(rr[t].key & 0x7fffffff);
    slotOfFace[rr[t].fid] = (int32_t)t;
  }

  /* throat_flux :864-944. The reference adds every face's contributions with unordered FP64
   * atomics; here each key's (pore's or throat's) terms are gathered in face order (z, y, x, d)
   * and summed in the fixed blocked order of blocked_sum, which the device reduction shares. */
  double* Q = (double*)calloc((size_t)(nt ? nt : 1), sizeof(double));
  double* A = (double*)calloc((size_t)(nt ? nt : 1), sizeof(double));
  double* Cx = (double*)calloc((size_t)(nt ? nt : 1), sizeof(double));
  double* Cy = (double*)calloc((size_t)(nt ? nt : 1), sizeof(double));
  double* Cz = (double*)calloc((size_t)(nt ? nt : 1), sizeof(double));
  double* resid = (double*)calloc((size_t)np, sizeof(double));
  {
    const double Ax = (double)spacing[1] * spacing[2], Ay = (double)spacing[0] * spacing[2],
                 Az = (double)spacing[0] * spacing[1];
    const float oxo = origin[0], oyo = origin[1], ozo = origin[2];
    const float sx = spacing[0], sy = spacing[1], sz = spacing[2];
    const double Lw[3] = {(double)Nx * sx, (double)Ny * sy, (double)Nz * sz};
    /* terms in face order: residual (pore, signed q) and throat (slot, q, area, C) records */
    int64_t nr = 0, ntr = 0;
    int64_t* rkey = (int64_t*)malloc(sizeof(int64_t) * (size_t)(2 * nf + 1));
    double* rq = (double*)malloc(sizeof(double) * (size_t)(2 * nf + 1));
    int64_t* tkey = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nf + 1));
    double* tv = (double*)malloc(sizeof(double) * 5 * (size_t)(nf + 1));
    for (int izz = 0; izz < Nz; ++izz)
      for (int iyy = 0; iyy < Ny; ++iyy)
        for (int ixx = 0; ixx < Nx; ++ixx) {
          const int c = (izz * Ny + iyy) * Nx + ixx;
          const int a = flowLab[c];
          for (int d = 0; d < 3; ++d) {
            const int nb = gidx1(ixx + (d == 0), iyy + (d == 1), izz + (d == 2), Nx, Ny, Nz);
            const int b = flowLab[nb];
            if (a == b) continue;
            const double vel = d == 0 ? u[nb] : (d == 1 ? v[nb] : w[nb]);
            const double opn = ho ? (d == 0 ? ox[nb] : (d == 1 ? oy[nb] : oz[nb])) : 1.0;
            const double area = d == 0 ? Ax : (d == 1 ? Ay : Az);
            const double q = opn * vel * area;
            if (a > 0) {
              rkey[nr] = a - 1;
              rq[nr++] = q;
            }
            if (b > 0) {
              rkey[nr] = b - 1;
              rq[nr++] = -q;
            }
            const int64_t rt = fpar[3 * (int64_t)c + d];
            if (rt >= 0) {
              const int64_t slot = slotOfFace[rt];
              const int lo = a < b ? a : b;
              if (slot >= 0) {
                const double w0 = opn * area;
                const int pk = peak[lo - 1];
                const int pkx = pk % Nx, pky = (pk / Nx) % Ny, pkz = pk / (Nx * Ny);
                double fp[3] = {oxo + (ixx + (d == 0 ? 0.5 : 0.0)) * (double)sx,
                                oyo + (iyy + (d == 1 ? 0.5 : 0.0)) * (double)sy,
                                ozo + (izz + (d == 2 ? 0.5 : 0.0)) * (double)sz};
                const double pc[3] = {oxo + pkx * (double)sx, oyo + pky * (double)sy,
                                      ozo + pkz * (double)sz};
                for (int a2 = 0; a2 < 3; ++a2) {
                  double dv = fp[a2] - pc[a2];
                  dv -= Lw[a2] * round(dv / Lw[a2]);
                  fp[a2] = dv;
                }
                tkey[ntr] = slot;
                tv[5 * ntr + 0] = (a < b ? q : -q);
                tv[5 * ntr + 1] = w0;
                tv[5 * ntr + 2] = w0 * fp[0];
                tv[5 * ntr + 3] = w0 * fp[1];
                tv[5 * ntr + 4] = w0 * fp[2];
                ++ntr;
              }
            }
          }
        }
    keyed_blocked_sums(rkey, rq, 1, nr, np, &resid);
    double* touts[5] = {Q, A, Cx, Cy, Cz};
    keyed_blocked_sums(tkey, tv, 5, ntr, nt, touts);
    free(rkey);
    free(rq);
    free(tkey);
    free(tv);
  }

  /* throat_dp (host in the reference) :958-995 */
  double* dp = (double*)malloc(sizeof(double) * (size_t)(nt ? nt : 1));
  {
    const double L[3] = {(double)Nx * spacing[0], (double)Ny * spacing[1],
                         (double)Nz * spacing[2]};
    const int N[3] = {Nx, Ny, Nz};
    for (int64_t t = 0; t < nt; ++t) {
      const int li = throats[2 * t], lj = throats[2 * t + 1];
      const pnmo_pore pi = pores[li - 1], pj = pores[lj - 1];
      const double aw = A[t];
      const int pk = peak[li - 1], pkj = peak[lj - 1];
      const double anc[3] = {origin[0] + (pk % Nx) * (double)spacing[0],
                             origin[1] + ((pk / Nx) % Ny) * (double)spacing[1],
                             origin[2] + (pk / (Nx * Ny)) * (double)spacing[2]};
      const double ancj[3] = {origin[0] + (pkj % Nx) * (double)spacing[0],
                              origin[1] + ((pkj / Nx) % Ny) * (double)spacing[1],
                              origin[2] + (pkj / (Nx * Ny)) * (double)spacing[2]};
      const double pip[3] = {(double)pi.x, (double)pi.y, (double)pi.z};
      const double pjp[3] = {(double)pj.x, (double)pj.y, (double)pj.z};
      double macro = 0.0;
      for (int a = 0; a < 3; ++a) {
        const double ct = (Cx[t] * (a == 0) + Cy[t] * (a == 1) + Cz[t] * (a == 2)) / aw;
        const long long Dc = llround((ancj[a] - anc[a]) / (double)spacing[a]) -
                             llround(ct / (double)spacing[a] + 1e-6);
        const long long k = llround((double)Dc / N[a]);
        macro += grad_p[a] * ((pjp[a] - pip[a]) - L[a] * (double)k);
      }
      dp[t] = (ppres[li - 1] - ppres[lj - 1]) - macro;
    }
  }

  out->np = np;
  out->nt = nt;
  out->pores = pores;
  out->pore_pressure = ppres;
  out->pore_residual = resid;
  out->throats = throats;
  out->throat_flow = Q;
  out->throat_area = A;
  out->throat_dp = dp;
  if (want_debug) {
    out->flow_lab = flowLab;
    out->fpar = fpar;
    out->peak = peak;
    out->cx = Cx;
    out->cy = Cy;
    out->cz = Cz;
  } else {
    free(flowLab);
    free(fpar);
    free(peak);
    free(Cx);
    free(Cy);
    free(Cz);
  }
  free(fpair);
  free(par);
  free(core);
  free(rr);
  free(slotOfFace);
  free(seg);
  free(roots);
  return 0;
}

void pnmo_flow_free(pnmo_flow_result* r) {
  free(r->pores);
  free(r->pore_pressure);
  free(r->pore_residual);
  free(r->throats);
  free(r->throat_flow);
  free(r->throat_area);
  free(r->throat_dp);
  free(r->flow_lab);
  free(r->fpar);
  free(r->peak);
  free(r->cx);
  free(r->cy);
  free(r->cz);
  memset(r, 0, sizeof(*r));
}
