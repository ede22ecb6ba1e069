// pnm_ref_omp.cpp — the CPU baseline: the reference's extract_pore_network_k
// (src/pore_extraction.hpp:1011-1025) restated one to one for OpenMP threads, as its
// Kokkos-OpenMP build would run it. Test and benchmark infrastructure only (bench.py's
// cpu_baseline and tests/), never part of the product.
//
// What is kept from the reference, construct by construct:
//   * Kokkos::View allocations are zero-filled in parallel (Kokkos initialises every View), and
//     the host vectors are copied into and out of them (uploadVec / downloadN, :37-57).
//   * MDRangePolicy<Rank<3>> over (ix, iy, iz) on a host space: Kokkos tiles the range with
//     tile 2 in every dimension but the innermost, which spans the whole extent, iterates
//     Iterate::Right (iz fastest) inside a tile and hands tiles to threads statically. The
//     restatement runs the same tiles in the same order (md_for below).
//   * RangePolicy: static chunks over the index range.
//   * atomic_min / atomic_fetch_add as CPU atomics; the `changed` flag as a relaxed store.
//   * parallel_scan: per-thread chunk sums, an exclusive scan over threads, a final pass.
//   * Fixpoint loops read the flag back after every sweep; deep_copy(labelsN, labels) each
//     flood sweep; the topology's host std::sort + std::unique.
// Arithmetic and comparisons are the reference's; the results equal the serial oracle's
// (oracle/pnm_oracle.c), pores up to the atomic-slot order.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <utility>
#include <vector>

namespace {

struct I3 {
  int x, y, z;
};
struct Pore {
  float x, y, z, radius;
};

inline int wrap(int a, int n) { return ((a % n) + n) % n; }
// get_idx (:59-64)
inline int get_idx(int x, int y, int z, I3 r) {
  return wrap(z, r.z) * r.y * r.x + wrap(y, r.y) * r.x + wrap(x, r.x);
}

template <class T>
T* view_alloc(size_t n) {  // a zero-initialised Kokkos::View<T*>
  T* p = static_cast<T*>(std::malloc(sizeof(T) * (n ? n : 1)));
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) p[i] = T{};
  return p;
}

template <class T>
void deep_copy(T* dst, const T* src, size_t n) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) dst[i] = src[i];
}

template <class T>
void fill(T* dst, T v, size_t n) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) dst[i] = v;
}

// MDRangePolicy<Rank<3>>({0,0,0}, {X,Y,Z}) on a host space: tiles of 2 x 2 x Z, tile order and
// in-tile order Iterate::Right (the last index fastest), tiles split statically over threads.
template <class F>
void md_for(I3 r, F&& f) {
  const int tx = (r.x + 1) / 2, ty = (r.y + 1) / 2;
  const int64_t tiles = (int64_t)tx * ty;
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < tiles; ++t) {
    const int bx = (int)(t / ty) * 2, by = (int)(t % ty) * 2;
    for (int ix = bx; ix < std::min(bx + 2, r.x); ++ix)
      for (int iy = by; iy < std::min(by + 2, r.y); ++iy)
        for (int iz = 0; iz < r.z; ++iz) f(ix, iy, iz);
  }
}

inline int ld(const int* p) { return __atomic_load_n(p, __ATOMIC_RELAXED); }
inline void atomic_min(int* p, int v) {
  int cur = ld(p);
  while (v < cur && !__atomic_compare_exchange_n(p, &cur, v, true, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED)) {
  }
}

// parallel_scan over [0, n): out[i] = sum_{j<i} in[j] (two passes over per-thread chunks).
void exclusive_scan(const int* in, int* out, int64_t n) {
  const int nt = omp_get_max_threads();
  std::vector<int64_t> part(nt + 1, 0);
#pragma omp parallel num_threads(nt)
  {
    const int t = omp_get_thread_num(), T = omp_get_num_threads();
    const int64_t lo = n * t / T, hi = n * (t + 1) / T;
    int64_t s = 0;
    for (int64_t i = lo; i < hi; ++i) s += in[i];
    part[t + 1] = s;
#pragma omp barrier
#pragma omp single
    for (int k = 1; k <= T; ++k) part[k] += part[k - 1];
    int run = (int)part[t];
    for (int64_t i = lo; i < hi; ++i) {
      const int v = in[i];
      out[i] = run;
      run += v;
    }
  }
}

// extractPoresView (:69-140)
std::vector<Pore> extract_pores(const float* sdf, I3 res, const float org[3],
                                const float spc[3]) {
  const float ox = org[0], oy = org[1], oz = org[2];
  const float sx = spc[0], sy = spc[1], sz = spc[2];
  const int max_pores = 1000000;
  Pore* pores = view_alloc<Pore>(max_pores);
  int counter = 0;
  md_for(res, [&](int ix, int iy, int iz) {
    const int ci = get_idx(ix, iy, iz, res);
    const float cv = sdf[ci];
    if (cv <= 0.0f) return;
    bool peak = true;
    for (int dz = -1; dz <= 1 && peak; ++dz)
      for (int dy = -1; dy <= 1 && peak; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          if (dx == 0 && dy == 0 && dz == 0) continue;
          const int ni = get_idx(ix + dx, iy + dy, iz + dz, res);
          const float nv = sdf[ni];
          if (nv > cv || (nv == cv && ni > ci)) {
            peak = false;
            break;
          }
        }
    if (!peak) return;
    float sw = 0.0f, px = 0.0f, py = 0.0f, pz = 0.0f;
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const float v = sdf[get_idx(ix + dx, iy + dy, iz + dz, res)];
          float w = v > 0.0f ? v : 0.0f;
          w = w * w;
          sw += w;
          px += dx * w;
          py += dy * w;
          pz += dz * w;
        }
    float fx = 0, fy = 0, fz = 0;
    if (sw > 1e-6f) {
      fx = px / sw;
      fy = py / sw;
      fz = pz / sw;
    }
    const int slot = __atomic_fetch_add(&counter, 1, __ATOMIC_RELAXED);
    if (slot < max_pores)
      pores[slot] = Pore{ox + (ix + fx) * sx, oy + (iy + fy) * sy, oz + (iz + fz) * sz, cv};
  });
  const int h = std::min(counter, max_pores);
  std::vector<Pore> out(pores, pores + h);  // downloadN
  std::free(pores);
  return out;
}

// segmentVolumeView (:157-386); returns the device-side seg (caller frees)
int* segment_volume(const float* sdf, I3 res, const float spc[3]) {
  const size_t n = (size_t)res.x * res.y * res.z;
  const float min_sp = std::min(spc[0], std::min(spc[1], spc[2]));
  const float thr = -1.5f * min_sp;
  int* labels = view_alloc<int>(n);
  int* roots = view_alloc<int>(n);
  int changed = 0;

  // 1. init markers (:173-179)
  md_for(res, [&](int ix, int iy, int iz) {
    const int i = get_idx(ix, iy, iz, res);
    labels[i] = (sdf[i] < thr) ? i : -1;
  });

  // 2. union-find CCL to fixpoint (:183-230)
  auto flatten = [&]() {
#pragma omp parallel for schedule(static)
    for (int64_t idx = 0; idx < (int64_t)n; ++idx) {
      int l = ld(labels + idx);
      if (l != -1) {
        while (l != ld(labels + l)) l = ld(labels + l);
        __atomic_store_n(labels + idx, l, __ATOMIC_RELAXED);
      }
    }
  };
  static const int dz_l[13] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0};
  static const int dy_l[13] = {-1, -1, -1, 0, 0, 0, 1, 1, 1, 1, 1, 1, 0};
  static const int dx_l[13] = {-1, 0, 1, -1, 0, 1, -1, 0, 1, -1, 0, 1, 1};
  int h_changed = 1;
  while (h_changed) {
    changed = 0;
    md_for(res, [&](int ix, int iy, int iz) {
      const int idx = get_idx(ix, iy, iz, res);
      const int my = ld(labels + idx);
      if (my == -1) return;
      for (int k = 0; k < 13; ++k) {
        const int ni = get_idx(ix + dx_l[k], iy + dy_l[k], iz + dz_l[k], res);
        const int nl = ld(labels + ni);
        if (nl != -1 && my != nl) {
          int rm = my;
          while (rm != ld(labels + rm)) rm = ld(labels + rm);
          int rn = nl;
          while (rn != ld(labels + rn)) rn = ld(labels + rn);
          if (rm != rn) {
            const int small = rm < rn ? rm : rn, large = rm < rn ? rn : rm;
            atomic_min(labels + large, small);
            __atomic_store_n(&changed, 1, __ATOMIC_RELAXED);
          }
        }
      }
    });
    flatten();
    h_changed = changed;
  }

  // 3. Jacobi flood to fixpoint (:237-269)
  int* labelsN = view_alloc<int>(n);
  h_changed = 1;
  while (h_changed) {
    changed = 0;
    deep_copy(labelsN, labels, n);
    md_for(res, [&](int ix, int iy, int iz) {
      const int idx = get_idx(ix, iy, iz, res);
      if (sdf[idx] >= 0.0f) return;
      if (labels[idx] != -1) return;
      int best = -1;
      for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dx = -1; dx <= 1; ++dx) {
            if (dx == 0 && dy == 0 && dz == 0) continue;
            const int nl = labels[get_idx(ix + dx, iy + dy, iz + dz, res)];
            if (nl != -1 && (best == -1 || nl < best)) best = nl;
          }
      if (best != -1) {
        labelsN[idx] = best;
        __atomic_store_n(&changed, 1, __ATOMIC_RELAXED);
      }
    });
    std::swap(labels, labelsN);
    h_changed = changed;
  }

  // 4. gradient-path basins (:271-311)
  md_for(res, [&](int ix, int iy, int iz) {
    const int ci = get_idx(ix, iy, iz, res);
    const bool ascent = (sdf[ci] > 0.0f);
    int walker = ci;
    const int MAX_STEPS = 512;
    for (int s = 0; s < MAX_STEPS; ++s) {
      int best = walker;
      float bv = sdf[walker];
      const int wx = walker % res.x, wy = (walker / res.x) % res.y, wz = walker / (res.x * res.y);
      for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dx = -1; dx <= 1; ++dx) {
            if (dx == 0 && dy == 0 && dz == 0) continue;
            const int ni = get_idx(wx + dx, wy + dy, wz + dz, res);
            const float nv = sdf[ni];
            if (ascent) {
              if (nv > bv) {
                bv = nv;
                best = ni;
              } else if (nv == bv && ni > best)
                best = ni;
            } else {
              if (nv < bv) {
                bv = nv;
                best = ni;
              } else if (nv == bv && ni > best)
                best = ni;
            }
          }
      if (best == walker) break;
      walker = best;
    }
    roots[ci] = walker;
  });

  // 5. renumbering (:313-382)
  int* seg = view_alloc<int>(n);
  int* minPoreIdx = view_alloc<int>(n);
  int* minSolidIdx = view_alloc<int>(n);
  int* poreFirst = view_alloc<int>(n);
  int* solidFirst = view_alloc<int>(n);
  int* poreRank = view_alloc<int>(n);
  int* solidRank = view_alloc<int>(n);
  int* poreId = view_alloc<int>(n);
  int* solidId = view_alloc<int>(n);
  const int kBig = 0x7fffffff;
  fill(minPoreIdx, kBig, n);
  fill(minSolidIdx, kBig, n);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    if (sdf[i] > 0.0f)
      atomic_min(minPoreIdx + roots[i], (int)i);
    else if (labels[i] != -1)
      atomic_min(minSolidIdx + labels[i], (int)i);
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    poreFirst[i] = (sdf[i] > 0.0f && minPoreIdx[roots[i]] == (int)i) ? 1 : 0;
    solidFirst[i] =
        (sdf[i] <= 0.0f && labels[i] != -1 && minSolidIdx[labels[i]] == (int)i) ? 1 : 0;
  }
  exclusive_scan(poreFirst, poreRank, (int64_t)n);
  exclusive_scan(solidFirst, solidRank, (int64_t)n);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    if (poreFirst[i]) poreId[roots[i]] = poreRank[i] + 1;
    if (solidFirst[i]) solidId[labels[i]] = -(solidRank[i] + 1);
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i)
    seg[i] = (sdf[i] > 0.0f) ? poreId[roots[i]] : (labels[i] == -1 ? 0 : solidId[labels[i]]);
  for (int* p : {labels, roots, labelsN, minPoreIdx, minSolidIdx, poreFirst, solidFirst, poreRank,
                 solidRank, poreId, solidId})
    std::free(p);
  return seg;
}

// extractTopologyView (:401-451)
std::vector<std::pair<int, int>> extract_topology(const int* seg, I3 res) {
  const size_t n = (size_t)res.x * res.y * res.z;
  const int max_pairs = (int)(n * 3);
  int* pairs = view_alloc<int>((size_t)max_pairs * 2);
  int cnt = 0;
  md_for(res, [&](int ix, int iy, int iz) {
    const int idx = get_idx(ix, iy, iz, res);
    const int my = seg[idx];
    const int dx_l[3] = {1, 0, 0}, dy_l[3] = {0, 1, 0}, dz_l[3] = {0, 0, 1};
    for (int k = 0; k < 3; ++k) {
      const int nl = seg[get_idx(ix + dx_l[k], iy + dy_l[k], iz + dz_l[k], res)];
      if (my != nl) {
        const int l1 = my < nl ? my : nl, l2 = my < nl ? nl : my;
        const int slot = __atomic_fetch_add(&cnt, 1, __ATOMIC_RELAXED);
        if (slot < max_pairs) {
          pairs[2 * slot] = l1;
          pairs[2 * slot + 1] = l2;
        }
      }
    }
  });
  const int h = std::min(cnt, max_pairs);
  std::vector<int> flat(pairs, pairs + 2 * (size_t)h);  // downloadN of the used slots
  std::free(pairs);
  std::vector<std::pair<int, int>> result;
  result.reserve(h);
  for (int i = 0; i < h; ++i) result.push_back({flat[2 * i], flat[2 * i + 1]});
  std::sort(result.begin(), result.end());
  result.erase(std::unique(result.begin(), result.end()), result.end());
  return result;
}

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

extern "C" {

int pnmr_threads(void) { return omp_get_max_threads(); }

// extract_pore_network_k (:1011-1025) on a host SDF. dims/origin/spacing XYZ. Writes up to
// `cap` pores (x, y, z, radius floats, atomic-slot order) and seg[n]; *pairs is malloc'ed
// (int32 pairs, sorted unique; pnmr_free). stage_ms (nullable): [upload, pores, segment,
// topology, download] wall milliseconds.
int64_t pnmr_extract_pore_network(const float* sdf_h, const int32_t dims[3], const float org[3],
                                  const float spc[3], float* pores_out, int64_t cap,
                                  int32_t* seg_out, int32_t** pairs, int64_t* npairs,
                                  double* stage_ms) {
  const I3 res{dims[0], dims[1], dims[2]};
  const size_t n = (size_t)res.x * res.y * res.z;
  *pairs = nullptr;
  *npairs = 0;
  if (n == 0) return 0;
  auto t = std::chrono::steady_clock::now();
  float* sdf = view_alloc<float>(n);  // Kokkos::View<float*> sdf("sdf", n) + uploadVec
  deep_copy(sdf, sdf_h, n);
  double ms[5];
  ms[0] = ms_since(t);
  t = std::chrono::steady_clock::now();
  const std::vector<Pore> pores = extract_pores(sdf, res, org, spc);
  ms[1] = ms_since(t);
  t = std::chrono::steady_clock::now();
  int* seg = segment_volume(sdf, res, spc);
  ms[2] = ms_since(t);
  t = std::chrono::steady_clock::now();
  const std::vector<std::pair<int, int>> conn = extract_topology(seg, res);
  ms[3] = ms_since(t);
  t = std::chrono::steady_clock::now();
  std::vector<int> seg_h(seg, seg + n);  // downloadN(seg)
  ms[4] = ms_since(t);
  std::memcpy(seg_out, seg_h.data(), sizeof(int) * n);
  std::free(seg);
  std::free(sdf);
  const int64_t np = (int64_t)pores.size();
  for (int64_t k = 0; k < std::min(np, cap); ++k) {
    pores_out[4 * k] = pores[k].x;
    pores_out[4 * k + 1] = pores[k].y;
    pores_out[4 * k + 2] = pores[k].z;
    pores_out[4 * k + 3] = pores[k].radius;
  }
  int32_t* pr = static_cast<int32_t*>(std::malloc(sizeof(int32_t) * 2 * (conn.size() + 1)));
  for (size_t k = 0; k < conn.size(); ++k) {
    pr[2 * k] = conn[k].first;
    pr[2 * k + 1] = conn[k].second;
  }
  *pairs = pr;
  *npairs = (int64_t)conn.size();
  if (stage_ms)
    for (int k = 0; k < 5; ++k) stage_ms[k] = ms[k];
  return np;
}

void pnmr_free(void* p) { std::free(p); }

}  // extern "C"
