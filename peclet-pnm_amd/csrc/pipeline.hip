// pipeline.hip — host orchestration of the plain pipeline on one HIP stream.
//
//   k_pre (one SDF pass: peaks, markers, ascent successor, flood band)
//   -> CCL union + flatten -> flood sweeps over the shrinking band
//   -> roots + first-appearance minima -> key ranking -> seg
//   -> topology (block-deduplicated pairs, radix sort, unique)
//
// The only host synchronisations are counter readbacks: after CCL (list sizes), once per flood
// sweep (the reference's `changed` flag, pore_extraction.hpp:266-268), after the root walk (cap
// flag), the solid-id count and the topology candidate/unique counts.
#include <algorithm>

#include "k_common.hpp"

namespace pnm {

static int nbits_u64(uint64_t v) {
  int b = 0;
  while (b < 64 && (v >> b) != 0) ++b;
  return b;
}

// Sort the peak list by voxel index and refine the first min(total, PNM_MAX_PORES) peaks.
static int64_t finish_pores(Ctx* c, const float* sdf, Vol v, const float org[3],
                            const float spc[3], uint32_t* peaks, int64_t npeaks,
                            pnm_pore_t* out_d, int64_t cap, int64_t* total) {
  *total = npeaks;
  const int64_t take = std::min<int64_t>(npeaks, PNM_MAX_PORES);
  if (take > cap) return -take;  // capacity error, caller reports
  if (npeaks > 1) radix_sort_u32(c, peaks, nullptr, npeaks, 0, nbits_u64((uint64_t)(v.n - 1)));
  launch_centroids(c, sdf, v, org, spc, peaks, take, out_d);
  return take;
}

int64_t pores_device(Ctx* c, const float* sdf, Vol v, const float org[3], const float spc[3],
                     pnm_pore_t* out_d, int64_t cap, int64_t* total) {
  uint32_t* peaks = c->arena.get_n<uint32_t>((size_t)v.n);
  uint64_t* ctr = c->arena.get_n<uint64_t>(8);
  fill_u64(c, ctr, 0, 8);
  PreOut o{};
  o.peak_list = peaks;
  o.cap_peaks = v.n;
  o.counters = ctr;
  launch_pre(c, sdf, v, 0.0f, o);
  uint64_t h[1];
  readback(c, ctr, 1, h);
  return finish_pores(c, sdf, v, org, spc, peaks, (int64_t)h[0], out_d, cap, total);
}

struct SegInternal {
  SegResult r;
  uint32_t* peaks;  // unsorted peak list (extract_pores predicate)
  int64_t npeaks;
};

static SegInternal segment_core(Ctx* c, const float* sdf, Vol v, const float spc[3],
                                int32_t* seg, int32_t* roots_out) {
  const int64_t n = v.n;
  const float mn = std::min(spc[0], std::min(spc[1], spc[2]));
  const float thr = -1.5f * mn;  // pore_extraction.hpp:163-164
  int32_t* labels = c->arena.get_n<int32_t>((size_t)n);
  int32_t* next = c->arena.get_n<int32_t>((size_t)n);
  int32_t* roots = roots_out ? roots_out : c->arena.get_n<int32_t>((size_t)n);
  int32_t* min_pore = c->arena.get_n<int32_t>((size_t)n);
  int32_t* min_solid = c->arena.get_n<int32_t>((size_t)n);
  uint32_t* peaks = c->arena.get_n<uint32_t>((size_t)n);
  uint32_t* band = c->arena.get_n<uint32_t>((size_t)n);
  uint32_t* band2 = c->arena.get_n<uint32_t>((size_t)n);
  int32_t* newlab = c->arena.get_n<int32_t>((size_t)n);
  uint32_t* skeys = c->arena.get_n<uint32_t>((size_t)n);
  uint64_t* ctr = c->arena.get_n<uint64_t>(8);
  fill_u64(c, ctr, 0, 8);

  PreOut o{};
  o.labels = labels;
  o.next = next;
  o.min_pore = min_pore;
  o.peak_list = peaks;
  o.band_list = band;
  o.counters = ctr;
  o.cap_peaks = n;
  launch_pre(c, sdf, v, thr, o);
  launch_ccl(c, v, labels, min_solid, skeys, ctr);
  uint64_t h[5];
  readback(c, ctr, 5, h);
  const int64_t npeaks = (int64_t)h[0], nband = (int64_t)h[1], nskeys = (int64_t)h[2];
  int64_t npos = (int64_t)h[4];

  // Jacobi flood until a sweep labels nothing (pore_extraction.hpp:238-269)
  {
    uint32_t* cur = band;
    uint32_t* nxt = band2;
    int64_t nb = nband;
    int64_t sweeps = 0;
    while (nb > 0) {
      const int64_t nb2 = flood_sweep(c, v, labels, cur, nb, nxt, newlab, ctr + 5);
      if (++sweeps > n + 2) fail(PNM_ENOCONVERGE, "flood did not converge");
      if (nb2 == nb) break;
      std::swap(cur, nxt);
      nb = nb2;
    }
  }

  launch_roots_min(c, sdf, v, next, labels, roots, min_pore, min_solid, ctr, true);
  readback(c, ctr + 3, 1, h);
  const uint32_t* pkeys = peaks;
  int64_t npk = npeaks;
  if (h[0]) {  // a walk stopped at the 512-step cap on a non-fixed point
    launch_min_full(c, sdf, v, roots, labels, min_pore, min_solid);
    uint32_t* full = c->arena.get_n<uint32_t>((size_t)n);
    npk = collect_keys_full(c, sdf, v, min_pore, full, ctr + 6);
    pkeys = full;
    npos = npk;
  }
  assign_ids(c, sdf, pkeys, npk, min_pore, false);
  const int64_t ns = assign_ids(c, sdf, skeys, nskeys, min_solid, true);
  launch_seg(c, sdf, v, roots, labels, min_pore, min_solid, seg);
  SegInternal s;
  s.r.n_pores = (int32_t)npos;
  s.r.n_solids = (int32_t)ns;
  s.peaks = peaks;
  s.npeaks = npeaks;
  return s;
}

SegResult segment_device(Ctx* c, const float* sdf, Vol v, const float spc[3], int32_t* seg,
                         int32_t* roots_out) {
  return segment_core(c, sdf, v, spc, seg, roots_out).r;
}

SegPeaks segment_with_peaks(Ctx* c, const float* sdf, Vol v, const float spc[3], int32_t* seg,
                            int32_t* roots_out) {
  const SegInternal s = segment_core(c, sdf, v, spc, seg, roots_out);
  return SegPeaks{s.r.n_pores, s.r.n_solids, s.peaks, s.npeaks};
}

// Fused pipeline: returns pore count (negative = capacity needed) and pair count via args.
int64_t network_device(Ctx* c, const float* sdf, Vol v, const float org[3], const float spc[3],
                       pnm_pore_t* pores_d, int64_t cap_pores, int64_t* total_pores,
                       int32_t* seg, int32_t* pairs_d, int64_t cap_pairs, int64_t* npairs,
                       bool* pairs_overflow) {
  SegInternal s = segment_core(c, sdf, v, spc, seg, nullptr);
  const int64_t np = finish_pores(c, sdf, v, org, spc, s.peaks, s.npeaks, pores_d, cap_pores,
                                  total_pores);
  int64_t needed = 0;
  const int64_t nt =
      topology(c, seg, v, -s.r.n_solids, s.r.n_pores, true, pairs_d, cap_pairs, &needed);
  *npairs = needed;
  *pairs_overflow = nt < 0;
  return np;
}

}  // namespace pnm
