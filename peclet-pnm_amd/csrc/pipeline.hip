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

// Sort the peak list by voxel index (unless it already is) and refine the first
// min(total, PNM_MAX_PORES) peaks.
static int64_t finish_pores(Ctx* c, const float* sdf, Vol v, const float org[3],
                            const float spc[3], uint32_t* peaks, int64_t npeaks, bool sorted,
                            pnm_pore_t* out_d, int64_t cap, int64_t* total) {
  *total = npeaks;
  const int64_t take = std::min<int64_t>(npeaks, PNM_MAX_PORES);
  if (take > cap) return -take;  // capacity error, caller reports
  if (!sorted && npeaks > 1)
    radix_sort_u32(c, peaks, nullptr, npeaks, 0, nbits_u64((uint64_t)(v.n - 1)));
  launch_centroids(c, sdf, v, org, spc, peaks, take, out_d);
  return take;
}

int64_t pores_device(Ctx* c, const float* sdf, Vol v, const float org[3], const float spc[3],
                     pnm_pore_t* out_d, int64_t cap, int64_t* total) {
  const int64_t nb = pre_blocks(v), scap = pre_stage_cap();
  uint32_t* stage = c->arena.get_n<uint32_t>((size_t)(nb * scap));
  uint32_t* cnt = c->arena.get_n<uint32_t>((size_t)nb);
  uint32_t* peaks = c->arena.get_n<uint32_t>((size_t)v.n);
  uint64_t* ctr = c->arena.get_n<uint64_t>(8);
  PreOut o{};
  o.peak_stage = stage;
  o.peak_cnt = cnt;
  launch_pre(c, sdf, v, 0.0f, o);
  compact_staged(c, stage, cnt, nb, scap, peaks, ctr);
  uint64_t h[1];
  readback(c, ctr, 1, h);
  return finish_pores(c, sdf, v, org, spc, peaks, (int64_t)h[0], false, out_d, cap, total);
}

struct SegInternal {
  SegResult r;  // counts read back only when segment_core is asked for them
  // bounds known on the host without a read back: pores <= pore_bound (pore keys), solids ==
  // solid_count (one id per component root)
  int64_t pore_bound, solid_count;
  uint32_t* peaks;  // peak list (extract_pores predicate) in ascending voxel order
  int64_t npeaks;
  // inputs of relabel_seg, for a caller that fuses it (seg_topology)
  const int32_t* segsrc;
  const int32_t* pore_id;
  const int32_t* solid_id;
};

// labels_out / ccl_out (nullable): caller buffers receiving the marker labels after the flood /
// after the CCL; labels_only stops after the flood (stage-level tests).
// seg == nullptr: everything up to the id assignment; the caller writes seg (fused).
static SegInternal segment_core(Ctx* c, const float* sdf, Vol v, const float spc[3],
                                int32_t* seg, int32_t* roots_out, int32_t* labels_out = nullptr,
                                int32_t* ccl_out = nullptr, bool labels_only = false,
                                bool read_counts = true) {
  const int64_t n = v.n;
  const float mn = std::min(spc[0], std::min(spc[1], spc[2]));
  const float thr = -1.5f * mn;  // pore_extraction.hpp:163-164
  const int64_t nseg = tile_segments(v), scap = tile_segment_cap();
  const int64_t nb1 = cdiv(n, 256);
  // One seg source word per voxel (kPoreFlag): the tile pass and the root resolver write the
  // flagged ascent roots of sdf>0 voxels into it, the flood the labels of the others. Separate
  // arrays instead, merged after the flood, when markers can be sdf>0 voxels (thr > 0: their
  // label and their root are both needed), when the caller wants the unflagged roots or the
  // labels of every voxel after the flood.
  const bool split = thr > 0.0f || labels_only || roots_out != nullptr;
  int32_t* segsrc = labels_out ? labels_out : c->arena.get_n<int32_t>((size_t)n);
  int32_t* roots =
      !split ? segsrc : (roots_out ? roots_out : c->arena.get_n<int32_t>((size_t)n));
  const uint32_t root_flag = split ? 0u : kPoreFlag;
  uint64_t* walk = c->arena.get_n<uint64_t>((size_t)n);
  int32_t* min_pore = c->arena.get_n<int32_t>((size_t)n);
  int32_t* min_solid = c->arena.get_n<int32_t>((size_t)n);
  // staging: the tile pass fills nseg segments of scap slots; flatten and the flood use nb1
  // blocks of 256
  const int64_t stage_words =
      std::max(std::max(nseg * scap, nb1 * 256), (int64_t)FS_BLOCKS * flood_seg_cap(n));
  uint32_t* stage_b = c->arena.get_n<uint32_t>((size_t)stage_words);
  uint32_t* stage_c = c->arena.get_n<uint32_t>((size_t)(nseg * scap));
  uint64_t* stage_o = c->arena.get_n<uint64_t>((size_t)(nseg * scap));  // open walks
  // counts: open walks from the tile pass; [solid keys | band] from the flood tile, contiguous
  // for one multi-list compaction
  uint32_t* cnt_c = c->arena.get_n<uint32_t>((size_t)nseg);
  uint32_t* cnt_pair2 = c->arena.get_n<uint32_t>((size_t)(2 * nseg));
  uint32_t* cnt_b =
      c->arena.get_n<uint32_t>((size_t)std::max(std::max(nseg, nb1), (int64_t)FS_BLOCKS));
  uint32_t* stage_r = c->arena.get_n<uint32_t>((size_t)(nseg * scap));
  uint32_t* cnt_r = c->arena.get_n<uint32_t>((size_t)nseg);
  uint32_t* peaks = c->arena.get_n<uint32_t>((size_t)n);
  uint32_t* band = c->arena.get_n<uint32_t>((size_t)n);
  uint32_t* band2 = c->arena.get_n<uint32_t>((size_t)n);
  int32_t* newlab = c->arena.get_n<int32_t>((size_t)n);
  uint32_t* skeys = c->arena.get_n<uint32_t>((size_t)n);
  uint64_t* ctr = c->arena.get_n<uint64_t>(64);
  // ctr: [0] peaks [1] (unused) [2] solid keys [3] cap flag [4] pores [5] solids [6] full keys
  //      [8..11] ring of flood band lengths
  fill_u64(c, ctr, 0, 64);

  int32_t* lab_ccl = c->arena.get_n<int32_t>((size_t)n);  // CCL labels (tile-local roots)
  TileOut o{};
  o.labels = lab_ccl;
  o.roots = roots;
  o.root_flag = root_flag;
  o.segsrc = segsrc;
  o.walk = walk;
  o.min_pore = min_pore;
  o.min_solid = min_solid;
  const int64_t nwx = cdiv(v.nx, 64), nwords = nwx * v.ny * v.nz;
  uint64_t* peak_map = c->arena.get_n<uint64_t>((size_t)nwords);
  o.peak_map = peak_map;
  int32_t* xface = c->arena.get_n<int32_t>((size_t)(2 * nwx * v.ny * v.nz));
  o.xface = reinterpret_cast<int2*>(xface);
  o.lroot_stage = stage_r;
  o.lroot_cnt = cnt_r;
  o.open_stage = stage_o;
  o.open_cnt = cnt_c;
  o.cap_flag = reinterpret_cast<unsigned long long*>(ctr + 3);
  launch_tile(c, sdf, v, thr, o);
  bitmap_list(c, peak_map, nwords, nwx, v.nx, peaks, ctr + 0);  // peaks in voxel order
  // the walks that left their tile, read where the tile pass staged them. The resolver shares
  // no data with the CCL and flood below (it writes the root words of sdf>0 voxels, which the
  // flood neither reads as labels nor writes, min_pore and the cap-flag word ctr[3]) and no
  // look-back scan state (a scan on each stream at once would race on it), so it runs beside
  // them on the second stream, joined before the counters are read back (so its outputs are
  // complete when the call returns, as every output is).
  if (!labels_only) {
    OnSide side(c);
    launch_roots_resolve(c, sdf, v, stage_o, cnt_c, nseg, (int)scap, walk, roots, root_flag,
                         min_pore, ctr + 3);
  }
  launch_ccl_seams(c, v, lab_ccl, xface);
  launch_lroot_flatten(c, v, lab_ccl, stage_r, cnt_r);
  if (ccl_out) {  // stage-level output: the flattened CCL labels before the flood
    launch_ccl_flatten(c, v, lab_ccl, min_solid, stage_b, cnt_b);
    launch_label_codes(c, v, lab_ccl, ccl_out);
  }
  // flatten + the first flood sweeps per tile; the band left over goes to the global sweeps
  launch_flood_tile(c, v, lab_ccl, segsrc, min_solid, stage_b, cnt_pair2, stage_c,
                    cnt_pair2 + nseg, xface, split);
  {
    const uint32_t* st[2] = {stage_b, stage_c};
    uint32_t* outs[2] = {skeys, band};
    uint64_t* tots[2] = {ctr + 2, ctr + 8};
    compact_staged_multi(c, 2, st, cnt_pair2, nseg, scap, outs, tots);
  }
  join_side(c);

  // Remaining Jacobi sweeps until a sweep labels nothing (pore_extraction.hpp:238-269), with
  // no cap on their number. Band lengths live on the device in a ring of four counter words
  // (ctr[8 + s % 4] is the length sweep s reads, ctr[8 + (s + 1) % 4] the one it writes); the
  // sweeps size their work from them, and the host reads the counter block every two sweeps to
  // stop when a sweep labels nothing. The first read also brings the list sizes of the tile
  // passes and the walk-cap flag (set by the resolver, joined above).
  uint64_t h[12];
  {
    uint32_t* cur = band;
    uint32_t* nxt = band2;
    int64_t s = 0;
    int64_t bound = n;  // band length bound for the staging segments (the band only shrinks)
    const size_t sweep_mark = c->arena.mark();  // every sweep pair reuses the same scratch
    while (true) {
      c->arena.rewind(sweep_mark);
      for (int k = 0; k < 2; ++k, ++s) {
        flood_sweep(c, v, segsrc, cur, ctr + 8 + (s & 3), bound, nxt, ctr + 8 + ((s + 1) & 3),
                    newlab, stage_b, cnt_b, min_solid);
        std::swap(cur, nxt);
      }
      readback(c, ctr, 12, h);
      const uint64_t before = h[8 + ((s - 1) & 3)], after = h[8 + (s & 3)];
      if (after == before || after == 0) break;
      bound = (int64_t)after;
    }
  }
  const int64_t npeaks = (int64_t)h[0], nskeys = (int64_t)h[2];
  h[0] = h[3];  // the walk-cap flag

  SegInternal out{};
  if (labels_only) return out;
  if (split) launch_merge_roots(c, sdf, v, roots, segsrc);
  if (h[0] & 2u) fail(PNM_EHIP, "internal error: ascent walk record missing");
  const uint32_t* pkeys = peaks;
  int64_t npk = npeaks;
  if (h[0] & 1u) {  // a walk stopped at the 512-step cap on a non-fixed point
    launch_min_full(c, sdf, v, segsrc, min_pore);
    uint32_t* full = c->arena.get_n<uint32_t>((size_t)n);
    npk = collect_keys_full(c, sdf, v, min_pore, full, ctr + 6);
    pkeys = full;
  }
  assign_ids_bitmap(c, sdf, v, pkeys, npk, min_pore, skeys, nskeys, min_solid, ctr + 4);
  if (seg) launch_seg(c, v, segsrc, min_pore, min_solid, seg);
  if (roots_out) launch_descent_roots(c, sdf, v, roots_out);  // rootsOut of sdf <= 0 voxels
  if (read_counts) {
    readback(c, ctr + 4, 2, h);
    out.r.n_pores = (int32_t)h[0];
    out.r.n_solids = (int32_t)h[1];
  }
  out.pore_bound = npk;
  out.solid_count = nskeys;
  out.peaks = peaks;
  out.npeaks = npeaks;
  out.segsrc = segsrc;
  out.pore_id = min_pore;
  out.solid_id = min_solid;
  return out;
}

SegResult segment_device(Ctx* c, const float* sdf, Vol v, const float spc[3], int32_t* seg,
                         int32_t* roots_out) {
  return segment_core(c, sdf, v, spc, seg, roots_out).r;
}

void labels_device(Ctx* c, const float* sdf, Vol v, const float spc[3], int32_t* labels,
                   int32_t* ccl_only) {
  segment_core(c, sdf, v, spc, nullptr, nullptr, labels, ccl_only, true);
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
  // the topology keys need only label bounds, so the id counts are not read back
  SegInternal s = segment_core(c, sdf, v, spc, nullptr, nullptr, nullptr, nullptr, false, false);
  const int64_t np = finish_pores(c, sdf, v, org, spc, s.peaks, s.npeaks, true, pores_d,
                                  cap_pores, total_pores);
  int64_t needed = 0;
  const int64_t nt = seg_topology(c, s.segsrc, s.pore_id, s.solid_id, seg, v, -s.solid_count,
                                  s.pore_bound, pairs_d, cap_pairs, &needed);
  *npairs = needed;
  *pairs_overflow = nt < 0;
  return np;
}

}  // namespace pnm
