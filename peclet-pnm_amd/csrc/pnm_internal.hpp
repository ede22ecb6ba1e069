// pnm_internal.hpp — shared host/device internals of libpnm_hip.so (gfx950).
//
// Context = one HIP stream + a slot arena of device scratch + pinned readback words.
// Kernels live in k_*.hip; each exposes host launchers declared here. Host orchestration of the
// reference pipelines lives in pipeline.hip, the C-ABI in capi.cpp.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "pnm_capi.h"

namespace pnm {

struct Error {
  pnm_status code;
  std::string msg;
};

[[noreturn]] void fail(pnm_status code, const std::string& msg);
void hip_check(hipError_t e, const char* what, const char* file, int line);
#define PNM_HIP(x) ::pnm::hip_check((x), #x, __FILE__, __LINE__)

// Grow-only slot arena: the k-th request of a call reuses the k-th slot when it is big enough.
// A pipeline makes the same sequence of requests every call, so after the first call of a given
// size no hipMalloc/hipFree happens. Memory is never initialised by the arena.
// A loop whose iterations make the same requests brackets them with mark()/rewind(), so its
// iterations share slots (stream order keeps an iteration's reuse after the previous one).
// Fixed mode (fix/unfix): requests are carved from one caller-given device region by a bump
// offset instead, and a request beyond its end fails with PNM_ECAPACITY; nothing is allocated.
class Arena {
 public:
  void reset() {
    next_ = 0;
    off_ = 0;
  }
  void* get(size_t bytes);
  template <class T>
  T* get_n(size_t count) {
    return static_cast<T*>(get(count * sizeof(T)));
  }
  size_t mark() const { return base_ ? off_ : next_; }
  void rewind(size_t m) {
    if (base_)
      off_ = m;
    else
      next_ = m;
  }
  void fix(void* base, size_t cap) {
    base_ = static_cast<char*>(base);
    cap_ = cap;
    off_ = 0;
    high_ = 0;
  }
  void unfix() { base_ = nullptr; }
  size_t high_water() const { return high_; }  // largest bump offset reached in fixed mode
  void release();
  int64_t bytes() const;

 private:
  struct Slot {
    void* p;
    size_t size;
  };
  std::vector<Slot> slots_;
  size_t next_ = 0;
  char* base_ = nullptr;  // fixed mode when set
  size_t cap_ = 0, off_ = 0, high_ = 0;
};

struct Ctx {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  Arena arena;
  uint64_t* pinned = nullptr;  // hipHostMalloc'ed readback words
  static constexpr int kPinnedWords = 64;
  // Error word in mapped pinned memory: kernels set it on an internal failure (a timed-out
  // wait); readback() reports it after every synchronisation.
  uint32_t* err_host = nullptr;
  uint32_t* err_dev = nullptr;
  // Look-back scan state: per-block status words (grow-only) and [ticket, done, epoch].
  uint64_t* lb_state = nullptr;
  int64_t lb_cap = 0;
  uint32_t* lb_ctl = nullptr;
  // Set while a call runs on caller-provided fixed scratch (no allocation allowed): the
  // look-back state must not grow and no second stream is forked.
  bool lb_fixed = false;
  bool no_fork = false;
  // Kernel timing: when `prof` is set every launch is bracketed by two HIP events on the
  // launch stream (only launches named in `prof_only`, newline-separated, when it is not
  // empty); pnm_ctx_profile_read aggregates them per kernel name.
  bool prof = false;
  std::string prof_only;
  bool timed(const char* name) const;
  struct Rec {
    const char* name;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  hipEvent_t take_event();
  // Second stream (default priority) for work that shares no buffers with the main chain, and
  // the events that fork it from `stream` and join it back; created on first use. side_pending:
  // work was forked and the main stream has not been made to wait for it yet.
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool side_pending = false;
  // Batched segmentation: autotuned launch shape and the cached graph of the last batch.
  bool batch_cfg_valid = false;
  int batch_block = 1024, batch_per_cu = 2, batch_groups = 0;
  uint64_t batch_shape = 0;
  float batch_tuned_ms = 0.0f;
  void* batch_graph = nullptr;  // hipGraphExec_t
  uint64_t batch_key = 0;
  // Workgroup count the scratch was sized for, per batch shape: later calls of the same shape
  // keep it (same arena requests, same buffers, the same captured graph).
  int batch_scratch_groups = 0;
  uint64_t batch_scratch_shape = 0;
  // Throat surrogate: activations of the last forward pass, kept for its backward pass.
  Arena sur;
  void* sur_acts = nullptr;
};

struct KTimer {
  KTimer(Ctx* c, const char* name);
  ~KTimer();
  Ctx* c;
  const char* name;
  hipEvent_t a = nullptr;
};

// Launch on the context's stream, timed when profiling is on, error-checked.
#define PNM_LAUNCH(c, name, kern, grid, block, shm, ...)                         \
  do {                                                                           \
    ::pnm::KTimer pnm_kt_((c), (name));                                          \
    hipLaunchKernelGGL(kern, grid, block, shm, (c)->stream, __VA_ARGS__);        \
    PNM_HIP(hipGetLastError());                                                  \
  } while (0)

// Launches made while an OnSide lives go to the context's second stream, after everything
// queued on the main stream so far. join_side makes the main stream wait for them (the end of
// every call does it too, so no forked work outlives the call that forked it).
// With ctx->no_fork set, an OnSide does nothing: its launches stay on the main stream.
struct OnSide {
  explicit OnSide(Ctx* c);
  ~OnSide();
  Ctx* ctx;
  hipStream_t main;
  bool active = false;
};
void join_side(Ctx* c);

// RAII: make ctx->device current, reset the arena for a new call.
struct CallScope {
  explicit CallScope(Ctx* c);
  ~CallScope();
  Ctx* ctx;
  int prev_device;
};

struct Vol {
  int nx, ny, nz;
  int64_t n;
};

// ---------- small device utilities (k_util.hip) ----------
void fill_u32(Ctx* c, uint32_t* p, uint32_t v, int64_t count);
void fill_u64(Ctx* c, uint64_t* p, uint64_t v, int64_t count);
// Copies `words` device uint64 counters to pinned memory and synchronises the stream.
void readback(Ctx* c, const void* dev, int words, uint64_t* host_out);
// Exclusive prefix sum of uint32 (n up to 2^31); returns nothing, out may alias in.
// Sorted flat indices of the set bits of a row-padded bitmap (word w: voxels (w / nwx) * nx +
// (w % nwx) * 64 + bit) and their count (uint64 at total_dev), in one launch.
void bitmap_list(Ctx* c, const uint64_t* map, int64_t nwords, int64_t nwx, int64_t nx,
                 uint32_t* out, uint64_t* total_dev);
// Rank directory of a uint32 bitmap: chunk_rank[c] = set bits in words [0, 32 c) (one entry per
// MAP_CHUNK_WORDS = 32 words), total at total_dev; one launch.
constexpr int MAP_CHUNK_WORDS = 32;
void map_rank(Ctx* c, const uint32_t* map, int64_t words, uint32_t* chunk_rank,
              uint64_t* total_dev);
// Exclusive scan of n uint32 values (in == out allowed); the total, when total_dev is given, is
// written there as uint64.
void scan_exclusive_u32(Ctx* c, const uint32_t* in, uint32_t* out, int64_t n,
                        uint64_t* total_dev = nullptr);
// Stable LSD radix sort by keys[begin_bit, end_bit); vals optional (nullptr). Result lands in
// keys/vals (the alt buffers are scratch).
void radix_sort_u64(Ctx* c, uint64_t* keys, uint32_t* vals, int64_t n, int begin_bit,
                    int end_bit);
void radix_sort_u32(Ctx* c, uint32_t* keys, uint32_t* vals, int64_t n, int begin_bit,
                    int end_bit);
// Gathers block-staged items (stage[b*cap + k], k < counts[b]) in block order into out; the
// total lands in *total_dev (device word).
void compact_staged(Ctx* c, const uint32_t* stage, const uint32_t* counts, int64_t nblocks,
                    int64_t cap, uint32_t* out, uint64_t* total_dev);
// K (<= 4) staged lists sharing one segment layout, counts contiguous (list k at
// counts + k * nseg), compacted with one scan and one gather; totals written per list.
void compact_staged_multi(Ctx* c, int K, const uint32_t* const* stage, const uint32_t* counts,
                          int64_t nseg, int64_t cap, uint32_t* const* out,
                          uint64_t* const* total_dev);
void compact_staged64(Ctx* c, const uint64_t* stage, const uint32_t* counts, int64_t nblocks,
                      int64_t cap, uint64_t* out, uint64_t* total_dev);
// Writes keys[i] with keys[i] != keys[i-1] to out; returns the count (synchronises).
// Unique of a sorted list into out; the count is left on the device (*count_dev), no host
// synchronisation.
void unique_sorted_u64_dev(Ctx* c, const uint64_t* keys, uint64_t* out, int64_t n,
                           uint64_t* count_dev);

// ---------- plain pipeline kernels (k_extract.hip, k_segment.hip, k_topology.hip) ----------
struct PreOut {
  int32_t* labels;        // markers (x-run start index) or -1 [n], nullable
  int32_t* next;          // ascent successor                  [n], nullable
  int32_t* min_pore;      // set to the peak's own index at sdf>0 peaks [n], nullable
  uint32_t* peak_stage;   // staged peaks (extract_pores predicate): pre_stage_cap() per segment
  uint32_t* peak_cnt;     // per-segment peak counts
  uint32_t* band_stage;   // staged flood candidates, nullable
  uint32_t* band_cnt;     // per-segment band counts, nullable
};
// Number of staging segments launch_pre fills for a volume, and the slots per segment.
int64_t pre_blocks(Vol v);
int64_t pre_stage_cap();
void launch_pre(Ctx* c, const float* sdf, Vol v, float thr, const PreOut& o);
void launch_centroids(Ctx* c, const float* sdf, Vol v, const float org[3], const float spc[3],
                      const uint32_t* peaks, int64_t count, pnm_pore_t* out);

// CCL of markers: x-run unions + flatten; flatten stages the component roots (solid keys) per
// 256-voxel block and sets min_solid[root] = root.
// Seg source word of a voxel, the one int32 relabel_seg (:378-381) needs per voxel: for an sdf>0
// voxel its ascent root with the top bit set (root | kPoreFlag; a root is a flat index below
// 2^31 - 1, so the word is never -1), for every other voxel its marker label after the flood (a
// flat index >= 0) or -1. As unsigned numbers root words and -1 both lie above every label, so
// the unsigned minimum of neighbour words is the smallest label among them, if there is one.
constexpr uint32_t kPoreFlag = 0x80000000u;

// Segmentation tile pass (k_tile.hip).
struct TileOut {
  // marker label = flat index of its tile-local CCL root; for a voxel that is no marker -1 if
  // !(sdf >= 0) (a flood candidate), else -2                                             [n]
  int32_t* labels;
  // Ascent roots of sdf>0 voxels whose walk ends inside their tile, as root | root_flag [n].
  // With root_flag = kPoreFlag this is the seg source array (see kPoreFlag), which the flood
  // fills in for the other voxels; the tile pass writes -1 into `segsrc` at sdf == 0 voxels.
  int32_t* roots;
  uint32_t root_flag;
  int32_t* segsrc;
  uint64_t* walk;         // walk records: sdf>0 voxels on tile boundaries or left open [n]
  int32_t* min_pore;      // first-appearance minimum over the tile, at in-tile roots   [n]
  int32_t* min_solid;     // seeded with its own index at every tile-local marker root  [n]
  uint64_t* peak_map;     // extract_pores peaks: bit x%64 of word (z*ny + y)*cdiv(nx,64) + x/64
  int2* xface;            // labels of each row chunk's first/last column, entry (chunk, z, y)
  uint32_t* lroot_stage;  // staged tile-local marker roots
  uint32_t* lroot_cnt;
  uint64_t* open_stage;   // staged walks that left the tile: record | (slice, column) << 50
  uint32_t* open_cnt;
  unsigned long long* cap_flag;  // set when a walk stops at the cap on a non-fixed point
};
int64_t tile_segments(Vol v);
int64_t tile_segment_cap();
void launch_tile(Ctx* c, const float* sdf, Vol v, float thr, const TileOut& o);
// Cross-tile and wrapping marker relations; xface: the tile pass's x-face labels (int2 per
// (chunk, z, y)).
void launch_ccl_seams(Ctx* c, Vol v, int32_t* labels, const int32_t* xface);
// After the seams: every tile-local root (staged by the tile pass) -> its component root.
void launch_lroot_flatten(Ctx* c, Vol v, int32_t* labels, const uint32_t* stage,
                          const uint32_t* cnt);
// Walks that left their tile, read from the tile pass's staging (nseg segments of seg_cap slots,
// counts in cnt): one wave per segment.
void launch_roots_resolve(Ctx* c, const float* sdf, Vol v, const uint64_t* stage,
                          const uint32_t* cnt, int64_t nseg, int seg_cap, const uint64_t* walk,
                          int32_t* roots, uint32_t root_flag, int32_t* min_pore,
                          uint64_t* cap_flag);
void launch_ccl_flatten(Ctx* c, Vol v, int32_t* labels, int32_t* min_solid, uint32_t* key_stage,
                        uint32_t* key_cnt);
// Flatten + the first flood sweep(s) per tile: reads the CCL labels (tile-local roots whose
// own entries already hold their component root), writes the flooded labels to a second array
// (segments as tile_segments / tile_segment_cap). xface: the tile pass's x-face labels.
// write_all = false: voxels that are neither labelled nor flood candidates (sdf >= 0) are not
// written (the seg source array holds their ascent roots, or -1 at sdf == 0, already).
void launch_flood_tile(Ctx* c, Vol v, const int32_t* labels_ccl, int32_t* labels,
                       int32_t* min_solid, uint32_t* key_stage, uint32_t* key_cnt,
                       uint32_t* band_stage, uint32_t* band_cnt, const int32_t* xface,
                       bool write_all);
// out[i] = max(labels[i], -1): the tile pass's CCL labels with the reference's -1.
void launch_label_codes(Ctx* c, Vol v, const int32_t* labels, int32_t* out);
// segsrc[i] = roots[i] | kPoreFlag at every sdf>0 voxel (the others keep their label).
void launch_merge_roots(Ctx* c, const float* sdf, Vol v, const int32_t* roots, int32_t* segsrc);
void launch_ccl(Ctx* c, Vol v, int32_t* labels, int32_t* min_solid, uint32_t* key_stage,
                uint32_t* key_cnt);
// Grid of the global flood sweeps (k_flood_sweep): block b takes the b-th share of the band.
constexpr unsigned FS_BLOCKS = 1024;
// One Jacobi flood sweep over the band list whose length is the device word *count (at most
// max_items; stage holds FS_BLOCKS * flood_seg_cap(max_items) slots, stage_cnt FS_BLOCKS);
// writes the remaining band list and its length. No host synchronisation.
int64_t flood_seg_cap(int64_t max_items);
void flood_sweep(Ctx* c, Vol v, int32_t* labels, const uint32_t* band, const uint64_t* count,
                 int64_t max_items, uint32_t* band_next, uint64_t* count_next, int32_t* newlab,
                 uint32_t* stage, uint32_t* stage_cnt, int32_t* min_solid);
// Capped-walk fallback: full reinitialisation of the pore minima from the seg source array.
void launch_min_full(Ctx* c, const float* sdf, Vol v, const int32_t* segsrc, int32_t* min_pore);
int64_t collect_keys_full(Ctx* c, const float* sdf, Vol v, const int32_t* min_pore,
                          uint32_t* key_list, uint64_t* counter);
// Pore and solid ids of both key sets ranked by first appearance (id arrays hold each key's
// first-appearance voxel on entry); writes counts_dev[0] = pore ids, counts_dev[1] = solid ids
// (no synchronisation).
void assign_ids_bitmap(Ctx* c, const float* sdf, Vol v, const uint32_t* pkeys, int64_t npk,
                       int32_t* pore_ids, const uint32_t* skeys, int64_t nsk, int32_t* solid_ids,
                       uint64_t* counts_dev);
void launch_seg(Ctx* c, Vol v, const int32_t* segsrc, const int32_t* pore_id,
                const int32_t* solid_id, int32_t* seg);
// Gradient-walk successors and capped chases, computed stand-alone. Modes: kWalkPores = the
// reference's step 4 (ascent from sdf>0, descent from the rest), kWalkAll = ascent from every
// voxel (flow basins), kWalkSolids = descent from !(sdf>0) voxels only (chase writes only those).
constexpr int kWalkPores = 0, kWalkAll = 1, kWalkSolids = 2;
void launch_next(Ctx* c, const float* sdf, Vol v, int32_t* next, int mode);
void launch_chase(Ctx* c, const float* sdf, Vol v, const int32_t* next, int32_t* roots,
                  int mode);
// Descent roots of the !(sdf>0) voxels into roots (the rest of rootsOut, :383-384).
void launch_descent_roots(Ctx* c, const float* sdf, Vol v, int32_t* roots);

// Topology: block-deduplicated candidate keys, then sort + unique. key = (l1 - lo) << bits |
// (l2 - lo). Returns count; pairs_d receives (l1,l2) int32 pairs.
int64_t topology(Ctx* c, const int32_t* seg, Vol v, int32_t lo_label, int32_t hi_label,
                 bool labels_known, int32_t* pairs_d, int64_t cap, int64_t* needed);
// relabel_seg fused into the topology candidates: writes seg and returns the pair count as
// topology() does.
int64_t seg_topology(Ctx* c, const int32_t* segsrc, const int32_t* pore_id,
                     const int32_t* solid_id, int32_t* seg, Vol v, int32_t lo_label,
                     int32_t hi_label, int32_t* pairs_d, int64_t cap, int64_t* needed);

// ---------- pipelines (pipeline.hip) ----------
struct SegResult {
  int32_t n_pores;   // pore labels 1..n_pores
  int32_t n_solids;  // solid labels -1..-n_solids
};
SegResult segment_device(Ctx* c, const float* sdf, Vol v, const float spc[3], int32_t* seg,
                         int32_t* roots_out);
// Marker labels after CCL (ccl_only, nullable) and after the flood (labels).
void labels_device(Ctx* c, const float* sdf, Vol v, const float spc[3], int32_t* labels,
                   int32_t* ccl_only);
// Segmentation that also hands back the (unsorted) device peak list of the extract_pores
// predicate; the sdf>0 entries are exactly the self-roots.
struct SegPeaks {
  int32_t n_pores;
  int32_t n_solids;
  const uint32_t* peaks;
  int64_t npeaks;
};
SegPeaks segment_with_peaks(Ctx* c, const float* sdf, Vol v, const float spc[3], int32_t* seg,
                            int32_t* roots_out);
// Returns the number of pores written; a negative value -k means k records did not fit `cap`.
int64_t pores_device(Ctx* c, const float* sdf, Vol v, const float org[3], const float spc[3],
                     pnm_pore_t* out_d, int64_t cap, int64_t* total);
int64_t network_device(Ctx* c, const float* sdf, Vol v, const float org[3], const float spc[3],
                       pnm_pore_t* pores_d, int64_t cap_pores, int64_t* total_pores,
                       int32_t* seg, int32_t* pairs_d, int64_t cap_pairs, int64_t* npairs,
                       bool* pairs_overflow);

// ---------- batched segmentation (k_batch.hip) ----------
void segment_batch_device(Ctx* c, int nvol, const float* const* sdf, int32_t* const* seg,
                          const int32_t* dims, const float spc[3], int flags, int32_t* out_np,
                          int32_t* out_ns);

// ---------- network flow (k_flow.hip) ----------
// Throat geometry without fields (binary openness from the SDF: a face is open iff both cells
// have sdf > 0): pores, throats, areas and centroids only, equal to extract_network_flow's under
// that openness.
struct GeoMode {
  const int32_t* seg = nullptr;  // a given segmentation of the SDF (nullable: segment here)
  float* centers = nullptr;      // nullable: per-throat surrogate inputs when nt <= cap,
  float* raw = nullptr;          //   centroids [nt][3] (voxel X,Y,Z) and raw scalars [nt][4]
  int64_t cap = 0;
};
// Fields are device pointers; ox/oy/oz all null = fully open. Results are malloc'ed host arrays.
// geo != nullptr: throat geometry only (fields unused; the flow, dp, pressure and residual
// arrays of `out` stay null).
void flow_device(Ctx* c, const float* sdf, Vol v, const float org[3], const float spc[3],
                 const double* u, const double* vv, const double* w, const double* p,
                 const double* ox, const double* oy, const double* oz, const double grad_p[3],
                 pnm_flow_result_t* out, const GeoMode* geo = nullptr);

// ---------- project MAC Stokes solver (k_stokes.hip) ----------
struct StokesInfo {
  int iterations;
  double rel_residual;  // ||b - M x|| / ||b||
};
StokesInfo stokes_device(Ctx* c, const float* sdf, Vol v, const float spc[3],
                         const double force[3], double mu, double rtol, int max_iter, double* u,
                         double* vv, double* w, double* p, double* ox, double* oy, double* oz);

// ---------- throat-conductance surrogate (k_surrogate.hip) ----------
int64_t surrogate_param_count();
void surrogate_init(uint64_t seed, float* host_params);
void surrogate_forward(Ctx* c, int B, const float* patches, const float* scalars,
                       const float* params, float* out);
void surrogate_backward(Ctx* c, int B, const float* dy, const float* params, float* grads);
void surrogate_loss(Ctx* c, int B, const float* y, const float* t, float delta, float* dy,
                    float* loss);
void surrogate_adam(Ctx* c, float* params, const float* grads, float* m, float* v, int step,
                    float lr, float b1, float b2, float eps);
void surrogate_patches(Ctx* c, const float* sdf, Vol v, const float* centers, int B, float scale,
                       float* out);
void surrogate_release(Ctx* c);
// Fused inference: log g (before the target offset) of B throats straight from the SDF: patch
// gather around centers (xyz voxel coordinates), the three convolutions, pooling and the head.
void surrogate_infer(Ctx* c, const float* sdf, Vol v, const float* centers, const float* scalars,
                     int B, float scale, const float* params, float* out);

}  // namespace pnm
