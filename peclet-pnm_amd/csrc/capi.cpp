// capi.cpp — the extern "C" boundary of libpnm_hip.so (declared in include/pnm_capi.h).
//
// Every entry point validates its arguments before touching the device, converts C++
// exceptions into a pnm_status plus a thread-local message, and runs on the context's stream.
#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <new>
#include <string>

#include "pnm_internal.hpp"

namespace pnm {

void fail(pnm_status code, const std::string& msg) { throw Error{code, msg}; }

void hip_check(hipError_t e, const char* what, const char* file, int line) {
  if (e == hipSuccess) return;
  const pnm_status code =
      (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? PNM_ENOMEM : PNM_EHIP;
  fail(code, std::string(what) + " failed: " + hipGetErrorString(e) + " (" + file + ":" +
                 std::to_string(line) + ")");
}

void* Arena::get(size_t bytes) {
  bytes = (bytes + 255) & ~size_t(255);
  if (bytes == 0) bytes = 256;
  if (base_) {
    if (bytes > cap_ - off_)
      fail(PNM_ECAPACITY, "fixed scratch exhausted: request of " + std::to_string(bytes) +
                              " bytes at offset " + std::to_string(off_) + " of " +
                              std::to_string(cap_));
    void* p = base_ + off_;
    off_ += bytes;
    high_ = std::max(high_, off_);
    return p;
  }
  if (next_ < slots_.size()) {
    Slot& s = slots_[next_];
    if (s.size >= bytes) {
      ++next_;
      return s.p;
    }
    PNM_HIP(hipFree(s.p));
    s.p = nullptr;
    s.size = 0;
    void* p = nullptr;
    PNM_HIP(hipMalloc(&p, bytes));
    s.p = p;
    s.size = bytes;
    ++next_;
    return p;
  }
  void* p = nullptr;
  PNM_HIP(hipMalloc(&p, bytes));
  slots_.push_back({p, bytes});
  ++next_;
  return p;
}

void Arena::release() {
  for (auto& s : slots_)
    if (s.p) (void)hipFree(s.p);
  slots_.clear();
  next_ = 0;
}

int64_t Arena::bytes() const {
  int64_t b = 0;
  for (const auto& s : slots_) b += (int64_t)s.size;
  return b;
}

hipEvent_t Ctx::take_event() {
  if (!pool.empty()) {
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  PNM_HIP(hipEventCreate(&e));
  return e;
}

bool Ctx::timed(const char* nm) const {
  if (!prof) return false;
  if (prof_only.empty()) return true;
  const size_t len = std::strlen(nm);
  size_t pos = 0;
  while (pos <= prof_only.size()) {
    size_t end = prof_only.find('\n', pos);
    if (end == std::string::npos) end = prof_only.size();
    if (end - pos == len && prof_only.compare(pos, len, nm) == 0) return true;
    pos = end + 1;
  }
  return false;
}

KTimer::KTimer(Ctx* ctx, const char* nm) : c(ctx), name(nm) {
  if (!c->timed(nm)) return;
  a = c->take_event();
  PNM_HIP(hipEventRecord(a, c->stream));
}

KTimer::~KTimer() {
  if (!a) return;
  hipEvent_t b = nullptr;
  try {
    b = c->take_event();
  } catch (...) {
    c->pool.push_back(a);
    return;
  }
  if (hipEventRecord(b, c->stream) != hipSuccess) {
    c->pool.push_back(a);
    c->pool.push_back(b);
    return;
  }
  c->recs.push_back({name, a, b});
}

OnSide::OnSide(Ctx* c) : ctx(c), main(c->stream) {
  if (c->no_fork) return;
  active = true;
  if (!c->side) {
    PNM_HIP(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    PNM_HIP(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    PNM_HIP(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
  }
  PNM_HIP(hipEventRecord(c->ev_fork, main));
  PNM_HIP(hipStreamWaitEvent(c->side, c->ev_fork, 0));
  c->stream = c->side;
  c->side_pending = true;
}

OnSide::~OnSide() {
  if (!active) return;
  (void)hipEventRecord(ctx->ev_join, ctx->side);
  ctx->stream = main;
}

void join_side(Ctx* c) {
  if (!c->side_pending) return;
  c->side_pending = false;
  PNM_HIP(hipStreamWaitEvent(c->stream, c->ev_join, 0));
}

CallScope::CallScope(Ctx* c) : ctx(c), prev_device(-1) {
  if (hipGetDevice(&prev_device) != hipSuccess) prev_device = -1;
  if (prev_device != c->device) PNM_HIP(hipSetDevice(c->device));
  c->arena.reset();
}

CallScope::~CallScope() {
  if (ctx->side_pending) {
    ctx->side_pending = false;
    (void)hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0);
  }
  if (prev_device >= 0 && prev_device != ctx->device) (void)hipSetDevice(prev_device);
}

}  // namespace pnm

using namespace pnm;

namespace {

thread_local std::string g_err;

template <class F>
pnm_status guarded(F&& f) {
  try {
    f();
    return PNM_OK;
  } catch (const Error& e) {
    g_err = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    return PNM_ENOMEM;
  } catch (const std::exception& e) {
    g_err = e.what();
    return PNM_EHIP;
  }
}

void need(bool ok, const char* msg) {
  if (!ok) fail(PNM_EINVAL, msg);
}

Vol check_dims(const int32_t* dims) {
  need(dims != nullptr, "dims must not be NULL");
  need(dims[0] >= 0 && dims[1] >= 0 && dims[2] >= 0, "dims must be non-negative");
  const int64_t n = (int64_t)dims[0] * dims[1] * dims[2];
  need(n < ((int64_t)1 << 31), "volume too large: Nx*Ny*Nz must be < 2^31 (int32 labels)");
  return Vol{dims[0], dims[1], dims[2], n};
}

Ctx* check_ctx(pnm_ctx_t ctx) {
  need(ctx != nullptr, "context is NULL (create one with pnm_ctx_create)");
  return reinterpret_cast<Ctx*>(ctx);
}

template <class T>
T* host_copy(const T* dev, int64_t count, hipStream_t s) {
  T* h = static_cast<T*>(std::malloc(sizeof(T) * (size_t)std::max<int64_t>(count, 1)));
  if (!h) throw std::bad_alloc();
  if (count > 0) {
    const hipError_t e = hipMemcpyAsync(h, dev, sizeof(T) * count, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) {
      PNM_HIP(hipStreamSynchronize(s));
    } else {
      std::free(h);
      PNM_HIP(e);
    }
  }
  return h;
}

void upload(void* dev, const void* host, size_t bytes, hipStream_t s) {
  if (bytes) PNM_HIP(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, s));
}

}  // namespace

extern "C" {

const char* pnm_version(void) { return "0.1.0"; }
const char* pnm_execution_space(void) { return "HIP"; }
const char* pnm_last_error(void) { return g_err.c_str(); }

int pnm_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

pnm_status pnm_ctx_create(pnm_ctx_t* out, int device) {
  return guarded([&] {
    need(out != nullptr, "out must not be NULL");
    *out = nullptr;
    const int nd = pnm_device_count();
    if (nd <= 0) fail(PNM_EHIP, "no HIP device visible (libpnm_hip needs an MI355X GPU)");
    need(device >= 0 && device < nd, "device index out of range");
    Ctx* c = new Ctx();
    c->device = device;
    try {
      int prev = -1;
      (void)hipGetDevice(&prev);
      PNM_HIP(hipSetDevice(device));
      PNM_HIP(hipStreamCreate(&c->own_stream));
      PNM_HIP(hipHostMalloc(reinterpret_cast<void**>(&c->pinned),
                            sizeof(uint64_t) * Ctx::kPinnedWords, hipHostMallocDefault));
      PNM_HIP(hipHostMalloc(reinterpret_cast<void**>(&c->err_host), 256, hipHostMallocMapped));
      *c->err_host = 0u;
      PNM_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&c->err_dev), c->err_host, 0));
      if (prev >= 0 && prev != device) (void)hipSetDevice(prev);
    } catch (...) {
      if (c->err_host) (void)hipHostFree(c->err_host);
      if (c->pinned) (void)hipHostFree(c->pinned);
      if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
      delete c;
      throw;
    }
    c->stream = c->own_stream;
    *out = reinterpret_cast<pnm_ctx_t>(c);
  });
}

void pnm_ctx_destroy(pnm_ctx_t ctx) {
  if (!ctx) return;
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  int prev = -1;
  (void)hipGetDevice(&prev);
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  c->arena.release();
  c->sur.release();
  surrogate_release(c);
  for (auto& r : c->recs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (auto e : c->pool) (void)hipEventDestroy(e);
  if (c->batch_graph) (void)hipGraphExecDestroy(static_cast<hipGraphExec_t>(c->batch_graph));
  if (c->pinned) (void)hipHostFree(c->pinned);
  if (c->err_host) (void)hipHostFree(c->err_host);
  if (c->lb_state) (void)hipFree(c->lb_state);
  if (c->lb_ctl) (void)hipFree(c->lb_ctl);
  if (c->side) {
    (void)hipStreamSynchronize(c->side);
    (void)hipStreamDestroy(c->side);
  }
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  if (prev >= 0) (void)hipSetDevice(prev);
  delete c;
}

pnm_status pnm_ctx_set_stream(pnm_ctx_t ctx, void* stream) {
  return guarded([&] {
    Ctx* c = check_ctx(ctx);
    hipStream_t next = stream ? static_cast<hipStream_t>(stream) : c->own_stream;
    if (next == c->stream) return;
    int prev = -1;
    (void)hipGetDevice(&prev);
    PNM_HIP(hipSetDevice(c->device));
    // The old stream is drained on the host: one event re-recorded on alternating streams and
    // waited on by the other did not order the work on this runtime (measured: the backward
    // pass on the new stream overlapped the forward pass still queued on the old one).
    PNM_HIP(hipStreamSynchronize(c->stream));
    if (prev >= 0 && prev != c->device) (void)hipSetDevice(prev);
    c->stream = next;
  });
}

void* pnm_ctx_stream(pnm_ctx_t ctx) {
  return ctx ? static_cast<void*>(reinterpret_cast<Ctx*>(ctx)->stream) : nullptr;
}

pnm_status pnm_ctx_trim(pnm_ctx_t ctx) {
  return guarded([&] {
    Ctx* c = check_ctx(ctx);
    CallScope cs(c);
    PNM_HIP(hipStreamSynchronize(c->stream));
    c->arena.release();
  });
}

int64_t pnm_ctx_arena_bytes(pnm_ctx_t ctx) {
  return ctx ? reinterpret_cast<Ctx*>(ctx)->arena.bytes() : 0;
}

void pnm_free(void* p) { std::free(p); }

pnm_status pnm_ctx_profile(pnm_ctx_t ctx, int enable) {
  return guarded([&] {
    Ctx* c = check_ctx(ctx);
    c->prof = enable != 0;
  });
}

pnm_status pnm_ctx_profile_only(pnm_ctx_t ctx, const char* names) {
  return guarded([&] {
    Ctx* c = check_ctx(ctx);
    c->prof_only = names ? names : "";
  });
}

pnm_status pnm_ctx_profile_read(pnm_ctx_t ctx, char* names, int64_t names_len, double* total_ms,
                                int64_t* launches, int64_t cap, int64_t* count) {
  return guarded([&] {
    Ctx* c = check_ctx(ctx);
    need(count != nullptr, "count must not be NULL");
    need(cap == 0 || (names && total_ms && launches), "output buffers required when cap > 0");
    int prev = -1;
    (void)hipGetDevice(&prev);
    PNM_HIP(hipSetDevice(c->device));
    std::vector<std::string> nm;
    std::vector<double> ms;
    std::vector<int64_t> cnt;
    for (auto& r : c->recs) {
      PNM_HIP(hipEventSynchronize(r.b));
      float t = 0.0f;
      PNM_HIP(hipEventElapsedTime(&t, r.a, r.b));
      size_t k = 0;
      while (k < nm.size() && nm[k] != r.name) ++k;
      if (k == nm.size()) {
        nm.emplace_back(r.name);
        ms.push_back(0.0);
        cnt.push_back(0);
      }
      ms[k] += t;
      cnt[k] += 1;
      c->pool.push_back(r.a);
      c->pool.push_back(r.b);
    }
    c->recs.clear();
    if (prev >= 0) (void)hipSetDevice(prev);
    *count = (int64_t)nm.size();
    std::string joined;
    for (size_t k = 0; k < nm.size(); ++k) {
      if ((int64_t)k < cap) {
        total_ms[k] = ms[k];
        launches[k] = cnt[k];
      }
      joined += nm[k];
      joined.push_back('\n');
    }
    if (names && names_len > 0) {
      const size_t m = std::min<size_t>(joined.size(), (size_t)names_len - 1);
      std::memcpy(names, joined.data(), m);
      names[m] = 0;
    }
    if ((int64_t)nm.size() > cap) fail(PNM_ECAPACITY, "profile buffers too small");
  });
}

// ------------------------- device-pointer entry points -------------------------

pnm_status pnm_extract_pores_d(pnm_ctx_t ctx, const float* sdf_d, const int32_t dims[3],
                               const float origin[3], const float spacing[3],
                               pnm_pore_t* pores_d, int64_t cap, int64_t* count,
                               int64_t* total) {
  return guarded([&] {
    const Vol v = check_dims(dims);
    need(origin && spacing && count, "origin, spacing and count must not be NULL");
    need(v.n == 0 || sdf_d, "sdf must not be NULL");
    need(cap == 0 || pores_d, "pores buffer must not be NULL when cap > 0");
    Ctx* c = check_ctx(ctx);
    *count = 0;
    if (total) *total = 0;
    if (v.n == 0) return;
    CallScope cs(c);
    int64_t tot = 0;
    const int64_t k = pores_device(c, sdf_d, v, origin, spacing, pores_d, cap, &tot);
    PNM_HIP(hipStreamSynchronize(c->stream));
    if (total) *total = tot;
    if (k < 0) {
      *count = -k;
      fail(PNM_ECAPACITY, "pores buffer too small: need " + std::to_string(-k) + " records");
    }
    *count = k;
  });
}

pnm_status pnm_segment_volume_d(pnm_ctx_t ctx, const float* sdf_d, const int32_t dims[3],
                                const float spacing[3], int32_t* seg_d, int32_t* roots_d) {
  return guarded([&] {
    const Vol v = check_dims(dims);
    need(spacing != nullptr, "spacing must not be NULL");
    need(v.n == 0 || (sdf_d && seg_d), "sdf and seg must not be NULL");
    Ctx* c = check_ctx(ctx);
    if (v.n == 0) return;
    CallScope cs(c);
    segment_device(c, sdf_d, v, spacing, seg_d, roots_d);
    PNM_HIP(hipStreamSynchronize(c->stream));
  });
}

pnm_status pnm_extract_topology_d(pnm_ctx_t ctx, const int32_t* seg_d, const int32_t dims[3],
                                  int32_t* pairs_d, int64_t cap, int64_t* count) {
  return guarded([&] {
    const Vol v = check_dims(dims);
    need(count != nullptr, "count must not be NULL");
    need(v.n == 0 || seg_d, "seg must not be NULL");
    need(cap == 0 || pairs_d, "pairs buffer must not be NULL when cap > 0");
    Ctx* c = check_ctx(ctx);
    *count = 0;
    if (v.n == 0) return;
    CallScope cs(c);
    int64_t needed = 0;
    const int64_t k = topology(c, seg_d, v, 0, 0, false, pairs_d, cap, &needed);
    PNM_HIP(hipStreamSynchronize(c->stream));
    *count = needed;
    if (k < 0)
      fail(PNM_ECAPACITY, "pairs buffer too small: need " + std::to_string(needed) + " pairs");
  });
}

pnm_status pnm_extract_pore_network_d(pnm_ctx_t ctx, const float* sdf_d, const int32_t dims[3],
                                      const float origin[3], const float spacing[3],
                                      pnm_pore_t* pores_d, int64_t cap_pores, int64_t* npores,
                                      int64_t* total_pores, int32_t* seg_d, int32_t* pairs_d,
                                      int64_t cap_pairs, int64_t* npairs) {
  return guarded([&] {
    const Vol v = check_dims(dims);
    need(origin && spacing && npores && npairs, "origin, spacing, npores, npairs required");
    need(v.n == 0 || (sdf_d && seg_d), "sdf and seg must not be NULL");
    need(cap_pores == 0 || pores_d, "pores buffer must not be NULL when cap_pores > 0");
    need(cap_pairs == 0 || pairs_d, "pairs buffer must not be NULL when cap_pairs > 0");
    Ctx* c = check_ctx(ctx);
    *npores = 0;
    *npairs = 0;
    if (total_pores) *total_pores = 0;
    if (v.n == 0) return;
    CallScope cs(c);
    int64_t tot = 0, np2 = 0;
    bool povf = false;
    const int64_t k = network_device(c, sdf_d, v, origin, spacing, pores_d, cap_pores, &tot,
                                     seg_d, pairs_d, cap_pairs, &np2, &povf);
    PNM_HIP(hipStreamSynchronize(c->stream));
    if (total_pores) *total_pores = tot;
    *npores = k < 0 ? -k : k;
    *npairs = np2;
    if (k < 0) fail(PNM_ECAPACITY, "pores buffer too small: need " + std::to_string(-k));
    if (povf) fail(PNM_ECAPACITY, "pairs buffer too small: need " + std::to_string(np2));
  });
}

pnm_status pnm_segment_batch_d(pnm_ctx_t ctx, int32_t nvol, const float* const* sdf,
                               int32_t* const* seg, const int32_t* dims, const float spacing[3],
                               int32_t* n_pores_d, int32_t* n_solids_d, int32_t flags) {
  return guarded([&] {
    need(nvol >= 0, "nvol must be non-negative");
    need(spacing != nullptr, "spacing must not be NULL");
    need(nvol == 0 || (sdf && seg && dims), "sdf, seg and dims arrays are required");
    const float mn = std::min(spacing[0], std::min(spacing[1], spacing[2]));
    need(mn > 0.0f, "batched segmentation needs min(spacing) > 0");
    need((flags & ~(PNM_BATCH_GRAPH | PNM_BATCH_AUTOTUNE)) == 0, "unknown flags");
    for (int k = 0; k < nvol; ++k) {
      const Vol v = check_dims(dims + 3 * k);
      need(v.n > 0, "every volume of a batch must be non-empty");
      need(sdf[k] && seg[k], "NULL volume pointer in the batch");
    }
    Ctx* c = check_ctx(ctx);
    if (nvol == 0) return;
    CallScope cs(c);
    segment_batch_device(c, nvol, sdf, seg, dims, spacing, flags, n_pores_d, n_solids_d);
  });
}

pnm_status pnm_batch_config(pnm_ctx_t ctx, int32_t* block, int32_t* per_cu, int32_t* groups,
                            float* tuned_ms) {
  return guarded([&] {
    Ctx* c = check_ctx(ctx);
    if (block) *block = c->batch_block;
    if (per_cu) *per_cu = c->batch_per_cu;
    if (groups) *groups = c->batch_groups;
    if (tuned_ms) *tuned_ms = c->batch_tuned_ms;
  });
}

pnm_status pnm_stage_labels_d(pnm_ctx_t ctx, const float* sdf_d, const int32_t dims[3],
                              const float spacing[3], int32_t* labels_d, int32_t* ccl_only_d) {
  return guarded([&] {
    const Vol v = check_dims(dims);
    need(spacing != nullptr, "spacing must not be NULL");
    need(v.n == 0 || (sdf_d && labels_d), "sdf and labels must not be NULL");
    Ctx* c = check_ctx(ctx);
    if (v.n == 0) return;
    CallScope cs(c);
    labels_device(c, sdf_d, v, spacing, labels_d, ccl_only_d);
    PNM_HIP(hipStreamSynchronize(c->stream));
  });
}

pnm_status pnm_stage_roots_d(pnm_ctx_t ctx, const float* sdf_d, const int32_t dims[3],
                             int all_voxels, int32_t* roots_d) {
  return guarded([&] {
    const Vol v = check_dims(dims);
    need(v.n == 0 || (sdf_d && roots_d), "sdf and roots must not be NULL");
    Ctx* c = check_ctx(ctx);
    if (v.n == 0) return;
    CallScope cs(c);
    int32_t* next = c->arena.get_n<int32_t>((size_t)v.n);
    const int mode = all_voxels ? kWalkAll : kWalkPores;
    launch_next(c, sdf_d, v, next, mode);
    launch_chase(c, sdf_d, v, next, roots_d, mode);
    PNM_HIP(hipStreamSynchronize(c->stream));
  });
}

pnm_status pnm_extract_network_flow_d(pnm_ctx_t ctx, const float* sdf_d, const int32_t dims[3],
                                      const float origin[3], const float spacing[3],
                                      const double* u_d, const double* v_d, const double* w_d,
                                      const double* p_d, const double* ox_d, const double* oy_d,
                                      const double* oz_d, const double grad_p[3],
                                      pnm_flow_result_t* out) {
  return guarded([&] {
    const Vol v = check_dims(dims);
    need(origin && spacing && grad_p && out, "origin, spacing, grad_p and out are required");
    need(v.n == 0 || (sdf_d && u_d && v_d && w_d && p_d), "sdf, u, v, w, p must not be NULL");
    need((ox_d == nullptr) == (oy_d == nullptr) && (ox_d == nullptr) == (oz_d == nullptr),
         "pass all three openness arrays (ox,oy,oz) or none");
    Ctx* c = check_ctx(ctx);
    std::memset(out, 0, sizeof(*out));
    if (v.n == 0) return;
    CallScope cs(c);
    flow_device(c, sdf_d, v, origin, spacing, u_d, v_d, w_d, p_d, ox_d, oy_d, oz_d, grad_p, out);
  });
}

pnm_status pnm_throat_geometry_d(pnm_ctx_t ctx, const float* sdf_d, const int32_t* seg_d,
                                 const int32_t dims[3], const float origin[3],
                                 const float spacing[3], float* centers_d, float* raw_d,
                                 int64_t cap, pnm_flow_result_t* out) {
  return guarded([&] {
    const Vol v = check_dims(dims);
    need(origin && spacing && out, "origin, spacing and out are required");
    need(v.n == 0 || sdf_d, "sdf must not be NULL");
    need((centers_d == nullptr) == (raw_d == nullptr) && cap >= 0,
         "pass both feature buffers (centers, raw) or neither, with cap >= 0");
    Ctx* c = check_ctx(ctx);
    std::memset(out, 0, sizeof(*out));
    if (v.n == 0) return;
    CallScope cs(c);
    GeoMode geo;
    geo.seg = seg_d;
    geo.centers = centers_d;
    geo.raw = raw_d;
    geo.cap = cap;
    const double zero[3] = {0.0, 0.0, 0.0};
    flow_device(c, sdf_d, v, origin, spacing, nullptr, nullptr, nullptr, nullptr, nullptr,
                nullptr, nullptr, zero, out, &geo);
  });
}

pnm_status pnm_stokes_solve_d(pnm_ctx_t ctx, const float* sdf_d, const int32_t dims[3],
                              const float spacing[3], const double force[3], double mu,
                              double rtol, int32_t max_iter, double* u_d, double* v_d,
                              double* w_d, double* p_d, double* ox_d, double* oy_d, double* oz_d,
                              int32_t* iterations, double* rel_residual) {
  return guarded([&] {
    const Vol v = check_dims(dims);
    need(spacing && force, "spacing and force are required");
    need(v.n == 0 || (sdf_d && u_d && v_d && w_d && p_d), "sdf, u, v, w, p must not be NULL");
    need(spacing[0] > 0.0f && spacing[1] > 0.0f && spacing[2] > 0.0f,
         "spacing must be positive");
    need(mu > 0.0 && rtol > 0.0 && max_iter >= 0, "mu and rtol must be positive, max_iter >= 0");
    Ctx* c = check_ctx(ctx);
    if (iterations) *iterations = 0;
    if (rel_residual) *rel_residual = 0.0;
    if (v.n == 0) return;
    CallScope cs(c);
    const StokesInfo si = stokes_device(c, sdf_d, v, spacing, force, mu, rtol, max_iter, u_d,
                                        v_d, w_d, p_d, ox_d, oy_d, oz_d);
    if (iterations) *iterations = si.iterations;
    if (rel_residual) *rel_residual = si.rel_residual;
  });
}

// ------------------------- throat-conductance surrogate -------------------------

int64_t pnm_sur_param_count(void) { return surrogate_param_count(); }

pnm_status pnm_sur_init(uint64_t seed, float* params, int64_t count) {
  return guarded([&] {
    need(params != nullptr, "params must not be NULL");
    need(count == surrogate_param_count(), "count must equal pnm_sur_param_count()");
    surrogate_init(seed, params);
  });
}

pnm_status pnm_sur_forward_d(pnm_ctx_t ctx, int32_t batch, const float* patches_d,
                             const float* scalars_d, const float* params_d, float* out_d) {
  return guarded([&] {
    need(batch > 0 && patches_d && scalars_d && params_d && out_d,
         "batch > 0 and all buffers are required");
    Ctx* c = check_ctx(ctx);
    CallScope cs(c);
    surrogate_forward(c, batch, patches_d, scalars_d, params_d, out_d);
  });
}

pnm_status pnm_sur_backward_d(pnm_ctx_t ctx, int32_t batch, const float* dout_d,
                              const float* params_d, float* grads_d) {
  return guarded([&] {
    need(batch > 0 && dout_d && params_d && grads_d, "batch > 0 and all buffers are required");
    Ctx* c = check_ctx(ctx);
    CallScope cs(c);
    surrogate_backward(c, batch, dout_d, params_d, grads_d);
  });
}

pnm_status pnm_sur_loss_d(pnm_ctx_t ctx, int32_t batch, const float* pred_d,
                          const float* target_d, float delta, float* dpred_d, float* loss_d) {
  return guarded([&] {
    need(batch > 0 && pred_d && target_d && dpred_d, "batch > 0 and buffers are required");
    Ctx* c = check_ctx(ctx);
    CallScope cs(c);
    surrogate_loss(c, batch, pred_d, target_d, delta, dpred_d, loss_d);
  });
}

pnm_status pnm_sur_adam_d(pnm_ctx_t ctx, float* params_d, const float* grads_d, float* m_d,
                          float* v_d, int32_t step, float lr, float beta1, float beta2,
                          float eps) {
  return guarded([&] {
    need(params_d && grads_d && m_d && v_d && step >= 1, "buffers required, step >= 1");
    Ctx* c = check_ctx(ctx);
    CallScope cs(c);
    surrogate_adam(c, params_d, grads_d, m_d, v_d, step, lr, beta1, beta2, eps);
  });
}

pnm_status pnm_sur_infer_d(pnm_ctx_t ctx, const float* sdf_d, const int32_t dims[3],
                           const float* centers_d, const float* scalars_d, int32_t batch,
                           float scale, const float* params_d, float* out_d) {
  return guarded([&] {
    const Vol v = check_dims(dims);
    need(v.n > 0 && sdf_d && centers_d && scalars_d && params_d && out_d && batch >= 0 &&
             scale > 0.0f,
         "non-empty volume, buffers and scale > 0 are required");
    Ctx* c = check_ctx(ctx);
    if (batch == 0) return;
    CallScope cs(c);
    surrogate_infer(c, sdf_d, v, centers_d, scalars_d, batch, scale, params_d, out_d);
  });
}

pnm_status pnm_sur_patches_d(pnm_ctx_t ctx, const float* sdf_d, const int32_t dims[3],
                             const float* centers_d, int32_t batch, float scale, float* out_d) {
  return guarded([&] {
    const Vol v = check_dims(dims);
    need(v.n > 0 && sdf_d && centers_d && out_d && batch >= 0 && scale > 0.0f,
         "non-empty volume, buffers and scale > 0 are required");
    Ctx* c = check_ctx(ctx);
    if (batch == 0) return;
    CallScope cs(c);
    surrogate_patches(c, sdf_d, v, centers_d, batch, scale, out_d);
  });
}

// ------------------------- host-pointer entry points -------------------------

pnm_status pnm_extract_pores(pnm_ctx_t ctx, const float* sdf, const int32_t dims[3],
                             const float origin[3], const float spacing[3], pnm_pore_t** pores,
                             int64_t* count, int64_t* total) {
  return guarded([&] {
    const Vol v = check_dims(dims);
    need(origin && spacing && pores && count, "origin, spacing, pores and count required");
    need(v.n == 0 || sdf, "sdf must not be NULL");
    Ctx* c = check_ctx(ctx);
    *pores = nullptr;
    *count = 0;
    if (total) *total = 0;
    if (v.n == 0) {
      *pores = static_cast<pnm_pore_t*>(std::malloc(sizeof(pnm_pore_t)));
      return;
    }
    CallScope cs(c);
    float* sdf_d = c->arena.get_n<float>((size_t)v.n);
    pnm_pore_t* pd = c->arena.get_n<pnm_pore_t>(PNM_MAX_PORES);
    upload(sdf_d, sdf, sizeof(float) * v.n, c->stream);
    int64_t tot = 0;
    const int64_t k = pores_device(c, sdf_d, v, origin, spacing, pd, PNM_MAX_PORES, &tot);
    *pores = host_copy(pd, k, c->stream);
    *count = k;
    if (total) *total = tot;
  });
}

pnm_status pnm_segment_volume(pnm_ctx_t ctx, const float* sdf, const int32_t dims[3],
                              const float spacing[3], int32_t* seg_out) {
  return guarded([&] {
    const Vol v = check_dims(dims);
    need(spacing != nullptr, "spacing must not be NULL");
    need(v.n == 0 || (sdf && seg_out), "sdf and seg must not be NULL");
    Ctx* c = check_ctx(ctx);
    if (v.n == 0) return;
    CallScope cs(c);
    float* sdf_d = c->arena.get_n<float>((size_t)v.n);
    int32_t* seg_d = c->arena.get_n<int32_t>((size_t)v.n);
    upload(sdf_d, sdf, sizeof(float) * v.n, c->stream);
    segment_device(c, sdf_d, v, spacing, seg_d, nullptr);
    PNM_HIP(hipMemcpyAsync(seg_out, seg_d, sizeof(int32_t) * v.n, hipMemcpyDeviceToHost,
                           c->stream));
    PNM_HIP(hipStreamSynchronize(c->stream));
  });
}

pnm_status pnm_extract_topology(pnm_ctx_t ctx, const int32_t* seg, const int32_t dims[3],
                                int32_t** pairs, int64_t* count) {
  return guarded([&] {
    const Vol v = check_dims(dims);
    need(pairs && count, "pairs and count required");
    need(v.n == 0 || seg, "seg must not be NULL");
    Ctx* c = check_ctx(ctx);
    *pairs = nullptr;
    *count = 0;
    if (v.n == 0) {
      *pairs = static_cast<int32_t*>(std::malloc(sizeof(int32_t) * 2));
      return;
    }
    CallScope cs(c);
    int32_t* seg_d = c->arena.get_n<int32_t>((size_t)v.n);
    int32_t* pd = c->arena.get_n<int32_t>((size_t)(6 * v.n));
    upload(seg_d, seg, sizeof(int32_t) * v.n, c->stream);
    int64_t needed = 0;
    const int64_t k = topology(c, seg_d, v, 0, 0, false, pd, 3 * v.n, &needed);
    *pairs = host_copy(pd, 2 * k, c->stream);
    *count = k;
  });
}

pnm_status pnm_extract_pore_network(pnm_ctx_t ctx, const float* sdf, const int32_t dims[3],
                                    const float origin[3], const float spacing[3],
                                    pnm_pore_t** pores, int64_t* npores, int64_t* total_pores,
                                    int32_t* seg_out, int32_t** pairs, int64_t* npairs) {
  return guarded([&] {
    const Vol v = check_dims(dims);
    need(origin && spacing && pores && npores && pairs && npairs, "output pointers required");
    need(v.n == 0 || (sdf && seg_out), "sdf and seg must not be NULL");
    Ctx* c = check_ctx(ctx);
    *pores = nullptr;
    *pairs = nullptr;
    *npores = 0;
    *npairs = 0;
    if (total_pores) *total_pores = 0;
    if (v.n == 0) {
      *pores = static_cast<pnm_pore_t*>(std::malloc(sizeof(pnm_pore_t)));
      *pairs = static_cast<int32_t*>(std::malloc(sizeof(int32_t) * 2));
      return;
    }
    CallScope cs(c);
    float* sdf_d = c->arena.get_n<float>((size_t)v.n);
    int32_t* seg_d = c->arena.get_n<int32_t>((size_t)v.n);
    pnm_pore_t* pd = c->arena.get_n<pnm_pore_t>(PNM_MAX_PORES);
    int32_t* prd = c->arena.get_n<int32_t>((size_t)(6 * v.n));
    upload(sdf_d, sdf, sizeof(float) * v.n, c->stream);
    int64_t tot = 0, np2 = 0;
    bool povf = false;
    const int64_t k = network_device(c, sdf_d, v, origin, spacing, pd, PNM_MAX_PORES, &tot,
                                     seg_d, prd, 3 * v.n, &np2, &povf);
    PNM_HIP(hipMemcpyAsync(seg_out, seg_d, sizeof(int32_t) * v.n, hipMemcpyDeviceToHost,
                           c->stream));
    PNM_HIP(hipStreamSynchronize(c->stream));
    *pores = host_copy(pd, k, c->stream);
    *npores = k;
    if (total_pores) *total_pores = tot;
    *pairs = host_copy(prd, 2 * np2, c->stream);
    *npairs = np2;
  });
}

pnm_status pnm_extract_network_flow(pnm_ctx_t ctx, const float* sdf, const int32_t dims[3],
                                    const float origin[3], const float spacing[3],
                                    const double* u, const double* v, const double* w,
                                    const double* p, const double* ox, const double* oy,
                                    const double* oz, const double grad_p[3],
                                    pnm_flow_result_t* out) {
  return guarded([&] {
    const Vol vol = check_dims(dims);
    need(origin && spacing && grad_p && out, "origin, spacing, grad_p and out are required");
    need(vol.n == 0 || (sdf && u && v && w && p), "sdf, u, v, w, p must not be NULL");
    need((ox == nullptr) == (oy == nullptr) && (ox == nullptr) == (oz == nullptr),
         "pass all three openness arrays (ox,oy,oz) or none");
    Ctx* c = check_ctx(ctx);
    std::memset(out, 0, sizeof(*out));
    if (vol.n == 0) return;
    CallScope cs(c);
    const size_t nd = (size_t)vol.n;
    float* sdf_d = c->arena.get_n<float>(nd);
    double* f[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    const double* hf[7] = {u, v, w, p, ox, oy, oz};
    upload(sdf_d, sdf, sizeof(float) * nd, c->stream);
    for (int k = 0; k < 7; ++k)
      if (hf[k]) {
        f[k] = c->arena.get_n<double>(nd);
        upload(f[k], hf[k], sizeof(double) * nd, c->stream);
      }
    flow_device(c, sdf_d, vol, origin, spacing, f[0], f[1], f[2], f[3], f[4], f[5], f[6], grad_p,
                out);
  });
}

void pnm_flow_result_free(pnm_flow_result_t* r) {
  if (!r) return;
  std::free(r->pores);
  std::free(r->pore_pressure);
  std::free(r->pore_residual);
  std::free(r->throats);
  std::free(r->throat_flow);
  std::free(r->throat_area);
  std::free(r->throat_dp);
  std::free(r->throat_center);
  std::memset(r, 0, sizeof(*r));
}

}  // extern "C"

// ------------------------- segmentation without allocation -------------------------
//
// pnm_segment_volume_noalloc_d runs the segment_volume pipeline on a fixed device region
// instead of the context's arena, for callers that cannot call hipMalloc. The region is the
// file-scope array `scratch` below: one copy per GPU, held until the process ends. Its device
// address is taken once per device (hipGetSymbolAddress, with that device current) on the first
// call there and kept in g_scratch_ptr; host code uses only that address, never the array's name
// (which is a host-side stand-in). Layout of the region:
//   [0, 256)                            look-back control words (ticket, done, epoch)
//   [256, 256 + 8 * kNoAllocLbWords)    look-back status words (epoch-tagged, zeroed at set-up)
//   [kNoAllocHeader, end)               bump arena for every other scratch array of the call
// Every call on a device shares that device's region and leaves its contents behind; the
// pipeline clears what it needs zeroed (counters, bitmaps) itself. Calls that overlap on one
// device, from any thread or context, race on the region.
namespace {

constexpr int64_t kNoAllocLbWords = 16384;
constexpr size_t kNoAllocHeader = (size_t)1 << 20;
// Bound on the pipeline's scratch for a tile-padded volume of P voxels (n <= P): 56 B per voxel
// of working arrays plus 25 B per padded voxel of tile staging, rounded up with headroom.
constexpr size_t kNoAllocArena = (size_t)82 * PNM_NOALLOC_MAX_PADDED + ((size_t)2 << 20);
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocHeader + kNoAllocArena];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kNoAllocMaxDevices] = {};

// Device address of this device's copy of `scratch`; the first call on a device takes it and
// initialises the look-back words. The caller has made c->device current.
char* noalloc_scratch(Ctx* c) {
  need(c->device >= 0 && c->device < kNoAllocMaxDevices, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  void*& p = g_scratch_ptr[c->device];
  if (!p) {
    void* q = nullptr;
    PNM_HIP(hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)));
    uint32_t* ctl = static_cast<uint32_t*>(q);
    fill_u32(c, ctl, 0u, 64);     // ticket, done
    fill_u32(c, ctl + 2, 1u, 1);  // epoch 1 (status words of epoch 0 read as "not ready")
    fill_u64(c, reinterpret_cast<uint64_t*>(static_cast<char*>(q) + 256), 0u, kNoAllocLbWords);
    PNM_HIP(hipStreamSynchronize(c->stream));
    p = q;
  }
  return static_cast<char*>(p);
}

// For the life of one call: the arena carves from the region, the look-back scans use the
// region's words, no second stream is forked and no timing events are taken.
struct FixedScratch {
  FixedScratch(Ctx* c, char* base)
      : ctx(c), ctl(c->lb_ctl), state(c->lb_state), cap(c->lb_cap), prof(c->prof) {
    c->lb_ctl = reinterpret_cast<uint32_t*>(base);
    c->lb_state = reinterpret_cast<uint64_t*>(base + 256);
    c->lb_cap = kNoAllocLbWords;
    c->lb_fixed = true;
    c->no_fork = true;
    c->prof = false;
    c->arena.fix(base + kNoAllocHeader, kNoAllocArena);
  }
  ~FixedScratch() {
    ctx->arena.unfix();
    ctx->lb_ctl = ctl;
    ctx->lb_state = state;
    ctx->lb_cap = cap;
    ctx->lb_fixed = false;
    ctx->no_fork = false;
    ctx->prof = prof;
  }
  Ctx* ctx;
  uint32_t* ctl;
  uint64_t* state;
  int64_t cap;
  bool prof;
};

}  // namespace

extern "C" pnm_status pnm_segment_volume_noalloc_d(pnm_ctx_t ctx, const float* sdf_d,
                                                   const int32_t dims[3], const float spacing[3],
                                                   int32_t* seg_d, int32_t* roots_d) {
  return guarded([&] {
    const Vol v = check_dims(dims);
    need(spacing != nullptr, "spacing must not be NULL");
    need(v.n == 0 || (sdf_d && seg_d), "sdf and seg must not be NULL");
    Ctx* c = check_ctx(ctx);
    if (v.n == 0) return;
    const int64_t padded = PNM_NOALLOC_PADDED_VOXELS(dims[0], dims[1], dims[2]);
    if (padded > PNM_NOALLOC_MAX_PADDED)
      fail(PNM_ECAPACITY, "volume too large for the no-allocation entry point: " +
                              std::to_string(padded) + " tile-padded voxels (Nx, Ny, Nz rounded " +
                              "up to 64, 8, 16) exceed the maximum of " +
                              std::to_string((int64_t)PNM_NOALLOC_MAX_PADDED));
    CallScope cs(c);
    char* base = noalloc_scratch(c);
    FixedScratch fx(c, base);
    segment_device(c, sdf_d, v, spacing, seg_d, roots_d);
    PNM_HIP(hipStreamSynchronize(c->stream));
  });
}
