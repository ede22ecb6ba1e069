/*
 * pnm_capi.h — C-ABI of the MI355X (gfx950) pore-network extraction engine, libpnm_hip.so.
 *
 * Plain pointers and sizes only (no C++ or torch types). Every entry point replaces one host
 * wrapper of the reference's src/pore_extraction.hpp (cited per function); the Python module
 * `peclet.pnm` (built over this header) mirrors src/pnm_bindings.cpp.
 *
 * Conventions (as the reference's core, pore_extraction.hpp:59-64):
 *   - dims/origin/spacing are X,Y,Z. The Python layer reorders the zyx arguments and casts
 *     origin/spacing to float exactly as pnm_bindings.cpp:116-118 does.
 *   - volumes are flat, x fastest: i = z*Ny*Nx + y*Nx + x (numpy C-order of (Nz,Ny,Nx));
 *     every stencil wraps periodically.
 *   - n = Nx*Ny*Nz must be < 2^31 (labels are int32 voxel indices, as in the reference).
 *
 * Two families of entry points:
 *   - host-pointer calls (`pnm_*`): inputs in host memory, results copied back to host, as the
 *     reference's `*_k` wrappers do;
 *   - device-pointer calls (`pnm_*_d`): inputs and the per-voxel outputs stay in HBM (pointers
 *     from hipMalloc/torch); only small lists come back. They run on the context's stream.
 *
 * Errors: every call returns a pnm_status; pnm_last_error() gives the message of the most recent
 * failure on the calling thread. Arguments are validated before any device work, so a bad call
 * fails the same way with or without a GPU.
 *
 * Threading: a context owns one HIP stream and one scratch arena; it is not re-entrant. Separate
 * contexts may be used from separate threads.
 */
#ifndef PNM_CAPI_H
#define PNM_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Pore record, layout-identical to `struct pnm::Pore {float x,y,z,radius;}`
 * (pore_extraction.hpp:24-26). */
typedef struct {
  float x, y, z, radius;
} pnm_pore_t;

typedef enum {
  PNM_OK = 0,
  PNM_EINVAL = 1,      /* bad argument (shape, length, null pointer, n >= 2^31) */
  PNM_ENOMEM = 2,      /* device or host allocation failed */
  PNM_EHIP = 3,        /* HIP runtime error, or no GPU */
  PNM_ECAPACITY = 4,   /* a caller-provided output buffer is too small (needed size reported) */
  PNM_ENOCONVERGE = 5  /* a fixpoint loop exceeded its iteration guard */
} pnm_status;

typedef struct pnm_ctx_s* pnm_ctx_t;

/* Maximum pores returned by pnm_extract_pores / pnm_extract_pore_network
 * (max_pores, pore_extraction.hpp:77). The reference truncates silently to an arbitrary subset;
 * this library keeps the lowest-index peaks and reports the untruncated count. */
#define PNM_MAX_PORES 1000000

/* ---- library / context ---- */
const char* pnm_version(void);
/* "HIP" — the value of the module attribute `execution_space` (pnm_bindings.cpp:84). */
const char* pnm_execution_space(void);
/* Message of the last failed call on this thread ("" if none). */
const char* pnm_last_error(void);
/* Number of visible HIP devices (0 when there is no GPU); never fails. */
int pnm_device_count(void);

/* Create a context on HIP device `device` (its own stream + arena). Replaces Kokkos::initialize
 * (pnm_bindings.cpp:75-76). */
pnm_status pnm_ctx_create(pnm_ctx_t* out, int device);
void pnm_ctx_destroy(pnm_ctx_t ctx);
/* Run this context's work on `stream` (a hipStream_t, e.g. torch's current stream); NULL
 * restores the context's own stream. The call first waits on the host for the work already
 * queued on the old stream, so asynchronous calls made before the switch are done with the
 * context's scratch before calls after it reuse that scratch. */
pnm_status pnm_ctx_set_stream(pnm_ctx_t ctx, void* stream);
void* pnm_ctx_stream(pnm_ctx_t ctx);
/* Release the context's scratch arena (it regrows on the next call). */
pnm_status pnm_ctx_trim(pnm_ctx_t ctx);
/* Bytes currently held by the context's scratch arena. */
int64_t pnm_ctx_arena_bytes(pnm_ctx_t ctx);

/* Kernel timing (no reference counterpart; the reference labels kernels for Kokkos Tools).
 * enable != 0: every kernel launch of this context is bracketed by two HIP events on the launch
 * stream. pnm_ctx_profile_read aggregates the recorded launches per kernel name since the last
 * read: names (newline-separated into `names`), total milliseconds and launch counts.
 * *count = number of distinct kernels (PNM_ECAPACITY when it exceeds cap). */
pnm_status pnm_ctx_profile(pnm_ctx_t ctx, int enable);
/* Restrict kernel timing to the launches whose kernel names are listed in `names`
 * (newline-separated); NULL or "" times every launch again. Launches that are not timed carry
 * no events, so the stream runs as it does with timing off. */
pnm_status pnm_ctx_profile_only(pnm_ctx_t ctx, const char* names);
pnm_status pnm_ctx_profile_read(pnm_ctx_t ctx, char* names, int64_t names_len, double* total_ms,
                                int64_t* launches, int64_t cap, int64_t* count);

/* Free a list returned by this library. */
void pnm_free(void* p);

/* ---- host-pointer entry points ---- */

/* extract_pores_k (pore_extraction.hpp:143-151) -> extractPoresView (:69-140).
 * Pores in ascending peak-voxel index; *pores is malloc'ed (pnm_free), *count <= PNM_MAX_PORES,
 * *total = untruncated number of peaks (may be NULL). n == 0 gives an empty list. */
pnm_status pnm_extract_pores(pnm_ctx_t ctx, const float* sdf, const int32_t dims[3],
                             const float origin[3], const float spacing[3], pnm_pore_t** pores,
                             int64_t* count, int64_t* total);

/* segment_volume_k (pore_extraction.hpp:389-397) -> segmentVolumeView (:157-386).
 * seg_out: caller-allocated int32[n]. */
pnm_status pnm_segment_volume(pnm_ctx_t ctx, const float* sdf, const int32_t dims[3],
                              const float spacing[3], int32_t* seg_out);

/* extract_topology_k (pore_extraction.hpp:454-461) -> extractTopologyView (:401-451).
 * *pairs: malloc'ed int32[2*count] of (min,max) pairs in signed lexicographic order. */
pnm_status pnm_extract_topology(pnm_ctx_t ctx, const int32_t* seg, const int32_t dims[3],
                                int32_t** pairs, int64_t* count);

/* extract_pore_network_k (pore_extraction.hpp:1011-1025): pores, seg, connections from one
 * upload of the SDF. seg_out: caller int32[n]; *pores and *pairs malloc'ed. */
pnm_status pnm_extract_pore_network(pnm_ctx_t ctx, const float* sdf, const int32_t dims[3],
                                    const float origin[3], const float spacing[3],
                                    pnm_pore_t** pores, int64_t* npores, int64_t* total_pores,
                                    int32_t* seg_out, int32_t** pairs, int64_t* npairs);

/* Result of pnm_extract_network_flow: struct NetworkFlow (pore_extraction.hpp:482-491).
 * All arrays are malloc'ed by the library; release with pnm_flow_result_free. */
typedef struct {
  int64_t np;             /* pore labels 1..np */
  int64_t nt;             /* throats (per connected face patch) */
  pnm_pore_t* pores;      /* [np]  pores[k] <-> label k+1 */
  double* pore_pressure;  /* [np]  trilinear p at the pore centre */
  double* pore_residual;  /* [np]  signed net outflow over the pore's boundary */
  int32_t* throats;       /* [2*nt] (li, lj), li < lj, sorted by (pair, patch root face) */
  double* throat_flow;    /* [nt]  positive li -> lj */
  double* throat_area;    /* [nt]  sum of openness * face area */
  double* throat_dp;      /* [nt]  P_i - P_j including the macroscopic gradient */
  double* throat_center;  /* [3*nt] X,Y,Z world position of the interface centroid: the lower
                             pore's peak-voxel centre plus the min-image centroid offset C/A
                             (the anchor throat_dp uses, :962-991); not part of the reference's
                             NetworkFlow, returned for throat-level models */
} pnm_flow_result_t;

/* extract_network_flow_k (pore_extraction.hpp:493-997). u,v,w,p: float64[n] MAC fields
 * (u(i,j,k) on the -x face of cell (i,j,k)); ox,oy,oz: all three or all NULL (fully open).
 * grad_p is X,Y,Z. */
pnm_status pnm_extract_network_flow(pnm_ctx_t ctx, const float* sdf, const int32_t dims[3],
                                    const float origin[3], const float spacing[3],
                                    const double* u, const double* v, const double* w,
                                    const double* p, const double* ox, const double* oy,
                                    const double* oz, const double grad_p[3],
                                    pnm_flow_result_t* out);
void pnm_flow_result_free(pnm_flow_result_t* r);

/* ---- device-pointer entry points (inputs and per-voxel outputs resident in HBM) ---- */

/* extractPoresView (pore_extraction.hpp:69-140) on a device SDF. pores_d: device buffer of
 * `cap` records; *count = number written (<= min(cap, PNM_MAX_PORES)), *total = peaks found. */
pnm_status pnm_extract_pores_d(pnm_ctx_t ctx, const float* sdf_d, const int32_t dims[3],
                               const float origin[3], const float spacing[3],
                               pnm_pore_t* pores_d, int64_t cap, int64_t* count,
                               int64_t* total);

/* segmentVolumeView (pore_extraction.hpp:157-386) on a device SDF; seg_d: device int32[n].
 * roots_d (nullable): device int32[n] receiving rootsOut (:383-384): the gradient-path root of
 * every voxel, the ascent walker for sdf>0 voxels and the descent walker for the others
 * (:273-311). */
pnm_status pnm_segment_volume_d(pnm_ctx_t ctx, const float* sdf_d, const int32_t dims[3],
                                const float spacing[3], int32_t* seg_d, int32_t* roots_d);

/* extractTopologyView (pore_extraction.hpp:401-451) on a device segmentation.
 * pairs_d: device int32[2*cap]; *count = unique pairs. PNM_ECAPACITY if count > cap (nothing
 * written, *count holds the needed size). */
pnm_status pnm_extract_topology_d(pnm_ctx_t ctx, const int32_t* seg_d, const int32_t dims[3],
                                  int32_t* pairs_d, int64_t cap, int64_t* count);

/* The fused pipeline (pore_extraction.hpp:1011-1025) fully device-resident: pores_d
 * (cap_pores records), seg_d (int32[n]) and pairs_d (int32[2*cap_pairs]) are device buffers.
 * Per-buffer PNM_ECAPACITY as above. */
pnm_status pnm_extract_pore_network_d(pnm_ctx_t ctx, const float* sdf_d, const int32_t dims[3],
                                      const float origin[3], const float spacing[3],
                                      pnm_pore_t* pores_d, int64_t cap_pores, int64_t* npores,
                                      int64_t* total_pores, int32_t* seg_d, int32_t* pairs_d,
                                      int64_t cap_pairs, int64_t* npairs);

/* extract_network_flow_k on device-resident SDF and fields; results come back to host. */
pnm_status pnm_extract_network_flow_d(pnm_ctx_t ctx, const float* sdf_d, const int32_t dims[3],
                                      const float origin[3], const float spacing[3],
                                      const double* u_d, const double* v_d, const double* w_d,
                                      const double* p_d, const double* ox_d, const double* oy_d,
                                      const double* oz_d, const double grad_p[3],
                                      pnm_flow_result_t* out);

/* Throat geometry of extract_network_flow without flow fields (project extension: the reference
 * computes it inside extract_network_flow_k, pore_extraction.hpp:655-944, and has no call for it
 * alone). Returns the pores, throats, throat_area and throat_center that
 * pnm_extract_network_flow_d returns under binary openness (a face is open iff both of its cells
 * have sdf > 0, as pnm_stokes_solve_d writes it); pore_pressure, pore_residual, throat_flow and
 * throat_dp stay NULL. seg_d (nullable): a segmentation of this SDF (pnm_segment_volume_d, or
 * one volume of pnm_segment_batch_d); NULL segments here. centers_d / raw_d (both or neither,
 * nullable): device float[cap][3] / float[cap][4] receiving each throat's surrogate inputs when
 * nt <= cap: the interface centroid in voxel units (X,Y,Z) and log area, log radius of each
 * pore, log min-image separation of the pore centres. Release with pnm_flow_result_free. */
pnm_status pnm_throat_geometry_d(pnm_ctx_t ctx, const float* sdf_d, const int32_t* seg_d,
                                 const int32_t dims[3], const float origin[3],
                                 const float spacing[3], float* centers_d, float* raw_d,
                                 int64_t cap, pnm_flow_result_t* out);

/* ---- batched segmentation (project extension; no single-call counterpart in the reference:
 * the result equals nvol separate segment_volume calls, pore_extraction.hpp:389-397, each volume
 * its own periodic box numbered on its own) ---- */
#define PNM_BATCH_GRAPH 1    /* replay the persistent launch from a cached hipGraph */
#define PNM_BATCH_AUTOTUNE 2 /* time launch shapes once per batch shape and keep the fastest */
/* sdf, seg: host arrays of nvol DEVICE pointers (float32[n_k] in, int32[n_k] out);
 * dims: host int32[3*nvol], X,Y,Z per volume; spacing: common, min(spacing) must be > 0.
 * n_pores_d / n_solids_d: nullable device int32[nvol] receiving each volume's label counts.
 * One persistent kernel segments all volumes (work queue, largest volumes first). */
pnm_status pnm_segment_batch_d(pnm_ctx_t ctx, int32_t nvol, const float* const* sdf,
                               int32_t* const* seg, const int32_t* dims, const float spacing[3],
                               int32_t* n_pores_d, int32_t* n_solids_d, int32_t flags);
/* Launch shape used by the last batched call: threads per workgroup, persistent workgroups per
 * CU, total workgroups; tuned_ms = the winning time of the last autotune (0 if none). */
pnm_status pnm_batch_config(pnm_ctx_t ctx, int32_t* block, int32_t* per_cu, int32_t* groups,
                            float* tuned_ms);

/* ---- project MAC Stokes solver (data generation for extract_network_flow; the reference takes
 * its fields from an external solver, README.md:56-100, whose MAC layout it states at
 * pore_extraction.hpp:476-477) ----
 * Solves mu*(-Lap u) + grad p = force, div u = 0 on the periodic MAC grid: u on the -x face of
 * each cell, v on -y, w on -z, p at cell centres. Fluid = sdf > 0; a face is open iff both cells
 * are fluid (binary openness); closed faces carry zero velocity. Preconditioned MINRES in FP64
 * until the preconditioned residual falls below rtol times its initial value (or max_iter).
 * dims/spacing/force are XYZ. u_d..p_d: device double[n] outputs. ox_d/oy_d/oz_d (nullable):
 * device double[n] receiving the binary openness (1.0 / 0.0) in extract_network_flow's layout.
 * iterations / rel_residual (nullable): MINRES iterations and ||b - Mx|| / ||b||. */
pnm_status pnm_stokes_solve_d(pnm_ctx_t ctx, const float* sdf_d, const int32_t dims[3],
                              const float spacing[3], const double force[3], double mu,
                              double rtol, int32_t max_iter, double* u_d, double* v_d,
                              double* w_d, double* p_d, double* ox_d, double* oy_d, double* oz_d,
                              int32_t* iterations, double* rel_residual);

/* ---- throat-conductance surrogate (project extension, no reference counterpart: a learned
 * g = Q/dp per throat, trained on extract_network_flow's own output, which stays its fallback)
 * Network: 16^3 SDF patch -> three 3^3 convolutions (1->16->32->32, ReLU, 2x average pools,
 * global average) + 4 throat scalars -> dense 36->64 (ReLU) -> 1 = log g. Convolutions run on
 * the BF16 MFMA with FP32 accumulation; parameters are one flat FP32 vector.
 * patches_d: float[batch][16][16][16] (z, y, x); scalars_d: float[batch][4]; out_d: [batch]. */
int64_t pnm_sur_param_count(void);
/* He-normal weights, zero biases, from a seed; params: host float[count]. */
pnm_status pnm_sur_init(uint64_t seed, float* params, int64_t count);
/* Forward pass; keeps its activations in the context for the next backward call. */
pnm_status pnm_sur_forward_d(pnm_ctx_t ctx, int32_t batch, const float* patches_d,
                             const float* scalars_d, const float* params_d, float* out_d);
/* Gradients of sum_b dout[b] * out[b] w.r.t. the parameters, for the context's last forward
 * pass (same batch); grads_d: float[pnm_sur_param_count()] (overwritten). */
pnm_status pnm_sur_backward_d(pnm_ctx_t ctx, int32_t batch, const float* dout_d,
                              const float* params_d, float* grads_d);
/* Mean loss over the batch: Huber with threshold delta (delta > 0) or squared error
 * (delta <= 0) of pred - target; dpred = d(loss)/d(pred); loss_d nullable. */
pnm_status pnm_sur_loss_d(pnm_ctx_t ctx, int32_t batch, const float* pred_d,
                          const float* target_d, float delta, float* dpred_d, float* loss_d);
/* One Adam step (bias-corrected, step >= 1) on the flat parameter vector. */
pnm_status pnm_sur_adam_d(pnm_ctx_t ctx, float* params_d, const float* grads_d, float* m_d,
                          float* v_d, int32_t step, float lr, float beta1, float beta2,
                          float eps);
/* Throat patches: out[b] = sdf / scale on the 16^3 voxels starting at floor(center_b) - 7
 * (periodic); centers_d: float[batch][3], X,Y,Z in voxel units. */
pnm_status pnm_sur_patches_d(pnm_ctx_t ctx, const float* sdf_d, const int32_t dims[3],
                             const float* centers_d, int32_t batch, float scale, float* out_d);
/* Fused inference: out_d[b] = the network's output for throat b (the same value as
 * pnm_sur_forward_d on that throat's pnm_sur_patches_d patch, up to FP32 summation order inside
 * the MFMA), computed straight from the device SDF without materialising patches or
 * activations. scalars_d: float[batch][4] (standardised throat scalars). */
pnm_status pnm_sur_infer_d(pnm_ctx_t ctx, const float* sdf_d, const int32_t dims[3],
                           const float* centers_d, const float* scalars_d, int32_t batch,
                           float scale, const float* params_d, float* out_d);

/* ---- stage-level device entry points (parity tests of single kernels; no reference
 * counterpart as separate calls: each is one step of segmentVolumeView) ---- */

/* ---- segmentation without allocation (no counterpart in the reference) ----
 * For callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): the same results as pnm_segment_volume_d, computed in
 * a fixed scratch region the library holds per GPU (about 175 MiB, a static device array kept
 * until the process ends) instead of the context's arena. The call allocates nothing: no
 * hipMalloc/hipFree, no second stream, no timing events. Volumes are limited to
 * PNM_NOALLOC_MAX_PADDED voxels after rounding Nx, Ny, Nz up to the 64 x 8 x 16 tile
 * (PNM_NOALLOC_PADDED_VOXELS; 128^3 fits); larger ones fail with PNM_ECAPACITY before any work.
 * Every call on a GPU shares that GPU's region: calls that overlap on one GPU (other threads,
 * other contexts, other streams) race and must be serialised by the caller. */
#define PNM_NOALLOC_MAX_PADDED (1 << 21)
#define PNM_NOALLOC_PADDED_VOXELS(nx, ny, nz)                         \
  ((((int64_t)(nx) + 63) / 64 * 64) * (((int64_t)(ny) + 7) / 8 * 8) * \
   (((int64_t)(nz) + 15) / 16 * 16))
pnm_status pnm_segment_volume_noalloc_d(pnm_ctx_t ctx, const float* sdf_d, const int32_t dims[3],
                                        const float spacing[3], int32_t* seg_d, int32_t* roots_d);

/* Steps 1-3 (pore_extraction.hpp:173-269): marker labels after CCL and flood. */
pnm_status pnm_stage_labels_d(pnm_ctx_t ctx, const float* sdf_d, const int32_t dims[3],
                              const float spacing[3], int32_t* labels_d, int32_t* ccl_only_d);
/* Gradient-walk roots: all_voxels = 0 -> step 4 (:271-311), ascent from sdf>0 voxels and
 * descent from the others; all_voxels = 1 -> nf::flow_basins ascent (:621-652) from every
 * voxel. */
pnm_status pnm_stage_roots_d(pnm_ctx_t ctx, const float* sdf_d, const int32_t dims[3],
                             int all_voxels, int32_t* roots_d);

#ifdef __cplusplus
}
#endif

#endif /* PNM_CAPI_H */
