/*
 * pnm_oracle.h — serial CPU oracle for the pore-network extraction path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker the HIP path is compared against: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The product
 * library (libpnm_hip.so) never links or calls it.
 *
 * It restates, serially and in plain C, the algorithm of src/pore_extraction.hpp of
 * computational-chemical-engineering/peclet-pnm. Every comparison, tie-break, accumulation order
 * and FP32/FP64 boundary follows that file (line numbers cited per function in pnm_oracle.c).
 * Compile with -ffp-contract=off and no fast-math so that FP32 centroids and FP64 sums are the
 * plain IEEE results of the written expression order.
 *
 * Conventions: dims/origin/spacing are XYZ (the Python layer reorders zyx -> xyz and casts to
 * float). Flat index i = z*Ny*Nx + y*Nx + x; every stencil wraps periodically.
 */
#ifndef PNM_ORACLE_H
#define PNM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  float x, y, z, radius;
} pnmo_pore;

/* Pores: 26-neighbour SDF maxima with squared-SDF-weighted 3x3x3 centroid (a4).
 * Emits pores in ascending peak-voxel index. Returns the total number of peaks found; writes
 * min(total, cap) records into `out` (out may be NULL when cap == 0). */
int64_t pnmo_extract_pores(const float* sdf, const int32_t dims[3], const float origin[3],
                           const float spacing[3], pnmo_pore* out, int64_t cap);

/* Segmentation stages (a6-a10), exposed individually for stage-level parity tests. */
void pnmo_init_markers(const float* sdf, const int32_t dims[3], const float spacing[3],
                       int32_t* labels);
int pnmo_ccl(const int32_t dims[3], int32_t* labels);                 /* returns sweeps */
int pnmo_flood(const float* sdf, const int32_t dims[3], int32_t* labels); /* returns sweeps */
/* ascent_all = 0: segmentation roots (ascent for sdf>0 voxels; other voxels get roots[i]=i,
 * since the reference never consumes their descent roots). ascent_all = 1: ascent from every
 * voxel (the flow-basin walk, a20). */
void pnmo_gradient_roots(const float* sdf, const int32_t dims[3], int ascent_all, int32_t* roots);
void pnmo_relabel(const float* sdf, const int32_t* roots, const int32_t* labels, int64_t n,
                  int32_t* seg);

/* Full segment_volume (a6-a11). roots_out / labels_out may be NULL. Returns 0. */
int pnmo_segment_volume(const float* sdf, const int32_t dims[3], const float spacing[3],
                        int32_t* seg, int32_t* roots_out, int32_t* labels_out);

/* Topology (a12): sorted unique (min,max) signed label pairs across +x/+y/+z faces.
 * Returns the number of pairs; *pairs receives a malloc'ed int32[2*count] (free with
 * pnmo_free). */
int64_t pnmo_extract_topology(const int32_t* seg, const int32_t dims[3], int32_t** pairs);

/* Network flow (a15-a28). */
typedef struct {
  int64_t np, nt;
  pnmo_pore* pores;       /* [np], label order */
  double* pore_pressure;  /* [np] */
  double* pore_residual;  /* [np] */
  int32_t* throats;       /* [2*nt] (li, lj) */
  double* throat_flow;    /* [nt] */
  double* throat_area;    /* [nt] */
  double* throat_dp;      /* [nt] */
  /* intermediates for stage-level tests (NULL unless want_debug) */
  int32_t* flow_lab;      /* [n] */
  int64_t* fpar;          /* [3n] */
  int32_t* peak;          /* [np] */
  double* cx;             /* [nt] */
  double* cy;
  double* cz;
} pnmo_flow_result;

int pnmo_extract_network_flow(const float* sdf, const int32_t dims[3], const float origin[3],
                              const float spacing[3], const double* u, const double* v,
                              const double* w, const double* p, const double* ox,
                              const double* oy, const double* oz, const double grad_p[3],
                              int want_debug, pnmo_flow_result* out);
void pnmo_flow_free(pnmo_flow_result* r);

void pnmo_free(void* p);

#ifdef __cplusplus
}
#endif

#endif
