/* sanitize_driver.c — runs every function of the serial oracle on small and degenerate volumes
 * (1x1x1, thin, tiny periodic boxes where neighbours alias, quantised values with ties, NaN and
 * infinities, fractional openness) in a build with AddressSanitizer and UndefinedBehaviour-
 * Sanitizer (`make -C oracle sanitize`). Any out-of-bounds access, leak or undefined operation
 * aborts with a report; tests/test_oracle.py runs it. Test infrastructure only. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "pnm_oracle.h"

static uint64_t rng_state = 88172645463325252ull;
static double urand(void) { /* xorshift64 */
  rng_state ^= rng_state << 13;
  rng_state ^= rng_state >> 7;
  rng_state ^= rng_state << 17;
  return (double)(rng_state >> 11) / 9007199254740992.0;
}

static void run_case(int nx, int ny, int nz, int special, int with_open) {
  const int32_t dims[3] = {nx, ny, nz};
  const int64_t n = (int64_t)nx * ny * nz;
  float* sdf = malloc(sizeof(float) * (size_t)n);
  double* f[7];
  for (int k = 0; k < 7; ++k) f[k] = malloc(sizeof(double) * (size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    sdf[i] = (float)(floor(urand() * 12.0) * 0.5 - 3.0);
    if (special && urand() < 0.05) sdf[i] = NAN;
    if (special && urand() < 0.02) sdf[i] = (urand() < 0.5) ? INFINITY : -INFINITY;
    for (int k = 0; k < 4; ++k) f[k][i] = urand() - 0.5;
    for (int k = 4; k < 7; ++k) f[k][i] = urand() < 0.2 ? 0.0 : urand();
  }
  const float org[3] = {0.5f, -1.0f, 2.0f}, spc[3] = {0.75f, 0.5f, 1.0f};
  pnmo_pore* pores = malloc(sizeof(pnmo_pore) * (size_t)(n + 1));
  (void)pnmo_extract_pores(sdf, dims, org, spc, pores, n + 1);
  int32_t* seg = malloc(sizeof(int32_t) * (size_t)n);
  int32_t* roots = malloc(sizeof(int32_t) * (size_t)n);
  int32_t* labels = malloc(sizeof(int32_t) * (size_t)n);
  pnmo_segment_volume(sdf, dims, spc, seg, roots, labels);
  pnmo_gradient_roots(sdf, dims, 1, roots);
  int32_t* pairs = NULL;
  (void)pnmo_extract_topology(seg, dims, &pairs);
  pnmo_free(pairs);
  const double gp[3] = {0.01, -0.02, 0.005};
  pnmo_flow_result r;
  pnmo_extract_network_flow(sdf, dims, org, spc, f[0], f[1], f[2], f[3],
                            with_open ? f[4] : NULL, with_open ? f[5] : NULL,
                            with_open ? f[6] : NULL, gp, 1, &r);
  pnmo_flow_free(&r);
  free(pores);
  free(seg);
  free(roots);
  free(labels);
  for (int k = 0; k < 7; ++k) free(f[k]);
  free(sdf);
}

int main(void) {
  static const int shapes[][3] = {{1, 1, 1}, {1, 1, 2}, {2, 1, 1}, {2, 2, 2}, {1, 3, 5},
                                  {3, 3, 3}, {7, 3, 9}, {9, 8, 10}, {16, 4, 3}, {12, 12, 12}};
  int cases = 0;
  for (size_t s = 0; s < sizeof(shapes) / sizeof(shapes[0]); ++s)
    for (int special = 0; special < 2; ++special)
      for (int open = 0; open < 2; ++open) {
        run_case(shapes[s][0], shapes[s][1], shapes[s][2], special, open);
        ++cases;
      }
  printf("sanitized oracle: %d cases clean\n", cases);
  return 0;
}
