// k_flow.hip — network-flow bookkeeping (extract_network_flow_k, pore_extraction.hpp:493-997).
#include "k_common.hpp"

namespace pnm {

void flow_device(Ctx* c, const float* sdf, Vol v, const float org[3], const float spc[3],
                 const double* u, const double* vv, const double* w, const double* p,
                 const double* ox, const double* oy, const double* oz, const double grad_p[3],
                 pnm_flow_result_t* out) {
  (void)c; (void)sdf; (void)v; (void)org; (void)spc; (void)u; (void)vv; (void)w; (void)p;
  (void)ox; (void)oy; (void)oz; (void)grad_p; (void)out;
  fail(PNM_EHIP, "extract_network_flow: not built yet");
}

}  // namespace pnm
