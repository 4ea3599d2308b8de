// Adjoint differentiation (kbwh_adjoint_gradient).
#include "state.h"

namespace kbwh {
void free_adjoint(kbwh_state *s) {
    if (s->adj_phi) (void)hipFree(s->adj_phi);
    if (s->adj_lam) (void)hipFree(s->adj_lam);
    if (s->adj_partial) (void)hipFree(s->adj_partial);
    s->adj_phi = s->adj_lam = nullptr;
    s->adj_partial = nullptr;
}
}  // namespace kbwh

extern "C" int32_t kbwh_adjoint_gradient(kbwh_state *s, const kbwh_op *ops, uint64_t num_ops,
                                         const double *params, uint32_t num_params,
                                         const uint8_t *paulis, const uint32_t *qubits,
                                         const uint32_t *offsets, const double *coeffs,
                                         uint32_t num_strings, double *energy, double *grad) {
    return kbwh::set_error(KBWH_ERR_INVALID_ARGUMENT, "adjoint not built");
}
