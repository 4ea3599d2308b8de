// Multi-gate tiled passes (KBWH_RUN_TILED).
#include "state.h"

namespace kbwh {

int32_t run_tiled(kbwh_state *s, const kbwh_op *ops, uint64_t num_ops, const double *params) {
    (void)s;
    (void)ops;
    (void)num_ops;
    (void)params;
    return set_error(KBWH_ERR_INVALID_ARGUMENT, "tiled mode not built");
}

}  // namespace kbwh
