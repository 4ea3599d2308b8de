// The tiled pass kernel k_tile in exact arithmetic (KBWH_RUN_TILED, Block
// states): FP contraction off, so every gate reproduces the per-gate path's
// operations.  Built on its own, like the contracted copy in tiled_fma.hip,
// so that only the two pass kernels carry their translation units' code
// generation flags.
#include <string.h>

#include "state.h"
#include "tile_core.h"
#include "tile_pass.h"

namespace kbwh {

hipError_t launch_tile_exact(double2 *amp, const void *spec, const void *store, const void *gates,
                             uint32_t num_gates, uint32_t ntiles, hipStream_t stream) {
    static_assert(sizeof(tile::TileSpec) == kTileSpecBytes && sizeof(tile::TGate) == kTGateBytes,
                  "tile records shared with tiled.hip");
    tile::TileSpec ts;
    memcpy(&ts, spec, sizeof ts);
    const tile::TGate *g = (const tile::TGate *)gates;
    const bool pf = ntiles < kPrefetchBelowTiles;
    if (store) {
        tile::TileSpec st;
        memcpy(&st, store, sizeof st);
        if (pf)
            k_tile<true, true><<<ntiles, tile::TT, 0, stream>>>((tile::dv2 *)amp, ts, g, num_gates, ntiles, st, 32u);
        else
            k_tile<true, false><<<ntiles, tile::TT, 0, stream>>>((tile::dv2 *)amp, ts, g, num_gates, ntiles, st, 32u);
    } else if (pf) {
        k_tile<false, true><<<ntiles, tile::TT, 0, stream>>>((tile::dv2 *)amp, ts, g, num_gates, ntiles, ts, 32u);
    } else {
        k_tile<false, false><<<ntiles, tile::TT, 0, stream>>>((tile::dv2 *)amp, ts, g, num_gates, ntiles, ts, 32u);
    }
    return hipGetLastError();
}

}  // namespace kbwh
