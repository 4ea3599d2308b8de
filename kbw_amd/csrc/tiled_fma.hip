// The tiled pass kernel built with FP contraction (KBWH_RUN_REORDER only).
//
// A dependency-ordered run already differs from the program-order result by
// rounding, so its passes use a copy of k_tile in which a product feeding an
// add is fused into one FMA: the two-term pair form and the complex products
// of phase gates take 4 FP64 operations per amplitude instead of 6.  Exact
// runs (KBWH_RUN_TILED, Block states) never use this copy.
#define KBWH_TILE_CONTRACT 1
#define KBWH_TILE_PASS_NAME k_tile_fma
#include <string.h>

#include "state.h"
#include "tile_core.h"
#include "tile_pass.h"

namespace kbwh {

hipError_t launch_tile_contracted(double2 *amp, const void *spec, const void *gates, uint32_t num_gates,
                                  uint32_t ntiles, hipStream_t stream, uint32_t set_shift) {
    static_assert(sizeof(tile::TileSpec) == kTileSpecBytes && sizeof(tile::TGate) == kTGateBytes,
                  "tile records shared with tiled.hip");
    tile::TileSpec ts;
    memcpy(&ts, spec, sizeof ts);
    const tile::TGate *g = (const tile::TGate *)gates;
    if (ntiles < kPrefetchBelowTiles)
        k_tile_fma<false, true><<<ntiles, tile::TT, 0, stream>>>((tile::dv2 *)amp, ts, g, num_gates, ntiles, ts,
                                                                 set_shift);
    else
        k_tile_fma<false, false><<<ntiles, tile::TT, 0, stream>>>((tile::dv2 *)amp, ts, g, num_gates, ntiles, ts,
                                                                  set_shift);
    return hipGetLastError();
}

}  // namespace kbwh
