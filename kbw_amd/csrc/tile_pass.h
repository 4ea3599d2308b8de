// The tiled pass kernel k_tile, shared by the exact tiled runs
// (tiled_pass.hip, built with FP contraction off) and the contracted copy used
// by dependency-ordered runs (tiled_fma.hip).  Include after tile_core.h.
//
// Both units are compiled with uniform branches left unstructurized
// (-structurizecfg-skip-uniform-regions): the per-gate dispatch below is a
// chain of uniform tests, and structurizing it into guarded blocks kept the
// loaded register tile live across every gate path, so each path wrote a
// second register set that was copied back after every gate.
#pragma once

// the kernel's name: the contracted copy is named apart in kernel traces
#ifndef KBWH_TILE_PASS_NAME
#define KBWH_TILE_PASS_NAME k_tile
#endif

namespace kbwh {
namespace {

using namespace tile;

// One tile per workgroup: load the tile into registers, apply the gates,
// store it back; overlap comes from the other resident workgroups.
// PERM: Block swaps folded into the pass; the tile is stored through st, whose
// qubits are those of ts in another order (position b of the tile goes to
// qubit st.qbit[b]), so exchanged qubits cost no extra traffic.
// Gate records: a state of few tiles (PF) has too few waves to cover a scalar
// load per gate, so the next record is loaded while the current gate is
// applied; on a large state each record is read at the top of its step (the
// other resident waves cover the latency), which saves copying the prefetched
// record through scalar registers on every gate.
constexpr uint32_t kPrefetchBelowTiles = 1024;

// Sets: a state that holds several independent states of the same circuit
// (parameter sets) one after another takes each set's own gate records:
// tile t reads gates + (t >> set_shift) * ng (set_shift >= 32: one set).
template <bool PERM, bool PF>
__global__ __launch_bounds__(TT) __attribute__((amdgpu_waves_per_eu(KBWH_TILE_WAVES, KBWH_TILE_WAVES))) void KBWH_TILE_PASS_NAME(
    dv2 *__restrict__ a, TileSpec ts, const TGate *__restrict__ gates, uint32_t ng, uint32_t ntiles,
    TileSpec st, uint32_t set_shift) {
    __shared__ dv2 lds[TT * kChunk];
    const uint32_t tid = threadIdx.x;
    const uint32_t tile = blockIdx.x;
    if (tile >= ntiles) return;
    if (set_shift < 32) gates += (size_t)(tile >> set_shift) * ng;
    const uint64_t base = tile_base(tile, ts);
    const uint64_t dt = thread_offset(base, tid, ts);
    dv2 v[TR];
    tile_load(v, a, dt, ts);
    // ng >= 1 (the host launches only non-empty passes); a do-while leaves no
    // path around the loop, so the loaded values need no copy kept for one.
    auto step = [&](const TGate &G) { apply_gate(v, lds, base, tid, G); };
    uint32_t i = 0;
    if constexpr (PF) {
        TGate G = gates[0];
        do {
            const TGate Gn = gates[i + 1 < ng ? i + 1 : i];
            step(G);
            G = Gn;
        } while (++i < ng);
    } else {
        // walk a record pointer: its fields are then loaded at fixed offsets
        // from it (an index kept the loads on a pointer into the middle of
        // the record, rebased with 64-bit adds for every field)
        const TGate *gp = gates, *const ge = gates + ng;
        do {
            const TGate G = *gp;
            step(G);
        } while (++gp != ge);
    }
    if constexpr (PERM) {
        // the tile goes to addresses that other waves of the workgroup
        // loaded: every load must have returned before any store is issued
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        tile_store(v, a, thread_offset(base, tid, st), st);
    } else {
        tile_store(v, a, dt, ts);
    }
}

}  // namespace
}  // namespace kbwh
