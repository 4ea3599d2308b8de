// Multi-gate tiled passes (KBWH_RUN_TILED).
//
// A pass applies a run of consecutive gates to the whole state in one HBM
// read and one HBM write.  The state is cut into tiles of 2^TL amplitudes
// (TL = 11 by default): a tile is the set of indices that vary over TL chosen
// qubits Q (the tile positions) with every other bit fixed by the tile
// number.  Q always holds qubits 0..2, so every 128-byte line of the state
// belongs to one tile and is read whole.  A workgroup of 256 threads holds
// one tile in registers, 2^(TL-8) amplitudes per thread:
//   tile position 0..5    -> lane bits      (exchange with cross-lane shuffles)
//   tile position 6..7    -> wave bits      (exchange through LDS)
//   tile position 8..TL-1 -> register index (no exchange)
// A non-diagonal gate (X, Y, H, RX, RY) needs its target in Q; diagonal gates
// (Z, P, RZ) and every control may sit anywhere (bits outside Q are constant
// over a tile).  Gates are applied in program order and every output equals
// the reference's per-pair closure (dense_v2.rs:166-277) in value, so a tiled
// run produces the same amplitudes as one pass per gate; only the number of
// HBM passes changes.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "state.h"
#include "tile_core.h"

#pragma clang fp contract(off)
namespace kbwh {

namespace {

using namespace tile;

// One tile per workgroup: load the tile into registers, apply the gates,
// store it back; overlap comes from the other resident workgroups.
__global__ __launch_bounds__(TT) __attribute__((amdgpu_waves_per_eu(KBWH_TILE_WAVES, KBWH_TILE_WAVES))) void k_tile(
    dv2 *__restrict__ a, TileSpec ts, const TGate *__restrict__ gates, uint32_t ng, uint32_t ntiles) {
    __shared__ dv2 lds[TT * kChunk];
    const uint32_t tid = threadIdx.x;
    const uint32_t tile = blockIdx.x;
    if (tile >= ntiles) return;
    uint64_t tmask = 0;
#pragma unroll
    for (int b = 0; b < TB; ++b)
        if ((tid >> b) & 1u) tmask |= ts.qbit[b];
    const uint64_t base = tile_base(tile, ts);
    const uint64_t dt = base | tmask;
    dv2 v[TR];
#pragma unroll
    for (int r = 0; r < TR; ++r) v[r] = __builtin_nontemporal_load(a + slot_offset(dt, r, ts));
    // ng >= 1 (the host launches only non-empty passes); a do-while leaves no
    // path around the loop, so the loaded values need no copy kept for one.
    // The next gate's record is loaded while the current gate is applied.
    uint32_t i = 0;
    TGate G = gates[0];
    do {
        const TGate Gn = gates[i + 1 < ng ? i + 1 : i];
        if ((base & G.ocm) == G.ocm) {  // a control outside the tile is 0 here
            if (G.kind == KZ || G.kind == KP || G.kind == KRZ)
                apply_diag(v, base, tid, G);
            else
                apply_pair(v, lds, tid, G);
        }
        G = Gn;
    } while (++i < ng);
    // recompute the store addresses instead of keeping eight of them live
    // through the gate loop
    uint64_t ds = dt;
    asm volatile("" : "+v"(ds));
#pragma unroll
    for (int r = 0; r < TR; ++r) __builtin_nontemporal_store(v[r], a + slot_offset(ds, r, ts));
}


}  // namespace

int32_t run_tiled(kbwh_state *s, const kbwh_op *ops, uint64_t num_ops, const double *params) {
    return run_tiled_on(s, s->amp, ops, num_ops, params);
}

int32_t run_tiled_on(kbwh_state *s, double2 *amp, const kbwh_op *ops, uint64_t num_ops,
                     const double *params) {
    if (s->n < TL) {
        // the whole state is smaller than one tile: one pass per gate
        if (amp == s->amp) return kbwh_run(s, ops, num_ops, params, KBWH_RUN_PER_GATE);
        for (uint64_t i = 0; i < num_ops; ++i) {
            const double angle = (params && ops[i].param >= 0) ? params[ops[i].param] : ops[i].angle;
            hipError_t e = launch_gate(amp, s->n, ops[i].kind, angle, ops[i].target,
                                       ops[i].ctrl_mask & ~(1ull << ops[i].target), s->stream);
            if (e != hipSuccess) return hip_status(e, "gate launch");
        }
        return KBWH_OK;
    }
    if (s->block && amp != s->amp) return set_error(KBWH_ERR_INVALID_ARGUMENT, "Block state on a side buffer");
    const uint32_t n = s->n;
    s->last_passes = 0;
    for (uint64_t i = 0; i < num_ops; ++i) {
        const kbwh_op &op = ops[i];
        if (op.kind > KRZ) return set_error(KBWH_ERR_INVALID_ARGUMENT, "unknown gate kind");
        if (op.target >= n) return set_error(KBWH_ERR_INVALID_ARGUMENT, "target out of range");
        if (n < 64 && (op.ctrl_mask >> n)) return set_error(KBWH_ERR_INVALID_ARGUMENT, "control out of range");
    }
    Pass cur;
    // Execute: build passes lazily, flushing before any Block swap.
    std::vector<TGate> gates;
    struct Launch {
        TileSpec spec;
        uint32_t first, count;
    };
    std::vector<Launch> launches;
    auto flush = [&]() -> int32_t {
        if (cur.ops.empty()) return KBWH_OK;
        int pos_of[64];
        uint32_t q_at[TL];
        assign_positions(n, cur, pos_of, q_at);
        Launch L;
        L.spec = make_tile_spec(q_at);
        L.first = (uint32_t)gates.size();
        for (const PhysOp &o : cur.ops) gates.push_back(make_tgate(o, pos_of, n));
        L.count = (uint32_t)cur.ops.size();
        launches.push_back(L);
        cur = Pass();
        return KBWH_OK;
    };
    auto launch_all = [&]() -> int32_t {
        if (launches.empty()) return KBWH_OK;
        const size_t bytes = gates.size() * sizeof(TGate);
        if (s->tile_gates_cap < bytes) {
            hipError_t e = hipStreamSynchronize(s->stream);
            if (e == hipSuccess && s->tile_gates) e = hipFree(s->tile_gates);
            s->tile_gates = nullptr;
            s->tile_gates_cap = 0;
            if (e == hipSuccess) e = hipMalloc(&s->tile_gates, bytes);
            if (e != hipSuccess) return hip_status(e, "tile gate buffer");
            s->tile_gates_cap = bytes;
        }
        hipError_t e = hipMemcpyAsync(s->tile_gates, gates.data(), bytes, hipMemcpyHostToDevice, s->stream);
        if (e != hipSuccess) return hip_status(e, "tile gate upload");
        s->last_passes += launches.size();
        const uint32_t ntiles = (uint32_t)(1ull << (n - TL));
        const uint32_t grid = ntiles;
        const TGate *dg = (const TGate *)s->tile_gates;
        for (const Launch &L : launches) {
            k_tile<<<grid, TT, 0, s->stream>>>((dv2 *)amp, L.spec, dg + L.first, L.count, ntiles);
            e = hipGetLastError();
            if (e != hipSuccess) return hip_status(e, "k_tile launch");
        }
        // the gate buffer is reused by the next run: the upload above is
        // stream-ordered, and the next upload waits for these kernels
        launches.clear();
        gates.clear();
        return KBWH_OK;
    };
    for (uint64_t i = 0; i < num_ops; ++i) {
        const kbwh_op &op = ops[i];
        const double angle = (params && op.param >= 0) ? params[op.param] : op.angle;
        uint32_t t = op.target;
        if (s->block && s->l2p[op.target] >= s->nlocal) {
            // a global target: finish the pending passes, then swap (block.rs:132-150)
            int32_t rc = flush();
            if (rc == KBWH_OK) rc = launch_all();
            if (rc) return rc;
        }
        int32_t rc = local_qubit(s, op.target, &t);
        if (rc) return rc;
        uint64_t m = physical_mask(s, op.ctrl_mask) & ~(1ull << t);
        const PhysOp po{(int)op.kind, t, m, angle};
        if (!fits(cur, po, s->tile_targets) && (rc = flush())) return rc;
        add_op(cur, po);
    }
    int32_t rc = flush();
    if (rc == KBWH_OK) rc = launch_all();
    s->last_gates = num_ops;
    return rc;
}

}  // namespace kbwh
