// Queued live gates (KBWH_FLAG_QUEUE): the latency path for small states.
//
// In the reference every live gate is a separate call into the backend
// (quantum_execution.rs:310-314), one pass over the state each.  With the
// queue flag, kbwh_gate validates the gate and appends it to the handle's
// queue; the queue is flushed before anything reads or replaces the state.
// A state of up to 12 qubits (64 KiB) is flushed in one launch: a single
// 1024-thread workgroup holds the whole state as one 12-qubit register tile
// (4 amplitudes per thread: 6 lane bits, 4 wave bits, 2 register bits) and
// applies every queued gate in order with the tiled pass's exact per-pair
// arithmetic, so results are the same bits as gate by gate.  The planner puts
// the most used targets on the register bits; only gates on wave bits touch
// LDS.  A state narrower than 12 qubits is padded with qubits that no gate
// touches (their amplitudes stay 0 and are neither loaded nor stored).
// Larger states are flushed through the tiled planner (also exact).
#include <string.h>

#include <vector>

#include "internal.h"
#include "state.h"
// the whole state of up to 12 qubits is one tile of 1024 threads
#undef KBWH_TILE_BITS
#undef KBWH_TILE_TBITS
#undef KBWH_TILE_WAVES
#define KBWH_TILE_BITS 12
#define KBWH_TILE_TBITS 10
#define KBWH_TILE_WAVES 4
#include "tile_core.h"

#pragma clang fp contract(off)

namespace kbwh {
namespace {

using namespace tile;

constexpr uint32_t kLiveQubits = TL;

__global__ __launch_bounds__(TT) void k_whole(dv2 *__restrict__ a, uint32_t n, TileSpec ts,
                                              const TGate *__restrict__ gates, uint32_t ng) {
    __shared__ dv2 lds[TT * kChunk];
    const uint32_t tid = threadIdx.x;
    const uint64_t size = 1ull << n;
    uint64_t dt = 0;
#pragma unroll
    for (int b = 0; b < TB; ++b)
        if ((tid >> b) & 1u) dt |= ts.qbit[b];
    dv2 v[TR];
#pragma unroll
    for (int r = 0; r < TR; ++r) {
        const uint64_t i = slot_offset(dt, r, ts);
        v[r] = i < size ? a[i] : mk2(0.0, 0.0);
    }
    // ng >= 1: the host flushes only a non-empty queue; the next gate's
    // record is loaded while the current gate is applied
    uint32_t g = 0;
    TGate G = gates[0];
    do {
        const TGate Gn = gates[g + 1 < ng ? g + 1 : g];
        if (G.kind == KZ || G.kind == KP || G.kind == KRZ)
            apply_diag(v, 0, tid, G);
        else
            apply_pair(v, lds, tid, G);
        G = Gn;
    } while (++g < ng);
#pragma unroll
    for (int r = 0; r < TR; ++r) {
        const uint64_t i = slot_offset(dt, r, ts);
        if (i < size) a[i] = v[r];
    }
}

}  // namespace

int32_t queue_gate(kbwh_state *s, int kind, double angle, uint32_t target, uint64_t ctrl_mask) {
    if (kind < KBWH_X || kind > KBWH_RZ) return set_error(KBWH_ERR_INVALID_ARGUMENT, "unknown gate kind");
    if (target >= s->n) return set_error(KBWH_ERR_INVALID_ARGUMENT, "target out of range");
    if (s->n < 64 && (ctrl_mask >> s->n)) return set_error(KBWH_ERR_INVALID_ARGUMENT, "control out of range");
    kbwh_op op;
    memset(&op, 0, sizeof op);
    op.kind = (uint8_t)kind;
    op.target = (uint8_t)target;
    op.param = -1;
    op.angle = angle;
    op.ctrl_mask = ctrl_mask;
    s->queue.push_back(op);
    return KBWH_OK;
}

int32_t flush_queue(kbwh_state *s) {
    if (s->queue.empty()) return KBWH_OK;
    std::vector<kbwh_op> q;
    q.swap(s->queue);
    if (s->block) {
        for (const kbwh_op &op : q) {
            int32_t rc = apply_gate(s, op.kind, op.angle, op.target, op.ctrl_mask);
            if (rc) return rc;
        }
        return KBWH_OK;
    }
    if (s->n > kLiveQubits) return run_tiled(s, q.data(), q.size(), nullptr);
    // one pass over a 12-qubit tile holding qubits 0..11 (n..11 are padding)
    Pass all;
    for (const kbwh_op &op : q) {
        const PhysOp po{(int)op.kind, op.target, op.ctrl_mask & ~(1ull << op.target), op.angle};
        add_op(all, po);
    }
    int pos_of[64];
    uint32_t q_at[TL];
    assign_positions(kLiveQubits, all, pos_of, q_at);
    const TileSpec ts = make_tile_spec(q_at);
    std::vector<TGate> tg;
    tg.reserve(all.ops.size());
    for (const PhysOp &o : all.ops) tg.push_back(make_tgate(o, pos_of, kLiveQubits));
    const size_t bytes = tg.size() * sizeof(TGate);
    if (s->tile_gates_cap < bytes) {
        hipError_t e = hipStreamSynchronize(s->stream);
        if (e == hipSuccess && s->tile_gates) e = hipFree(s->tile_gates);
        s->tile_gates = nullptr;
        s->tile_gates_cap = 0;
        if (e == hipSuccess) e = hipMalloc(&s->tile_gates, bytes);
        if (e != hipSuccess) return hip_status(e, "queue buffer");
        s->tile_gates_cap = bytes;
    }
    // the host copy is local: wait for the upload before it goes away
    hipError_t e = hipMemcpyAsync(s->tile_gates, tg.data(), bytes, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) {
        k_whole<<<1, TT, 0, s->stream>>>((dv2 *)s->amp, s->n, ts, (const TGate *)s->tile_gates,
                                         (uint32_t)tg.size());
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    s->last_passes = 1;
    s->last_gates = q.size();
    return hip_status(e, "queued gates");
}

}  // namespace kbwh
