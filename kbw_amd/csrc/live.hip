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
//
// A Block state queues gates as the reference's Block does (block.rs:179-208):
// the target's map and LRU stamp, and the swap of a global target, are decided
// when the gate is given; the gate is recorded on physical qubits and the swap
// as an event before it.  The reference flushes before each swap and at
// measure_p1 and dump (block.rs:132, 302-320); here a flush happens before any
// read of the state and runs gates and swaps together in exact tiled passes
// (swaps folded into passes), the same amplitudes either way: a swap moves
// amplitudes and never recomputes one.  measure_collapse does not flush, as in
// block.rs:311-316: it applies the queue up to its last swap (the part the
// reference has already applied) and leaves the gates after it queued; a
// target that needs a swap flushes everything first.
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
    if (s->block) {
        // block.rs:233-300: the target is made local now (the LRU stamp and,
        // for a global target, the swap and the new map), then the controls
        // are read through the updated map; the amplitudes move at the flush
        uint32_t t = target, sl = kNoSwap, sg = kNoSwap;
        map_qubit(s, target, &t, &sl, &sg);
        if (sl != kNoSwap) s->block_queue.push_back(kbwh_phys_event{KX, 0, 0, 0.0, sl, sg});
        const uint64_t m = physical_mask(s, ctrl_mask) & ~(1ull << t);
        s->block_queue.push_back(kbwh_phys_event{kind, t, m, angle, kNoSwap, kNoSwap});
        return KBWH_OK;
    }
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

bool queue_pending(const kbwh_state *s) { return !s->queue.empty() || !s->block_queue.empty(); }

// A Block state's queue up to and including its last swap.  The reference
// flushes before every swap (block.rs:132), so at any moment it holds only the
// gates given after the last swap; an operation that does not flush
// (measure_collapse, block.rs:311-316) must see the state with everything
// before that point applied.
int32_t flush_through_last_swap(kbwh_state *s) {
    std::vector<kbwh_phys_event> &q = s->block_queue;
    size_t last = q.size();
    for (size_t i = q.size(); i-- > 0;)
        if (q[i].swap_local != kNoSwap) {
            last = i;
            break;
        }
    if (last == q.size()) return KBWH_OK;
    std::vector<kbwh_phys_event> head(q.begin(), q.begin() + (ptrdiff_t)last + 1);
    q.erase(q.begin(), q.begin() + (ptrdiff_t)last + 1);
    return run_events(s, s->amp, head);
}

int32_t flush_queue(kbwh_state *s) {
    if (!s->block_queue.empty()) {
        // a Block state's queue (block.rs:179-208, flushed where the reference
        // flushes or earlier): gates and swaps in program order, in exact
        // tiled passes with the swaps folded in
        std::vector<kbwh_phys_event> ev;
        ev.swap(s->block_queue);
        uint64_t gates = 0;
        for (const kbwh_phys_event &E : ev) gates += E.swap_local == kNoSwap;
        const int32_t rc = run_events(s, s->amp, ev);
        s->last_gates = gates;
        return rc;
    }
    if (s->queue.empty()) return KBWH_OK;
    std::vector<kbwh_op> q;
    q.swap(s->queue);
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
        if (e == hipSuccess) e = dev_alloc(&s->tile_gates, bytes);
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
