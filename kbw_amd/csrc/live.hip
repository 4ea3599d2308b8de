// Queued live gates (KBWH_FLAG_QUEUE): the latency path for small states.
//
// In the reference every live gate is a separate call into the backend
// (quantum_execution.rs:310-314), one pass over the state each.  With the
// queue flag, kbwh_gate validates the gate and appends it to the handle's
// queue; the queue is flushed before anything reads or replaces the state.
// A state of up to 12 qubits (64 KiB) is flushed in one launch: a single
// 1024-thread workgroup holds the whole state in LDS and applies every queued
// gate in order, a barrier between gates, with the per-gate kernels' exact
// arithmetic, so results are the same bits as gate by gate.  Larger states
// are flushed through the tiled planner (also exact); Block states apply the
// queue gate by gate (their map and LRU follow each gate).
#include <string.h>

#include <vector>

#include "gate_math.h"
#include "internal.h"
#include "state.h"

#pragma clang fp contract(off)

namespace kbwh {
namespace {

constexpr uint32_t kLiveQubits = 12;
constexpr int kLiveThreads = 1024;

struct LGate {
    uint32_t kind, target;
    uint32_t ctrl, pad;  // controls without the target (n <= 12)
    double c, s;
    double2 ph0, ph1;
};
static_assert(sizeof(LGate) == 64, "LGate layout");

// The gate's kind is dispatched on its bits (nested two-way branches, one
// predecessor per case) rather than with a switch.
__device__ __forceinline__ void live_pair(double2 &t0, double2 &t1, const LGate &G) {
    const uint32_t k = G.kind;
    if (k & 4u) {
        if (k & 2u) {
            if (k & 1u)
                gm::diag_math<gm::KRZ>(t0, t1, G.ph0, G.ph1);
            else
                gm::pair_math<gm::KRY>(t0, t1, G.c, G.s);
        } else {
            if (k & 1u)
                gm::pair_math<gm::KRX>(t0, t1, G.c, G.s);
            else
                gm::diag_math<gm::KP>(t0, t1, G.ph0, G.ph1);
        }
    } else {
        if (k & 2u) {
            if (k & 1u)
                gm::pair_math<gm::KH>(t0, t1, G.c, G.s);
            else
                gm::diag_math<gm::KZ>(t0, t1, G.ph0, G.ph1);
        } else {
            if (k & 1u)
                gm::pair_math<gm::KY>(t0, t1, G.c, G.s);
            else
                gm::pair_math<gm::KX>(t0, t1, G.c, G.s);
        }
    }
}

__global__ __launch_bounds__(kLiveThreads) void k_live(double2 *__restrict__ a, uint32_t n,
                                                        const LGate *__restrict__ gates, uint32_t ng) {
    __shared__ double2 st[1u << kLiveQubits];
    const uint32_t size = 1u << n, tid = threadIdx.x;
    for (uint32_t i = tid; i < size; i += kLiveThreads) st[i] = a[i];
    __syncthreads();
    const uint32_t npairs = size >> 1;
    for (uint32_t g = 0; g < ng; ++g) {
        const LGate G = gates[g];
        const uint32_t tb = 1u << G.target, lo = tb - 1u;
        for (uint32_t p = tid; p < npairs; p += kLiveThreads) {
            const uint32_t i = ((p & ~lo) << 1) | (p & lo);
            if ((i & G.ctrl) != G.ctrl) continue;
            double2 t0 = st[i], t1 = st[i | tb];
            live_pair(t0, t1, G);
            st[i] = t0;
            st[i | tb] = t1;
        }
        __syncthreads();
    }
    for (uint32_t i = tid; i < size; i += kLiveThreads) a[i] = st[i];
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
    std::vector<LGate> lg(q.size());
    for (size_t i = 0; i < q.size(); ++i) {
        const kbwh_op &op = q[i];
        const Coef cf = make_coef(op.kind, op.angle);
        LGate &g = lg[i];
        memset(&g, 0, sizeof g);
        g.kind = op.kind;
        g.target = op.target;
        g.ctrl = (uint32_t)(op.ctrl_mask & ~(1ull << op.target));
        g.c = cf.c;
        g.s = cf.s;
        g.ph0 = cf.ph0;
        g.ph1 = cf.ph1;
    }
    const size_t bytes = lg.size() * sizeof(LGate);
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
    hipError_t e = hipMemcpyAsync(s->tile_gates, lg.data(), bytes, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) {
        k_live<<<1, kLiveThreads, 0, s->stream>>>(s->amp, s->n, (const LGate *)s->tile_gates, (uint32_t)lg.size());
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    s->last_passes = 1;
    s->last_gates = q.size();
    return hip_status(e, "queued gates");
}

}  // namespace kbwh
