// Dependency-ordered runs in 12-qubit tiles on pass-specialised kernels
// (KBWH_RUN_COMPILED on flat states of 12 qubits and more).
//
// A pass reads and writes the whole state once, so the run's HBM traffic is
// the pass count times twice the state.  A 12-qubit tile (16 amplitudes per
// thread, 256 threads: lane bits 0..5, wave bits 6..7, register bits 8..11)
// leaves 9 free target positions beside qubits 0..2 instead of the 11-qubit
// tile's 8, and the plan search (tile_core.h) finds plans with fewer passes:
// on the 30-qubit, 20-layer brickwork circuit 16 passes instead of the
// 11-qubit tile's 21 (greedy) or 19 (searched).  The register tile is twice as
// large, which only the specialised kernels carry without spilling (built for
// 4 waves per SIMD, or 3 or 2 where that spills); the generic kernel keeps
// 11-qubit tiles.  The shape is kWideTileQubits / kWideTileTBits: 13-qubit
// tiles of 512 threads plan the same circuit in 14 passes, but their passes
// carry more gates (compute-bound beyond about 60) and their tile sets more
// scattered high qubits (memory floors up to 12.9 ms), and the run took 124
// ms against 118.  When a pass kernel cannot be compiled the run falls back to
// the 11-qubit passes before anything is launched.
#include <string.h>

#include <vector>

#include "internal.h"
#include "state.h"
#undef KBWH_TILE_BITS
#undef KBWH_TILE_TBITS
#undef KBWH_TILE_WAVES
#define KBWH_TILE_BITS 12
#define KBWH_TILE_TBITS 8
#define KBWH_TILE_WAVES 3
#include "tile_core.h"

namespace kbwh {
namespace {

using namespace tile;

static_assert(TL == (int)kWideTileQubits && TB == (int)kWideTileTBits, "wide tile shape");

// Passes longer than this are cut (a kernel's straight line compiles slowly
// past it); the cut keeps the order, so the pieces are passes of their own.
constexpr uint32_t kMaxWidePassGates = 256;

// Default randomised plans tried for a compiled run (about 0.1 s of host time
// for 900 gates on 30 qubits, once per circuit structure).
constexpr uint32_t kCompiledPlanTrials = 256;

struct WidePass {
    TileSpec spec;
    uint32_t first, count;
};

// The plan of ops in wide passes: records and one TileSpec per pass.
int32_t plan_wide(const kbwh_state *s, uint32_t n, const kbwh_op *ops, uint64_t num_ops, const double *params,
                  uint32_t trials, std::vector<TGate> &gates, std::vector<WidePass> &passes) {
    std::vector<PhysOp> phys(num_ops);
    for (uint64_t i = 0; i < num_ops; ++i) {
        const kbwh_op &op = ops[i];
        if (op.kind > KRZ || op.target >= n || (n < 64 && (op.ctrl_mask >> n)))
            return set_error(KBWH_ERR_INVALID_ARGUMENT, "op out of range");
        const double angle = (params && op.param >= 0) ? params[op.param] : op.angle;
        phys[i] = PhysOp{(int)op.kind, op.target, op.ctrl_mask & ~(1ull << op.target), angle};
    }
    fuse_diagonals(phys, n);
    const uint32_t cap = std::min<uint32_t>(s ? s->tile_targets : 64u, TL - 3);
    std::vector<uint32_t> order, ends;
    schedule_passes_cached(phys, n, cap, trials, order, ends);
    if (order.size() != phys.size()) return set_error(KBWH_ERR_DEVICE, "wide schedule incomplete");
    gates.clear();
    passes.clear();
    size_t k = 0;
    for (const uint32_t end : ends) {
        while (k < end) {
            const size_t stop = std::min<size_t>(end, k + kMaxWidePassGates);
            Pass cur;
            for (size_t i = k; i < stop; ++i) add_op(cur, phys[order[i]]);
            int pos_of[64];
            uint32_t q_at[TL];
            assign_positions(n, cur, pos_of, q_at);
            WidePass P;
            P.spec = make_tile_spec(q_at);
            P.first = (uint32_t)gates.size();
            for (const PhysOp &o : cur.ops) gates.push_back(make_tgate(o, pos_of, n));
            P.count = (uint32_t)(gates.size() - P.first);
            passes.push_back(P);
            k = stop;
        }
    }
    return KBWH_OK;
}

uint32_t trials_for(const kbwh_state *s) {
    return s->plan_trials < 0 ? kCompiledPlanTrials : (uint32_t)s->plan_trials;
}

}  // namespace

int32_t run_compiled_wide(kbwh_state *s, double2 *amp, const kbwh_op *ops, uint64_t num_ops, const double *params,
                          bool *done) {
    *done = false;
    const uint32_t n = s->n;
    if (n < (uint32_t)TL || n > 40) return KBWH_OK;
    std::vector<TGate> gates;
    std::vector<WidePass> passes;
    int32_t rc = plan_wide(s, n, ops, num_ops, params, trials_for(s), gates, passes);
    if (rc) return rc;
    s->last_passes = 0;
    s->last_gates = num_ops;
    *done = true;
    if (passes.empty()) return KBWH_OK;
    std::vector<JitPass> lst;
    lst.reserve(passes.size());
    for (const WidePass &P : passes) lst.push_back(JitPass{&gates[P.first], P.count, kJitWide});
    std::vector<void *> fns;
    if ((rc = jit_prepare(s->device, lst, fns))) return rc;
    for (void *f : fns)
        if (!f) {
            *done = false;  // the 11-qubit passes run this circuit
            return KBWH_OK;
        }
    const size_t bytes = gates.size() * sizeof(TGate);
    if (s->tile_gates_cap < bytes) {
        hipError_t e = hipStreamSynchronize(s->stream);
        if (e == hipSuccess && s->tile_gates) e = hipFree(s->tile_gates);
        s->tile_gates = nullptr;
        s->tile_gates_cap = 0;
        if (e == hipSuccess) e = dev_alloc(&s->tile_gates, bytes);
        if (e != hipSuccess) return hip_status(e, "tile gate buffer");
        s->tile_gates_cap = bytes;
    }
    // the records stay valid for the stream: the next run's upload is
    // stream-ordered after these kernels
    hipError_t e = hipMemcpyAsync(s->tile_gates, gates.data(), bytes, hipMemcpyHostToDevice, s->stream);
    if (e != hipSuccess) return hip_status(e, "tile gate upload");
    const uint32_t ntiles = (uint32_t)(1ull << (n - TL));
    const TGate *dg = (const TGate *)s->tile_gates;
    std::vector<OrderTrial> trials;
    trials.reserve(passes.size());
    for (size_t i = 0; i < passes.size() && e == hipSuccess; ++i) {
        trials.push_back(order_pick(s->device, fns[i], &passes[i].spec, sizeof(TileSpec), ntiles));
        OrderTrial &t = trials.back();
        if (t.slot) e = hipEventRecord(t.ev[0], s->stream);
        if (e == hipSuccess)
            e = jit_launch_spec(fns[i], amp, &passes[i].spec, sizeof(TileSpec), dg + passes[i].first, ntiles,
                                s->stream, t.order, (uint32_t)TT);
        if (e == hipSuccess && t.slot) e = hipEventRecord(t.ev[1], s->stream);
    }
    // the host records are local: the upload must finish before they go away
    const hipError_t se = hipStreamSynchronize(s->stream);
    for (OrderTrial &t : trials) {
        if (e != hipSuccess || se != hipSuccess) {
            // a failed run reports no times; its trials' events go away
            if (t.slot) {
                (void)hipEventDestroy(t.ev[0]);
                (void)hipEventDestroy(t.ev[1]);
            }
            continue;
        }
        order_report(t);
    }
    if (e != hipSuccess) return hip_status(e, "wide pass launch");
    s->last_passes = passes.size();
    return hip_status(se, "wide passes");
}

int32_t wide_pass_records(const kbwh_op *ops, uint64_t num_ops, uint32_t num_qubits, uint32_t pass,
                          std::vector<unsigned char> &recs, uint32_t &count) {
    if (num_qubits < (uint32_t)TL || num_qubits > 40) return set_error(KBWH_ERR_INVALID_ARGUMENT, "12..40 qubits");
    std::vector<TGate> gates;
    std::vector<WidePass> passes;
    const int32_t rc = plan_wide(nullptr, num_qubits, ops, num_ops, nullptr, kCompiledPlanTrials, gates, passes);
    if (rc) return rc;
    if (pass >= passes.size()) return set_error(KBWH_ERR_INVALID_ARGUMENT, "no such pass");
    const WidePass &P = passes[pass];
    recs.assign((const unsigned char *)&gates[P.first], (const unsigned char *)(&gates[P.first] + P.count));
    count = P.count;
    return KBWH_OK;
}

}  // namespace kbwh
