// Run-configuration autotuner (kbwh_autotune, kbwh_set_run_config).
//
// The cost of a gate list depends on how its gates fall into passes (their
// targets, controls and order), so the configuration is chosen by timing the
// handle's own gate list: each candidate runs once to warm up and once timed
// with HIP events on a scratch buffer the size of the state (pass costs do
// not depend on amplitude values), and the fastest one is kept for
// KBWH_RUN_AUTO.  The handle's state is not touched.
#include <string.h>

#include <algorithm>

#include "internal.h"
#include "state.h"

namespace kbwh {
namespace {

constexpr uint64_t kTuneOps = 256;

struct Candidate {
    uint32_t mode, fusion_width, tile_targets;
    bool exact;
};

const Candidate kCandidates[] = {
    {KBWH_RUN_PER_GATE, 5, 64, true}, {KBWH_RUN_TILED, 5, 64, true}, {KBWH_RUN_TILED, 5, 6, true},
    {KBWH_RUN_TILED, 5, 4, true},     {KBWH_RUN_FUSED, 4, 64, false}, {KBWH_RUN_FUSED, 5, 64, false},
    {KBWH_RUN_REORDER, 5, 64, false},
};

int32_t run_on(kbwh_state *s, double2 *buf, const Candidate &c, const kbwh_op *ops, uint64_t n,
               const double *params) {
    switch (c.mode) {
    case KBWH_RUN_PER_GATE:
        for (uint64_t i = 0; i < n; ++i) {
            const double angle = (params && ops[i].param >= 0) ? params[ops[i].param] : ops[i].angle;
            hipError_t e = launch_gate(buf, s->n, ops[i].kind, angle, ops[i].target,
                                       ops[i].ctrl_mask & ~(1ull << ops[i].target), s->stream);
            if (e != hipSuccess) return hip_status(e, "gate launch");
        }
        return KBWH_OK;
    case KBWH_RUN_TILED:
    case KBWH_RUN_REORDER: {
        const uint32_t keep = s->tile_targets;
        s->tile_targets = c.tile_targets;
        const int32_t rc = run_tiled_on(s, buf, ops, n, params, c.mode == KBWH_RUN_REORDER);
        s->tile_targets = keep;
        return rc;
    }
    default:
        return run_fused(s, buf, ops, n, params, (int)c.fusion_width);
    }
}

}  // namespace
}  // namespace kbwh

using namespace kbwh;

extern "C" {

int32_t kbwh_set_run_config(kbwh_state *s, uint32_t fusion_width, uint32_t tile_targets) {
    if (!s) return set_error(KBWH_ERR_INVALID_ARGUMENT, "null handle");
    if (fusion_width != 4 && fusion_width != 5) return set_error(KBWH_ERR_INVALID_ARGUMENT, "fusion width must be 4 or 5");
    if (tile_targets < 1 || tile_targets > 64) return set_error(KBWH_ERR_INVALID_ARGUMENT, "tile targets must be 1..64");
    s->fusion_width = fusion_width;
    s->tile_targets = tile_targets;
    return KBWH_OK;
}

int32_t kbwh_set_compiled(kbwh_state *s, int32_t enable) {
    if (!s) return set_error(KBWH_ERR_INVALID_ARGUMENT, "null handle");
    s->compiled_passes = enable != 0;
    return KBWH_OK;
}

int32_t kbwh_set_plan_search(kbwh_state *s, int32_t trials) {
    if (!s) return set_error(KBWH_ERR_INVALID_ARGUMENT, "null handle");
    if (trials > 65536) return set_error(KBWH_ERR_INVALID_ARGUMENT, "at most 65536 trials");
    s->plan_trials = trials < 0 ? -1 : trials;
    return KBWH_OK;
}

int32_t kbwh_autotune(kbwh_state *s, const kbwh_op *ops, uint64_t num_ops, const double *params,
                      int32_t allow_inexact, kbwh_tune_result *out) {
    if (!s || !out || (num_ops && !ops)) return set_error(KBWH_ERR_INVALID_ARGUMENT, "null argument");
    if (s->block) return set_error(KBWH_ERR_INVALID_ARGUMENT, "autotuning needs a state without blocks");
    for (uint64_t i = 0; i < num_ops; ++i) {
        if (ops[i].kind > 7 || ops[i].target >= s->n || (s->n < 64 && (ops[i].ctrl_mask >> s->n)))
            return set_error(KBWH_ERR_INVALID_ARGUMENT, "op out of range");
    }
    memset(out, 0, sizeof *out);
    const uint64_t n = std::min(num_ops, kTuneOps);
    const uint64_t keep_passes = s->last_passes, keep_gates = s->last_gates;
    double2 *buf = nullptr;
    hipError_t e = hipStreamSynchronize(s->stream);
    if (e == hipSuccess) e = dev_alloc(&buf, s->size * sizeof(double2));
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
        (void)hipGetLastError();
        return set_error(KBWH_ERR_OUT_OF_MEMORY, "no device memory for the tuning buffer");
    }
    if (e == hipSuccess) e = hipMemsetAsync(buf, 0, s->size * sizeof(double2), s->stream);
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&t0);
    if (e == hipSuccess) e = hipEventCreate(&t1);
    int32_t rc = hip_status(e, "tuning setup");
    double best = 1e300;
    uint32_t count = 0;
    for (const Candidate &c : kCandidates) {
        if (rc) break;
        const uint32_t idx = count++;
        if (!c.exact && !allow_inexact) continue;
        rc = run_on(s, buf, c, ops, n, params);  // warm-up
        if (rc) break;
        e = hipEventRecord(t0, s->stream);
        if (e == hipSuccess) rc = run_on(s, buf, c, ops, n, params);
        if (rc) break;
        if (e == hipSuccess) e = hipEventRecord(t1, s->stream);
        if (e == hipSuccess) e = hipEventSynchronize(t1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
        if (e != hipSuccess) {
            rc = hip_status(e, "tuning run");
            break;
        }
        out->ms[idx] = ms;
        if (ms < best) {
            best = ms;
            out->mode = c.mode;
            out->fusion_width = c.fusion_width;
            out->tile_targets = c.tile_targets;
        }
    }
    out->num_candidates = count;
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    if (buf) {
        (void)hipStreamSynchronize(s->stream);
        (void)hipFree(buf);
    }
    s->last_passes = keep_passes;
    s->last_gates = keep_gates;
    if (rc) return rc;
    s->auto_mode = out->mode;
    if (out->mode == KBWH_RUN_FUSED) s->fusion_width = out->fusion_width;
    if (out->mode == KBWH_RUN_TILED || out->mode == KBWH_RUN_REORDER) s->tile_targets = out->tile_targets;
    return KBWH_OK;
}

}  // extern "C"
