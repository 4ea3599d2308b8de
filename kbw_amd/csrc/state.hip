// Host side of the engine: the state handle, the Block logical/physical
// mapping, the measure loop and the C ABI declared in include/kbwh.h.
//
// One handle owns one device buffer of 2^n complex f64 amplitudes (16 GiB at
// 30 qubits, 128 GiB at 33: a single HBM3E allocation on one MI355X) and one
// HIP stream.  Block mode keeps the same single buffer: the reference's
// 2^G blocks of 2^L local amplitudes (block.rs:89-101) are the top G bits of
// the physical index, a global control is a high control bit and a global
// target is brought down by an in-HBM bit swap.
#include <stdlib.h>
#include <string.h>

#include <atomic>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/kbwh.h"
#include "internal.h"
#include "state.h"

using namespace kbwh;

namespace {

thread_local std::string g_err;

int32_t fail(int32_t code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(e_ == hipErrorOutOfMemory ? KBWH_ERR_OUT_OF_MEMORY : KBWH_ERR_DEVICE, \
                        std::string(#expr) + ": " + hipGetErrorString(e_));              \
    } while (0)

const char kBuildInfo[] = "kbwh v0.1.0 [hip gfx950, dense c128]";

uint32_t env_block_size() {
    const char *v = getenv("KBW_BLOCK_SIZE");
    if (!v || !*v) return 20;
    char *end = nullptr;
    unsigned long x = strtoul(v, &end, 10);
    if (end == v || *end) return 20;  // block.rs:224-227: unparsable -> 20
    return (uint32_t)x;
}

}  // namespace

namespace kbwh {

int32_t set_error(int32_t code, const std::string &msg) { return fail(code, msg); }

// injected allocation failures: allocations left before the next one fails
// (negative: none pending)
static std::atomic<int64_t> g_fail_after{-1};

void inject_alloc_failure(int64_t after) { g_fail_after.store(after < 0 ? -1 : after); }

hipError_t dev_alloc_raw(void **p, size_t bytes) {
    int64_t left = g_fail_after.load();
    while (left >= 0) {
        if (g_fail_after.compare_exchange_weak(left, left - 1)) {
            if (left == 0) {
                *p = nullptr;
                return hipErrorOutOfMemory;
            }
            break;
        }
    }
    const hipError_t e = hipMalloc(p, bytes);
    // the failure is returned here; clear it from the runtime's last-error
    // slot, or the next launch check (hipGetLastError) would report it again
    if (e != hipSuccess) (void)hipGetLastError();
    return e;
}

int32_t hip_status(hipError_t e, const char *what) {
    if (e == hipSuccess) return KBWH_OK;
    return fail(e == hipErrorOutOfMemory ? KBWH_ERR_OUT_OF_MEMORY : KBWH_ERR_DEVICE,
                std::string(what) + ": " + hipGetErrorString(e));
}

int32_t grow_device(kbwh_state *s, void **buf, size_t *cap, size_t bytes) {
    if (*cap >= bytes) return KBWH_OK;
    if (*buf) {
        HIPCHK(hipStreamSynchronize(s->stream));
        HIPCHK(hipFree(*buf));
    }
    *buf = nullptr;
    *cap = 0;
    HIPCHK(dev_alloc(buf, bytes));
    *cap = bytes;
    return KBWH_OK;
}

int32_t ensure_scratch(kbwh_state *s, size_t doubles) {
    if (s->scratch_cap >= doubles) return KBWH_OK;
    if (s->scratch) HIPCHK(hipFree(s->scratch));
    s->scratch = nullptr;
    s->scratch_cap = 0;
    HIPCHK(dev_alloc(&s->scratch, doubles * sizeof(double)));
    s->scratch_cap = doubles;
    return KBWH_OK;
}

// block.rs:104-109
static void touch(kbwh_state *s, uint32_t phys) {
    if (phys < s->nlocal) {
        s->counter += 1;
        s->last_access[phys] = s->counter;
    }
}

// block.rs:123-156: a global target is swapped with the least recently used
// local qubit (first minimum, block.rs:111-117).  Only the maps change here;
// *swap_local / *swap_global name the physical bits whose amplitudes the
// caller must exchange (kNoSwap when the target is local).
void map_qubit(kbwh_state *s, uint32_t logical, uint32_t *out, uint32_t *swap_local, uint32_t *swap_global) {
    *swap_local = *swap_global = kNoSwap;
    if (!s->block) {
        *out = logical;
        return;
    }
    const uint32_t phys = s->l2p[logical];
    if (phys < s->nlocal) {
        touch(s, phys);
        *out = phys;
        return;
    }
    uint32_t lq = 0;
    for (uint32_t i = 1; i < s->nlocal; ++i)
        if (s->last_access[i] < s->last_access[lq]) lq = i;
    s->swaps++;
    const uint32_t old_logical = s->p2l[lq];
    s->l2p[logical] = lq;
    s->p2l[lq] = logical;
    s->l2p[old_logical] = phys;
    s->p2l[phys] = old_logical;
    touch(s, lq);
    *out = lq;
    *swap_local = lq;
    *swap_global = phys;
}

int32_t local_qubit(kbwh_state *s, uint32_t logical, uint32_t *out) {
    uint32_t lq, g;
    map_qubit(s, logical, out, &lq, &g);
    if (lq != kNoSwap) HIPCHK(launch_swap_bits(s->amp, s->n, lq, g, s->stream));
    return KBWH_OK;
}

uint64_t physical_mask(const kbwh_state *s, uint64_t logical_mask) {
    if (!s->block) return logical_mask;
    uint64_t m = 0;
    for (uint32_t q = 0; q < s->n; ++q)
        if (logical_mask >> q & 1ull) m |= 1ull << s->l2p[q];
    return m;
}

int32_t apply_gate(kbwh_state *s, int kind, double angle, uint32_t target, uint64_t ctrl_mask) {
    if (kind < KBWH_X || kind > KBWH_RZ) return fail(KBWH_ERR_INVALID_ARGUMENT, "unknown gate kind");
    if (target >= s->n) return fail(KBWH_ERR_INVALID_ARGUMENT, "target out of range");
    if (s->n < 64 && (ctrl_mask >> s->n)) return fail(KBWH_ERR_INVALID_ARGUMENT, "control out of range");
    uint32_t t = 0;
    int32_t rc = local_qubit(s, target, &t);
    if (rc) return rc;
    const uint64_t m = physical_mask(s, ctrl_mask);
    HIPCHK(launch_gate(s->amp, s->n, kind, angle, t, m, s->stream));
    return KBWH_OK;
}

int32_t reduce_from(kbwh_state *s, const double *src, double *out, uint32_t count) {
    if (count > kResultSlots) return fail(KBWH_ERR_CAPACITY, "more results than the host buffer holds");
    HIPCHK(hipMemcpyAsync(s->host_result, src, count * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    memcpy(out, s->host_result, count * sizeof(double));
    return KBWH_OK;
}

int32_t reduce_to_host(kbwh_state *s, double *out, uint32_t count) { return reduce_from(s, s->result, out, count); }

int32_t measure_p1(kbwh_state *s, uint32_t target, double *p1) {
    if (target >= s->n) return fail(KBWH_ERR_INVALID_ARGUMENT, "target out of range");
    uint32_t t = 0;
    int32_t rc = local_qubit(s, target, &t);  // block.rs:302-309
    if (rc) return rc;
    uint32_t nb = 0;
    HIPCHK(launch_p1_partial(s->amp, s->n, t, s->partial, &nb, s->stream));
    HIPCHK(launch_sum_final(s->partial, nb, s->result, s->stream));
    return reduce_to_host(s, p1, 1);
}

int32_t measure_collapse(kbwh_state *s, uint32_t target, int result, double p) {
    if (target >= s->n) return fail(KBWH_ERR_INVALID_ARGUMENT, "target out of range");
    uint32_t t = 0;
    int32_t rc = local_qubit(s, target, &t);  // block.rs:311-316
    if (rc) return rc;
    HIPCHK(launch_collapse(s->amp, s->n, t, result ? 1 : 0, p, s->stream));
    return KBWH_OK;
}

// rand's UniformFloat::new(0, total) scale: decreased one ulp at a time while
// (1 - 2^-52) * scale + 0 >= total.
double uniform_scale(double total) {
    const double max_rand = 1.0 - 2.220446049250313080847e-16;
    double scale = total - 0.0;
    for (int it = 0; it < 64 && scale * max_rand + 0.0 >= total; ++it) {
        uint64_t b;
        memcpy(&b, &scale, 8);
        b -= 1;
        memcpy(&scale, &b, 8);
    }
    return scale;
}

}  // namespace kbwh

// ============================================================================
// C ABI
// ============================================================================
extern "C" {

int32_t kbwh_new(uint32_t num_qubits, uint32_t flags, uint32_t block_size, kbwh_state **out) {
    if (!out) return fail(KBWH_ERR_INVALID_ARGUMENT, "out is null");
    *out = nullptr;
    const bool block = (flags & KBWH_FLAG_BLOCK) != 0;
    uint32_t nlocal = num_qubits;
    if (block) {
        if (num_qubits <= 1) return fail(KBWH_ERR_EXECUTION_FAILED, "Block needs at least 2 qubits");
        uint32_t bs = block_size ? block_size : env_block_size();
        nlocal = std::min(num_qubits - 1, bs);
        if (nlocal > 32) return fail(KBWH_ERR_EXECUTION_FAILED, "more than 32 local qubits");
        if (num_qubits > 40) return fail(KBWH_ERR_EXECUTION_FAILED, "more than 40 qubits");
    } else if (num_qubits > 32) {
        return fail(KBWH_ERR_EXECUTION_FAILED, "dense state limited to 32 qubits");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(KBWH_ERR_NO_DEVICE, "no HIP device visible");
    kbwh_state *s = new kbwh_state();
    s->n = num_qubits;
    s->block = block;
    s->block_size_arg = block_size;
    s->use_queue = (flags & KBWH_FLAG_QUEUE) != 0;
    s->nlocal = nlocal;
    s->size = 1ull << num_qubits;
    hipError_t e = hipGetDevice(&s->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = dev_alloc(&s->amp, s->size * sizeof(double2));
    if (e == hipSuccess) e = dev_alloc(&s->partial, (4096 + 64) * sizeof(double));
    if (e == hipSuccess) e = dev_alloc(&s->result, kResultSlots * sizeof(double));
    if (e == hipSuccess) e = hipHostMalloc(&s->host_result, kResultSlots * sizeof(double), 0);
    if (e == hipSuccess) e = launch_init(s->amp, s->size, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
        int32_t rc = hip_status(e, "kbwh_new");
        kbwh_free(s);
        return rc;
    }
    s->l2p.resize(num_qubits);
    s->p2l.resize(num_qubits);
    for (uint32_t i = 0; i < num_qubits; ++i) s->l2p[i] = s->p2l[i] = i;
    s->last_access.assign(block ? nlocal : 0, 0);
    *out = s;
    return KBWH_OK;
}

int32_t kbwh_free(kbwh_state *s) {
    if (!s) return KBWH_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->amp) (void)hipFree(s->amp);
    if (s->partial) (void)hipFree(s->partial);
    if (s->result) (void)hipFree(s->result);
    if (s->scratch) (void)hipFree(s->scratch);
    if (s->tile_gates) (void)hipFree(s->tile_gates);
    if (s->host_result) (void)hipHostFree(s->host_result);
    free_adjoint(s);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return KBWH_OK;
}

int32_t kbwh_clear(kbwh_state *s) {
    if (!s) return fail(KBWH_ERR_INVALID_ARGUMENT, "null handle");
    s->queue.clear();
    s->block_queue.clear();
    HIPCHK(launch_init(s->amp, s->size, s->stream));
    for (uint32_t i = 0; i < s->n; ++i) s->l2p[i] = s->p2l[i] = i;
    std::fill(s->last_access.begin(), s->last_access.end(), 0);
    s->counter = 0;
    return KBWH_OK;
}

int32_t kbwh_num_qubits(const kbwh_state *s, uint32_t *n) {
    if (!s || !n) return fail(KBWH_ERR_INVALID_ARGUMENT, "null argument");
    *n = s->n;
    return KBWH_OK;
}

int32_t kbwh_gate(kbwh_state *s, int32_t kind, double angle, uint32_t target,
                  const uint32_t *control, uint32_t num_control) {
    if (!s || (num_control && !control)) return fail(KBWH_ERR_INVALID_ARGUMENT, "null argument");
    uint64_t mask = 0;
    for (uint32_t i = 0; i < num_control; ++i) {
        if (control[i] >= s->n) return fail(KBWH_ERR_INVALID_ARGUMENT, "control out of range");
        mask |= 1ull << control[i];
    }
    return s->use_queue ? queue_gate(s, kind, angle, target, mask) : apply_gate(s, kind, angle, target, mask);
}

int32_t kbwh_gate_mask(kbwh_state *s, int32_t kind, double angle, uint32_t target,
                       uint64_t ctrl_mask) {
    if (!s) return fail(KBWH_ERR_INVALID_ARGUMENT, "null handle");
    return s->use_queue ? queue_gate(s, kind, angle, target, ctrl_mask) : apply_gate(s, kind, angle, target, ctrl_mask);
}

int32_t kbwh_inject_alloc_failure(int64_t after) {
    kbwh::inject_alloc_failure(after);
    return KBWH_OK;
}

int32_t kbwh_run(kbwh_state *s, const kbwh_op *ops, uint64_t num_ops, const double *params,
                 uint32_t mode) {
    if (!s || (num_ops && !ops)) return fail(KBWH_ERR_INVALID_ARGUMENT, "null argument");
    FLUSH(s);
    if (mode == KBWH_RUN_AUTO) mode = s->auto_mode;
    if (mode == KBWH_RUN_TILED) return run_tiled(s, ops, num_ops, params);
    if (mode == KBWH_RUN_REORDER) return run_tiled_on(s, s->amp, ops, num_ops, params, true);
    if (mode == KBWH_RUN_COMPILED) return run_tiled_on(s, s->amp, ops, num_ops, params, true, true);
    if (mode == KBWH_RUN_FUSED) return run_fused(s, s->amp, ops, num_ops, params, (int)s->fusion_width);
    if (mode != KBWH_RUN_PER_GATE) return fail(KBWH_ERR_INVALID_ARGUMENT, "unknown run mode");
    s->last_passes = num_ops;
    s->last_gates = num_ops;
    for (uint64_t i = 0; i < num_ops; ++i) {
        const double angle = (params && ops[i].param >= 0) ? params[ops[i].param] : ops[i].angle;
        int32_t rc = apply_gate(s, ops[i].kind, angle, ops[i].target, ops[i].ctrl_mask);
        if (rc) return rc;
    }
    return KBWH_OK;
}

int32_t kbwh_measure_p1(kbwh_state *s, uint32_t target, double *p1) {
    if (!s || !p1) return fail(KBWH_ERR_INVALID_ARGUMENT, "null argument");
    FLUSH(s);
    return measure_p1(s, target, p1);
}

int32_t kbwh_measure_collapse(kbwh_state *s, uint32_t target, int32_t result, double p) {
    if (!s) return fail(KBWH_ERR_INVALID_ARGUMENT, "null handle");
    if (s->block && s->use_queue) {
        // block.rs:311-316: get_local_qubit (a global target flushes the
        // queue and swaps), then every block collapses; the gates given after
        // the last swap stay queued and are applied after the collapse, as in
        // the reference
        if (target >= s->n) return fail(KBWH_ERR_INVALID_ARGUMENT, "target out of range");
        uint32_t t = 0, sl = kNoSwap, sg = kNoSwap;
        map_qubit(s, target, &t, &sl, &sg);
        int32_t rc;
        if (sl != kNoSwap) {
            s->block_queue.push_back(kbwh_phys_event{KBWH_X, 0, 0, 0.0, sl, sg});
            rc = flush_queue(s);
        } else {
            // the reference's queue holds only the gates after its last swap
            rc = flush_through_last_swap(s);
        }
        if (rc) return rc;
        HIPCHK(launch_collapse(s->amp, s->n, t, result ? 1 : 0, p, s->stream));
        return KBWH_OK;
    }
    FLUSH(s);
    return measure_collapse(s, target, result, p);
}

// quantum_execution.rs:272-292
int32_t kbwh_measure(kbwh_state *s, const uint32_t *qubits, uint32_t nq, const double *uniforms,
                     uint64_t *result) {
    if (!s || !result || (nq && (!qubits || !uniforms)))
        return fail(KBWH_ERR_INVALID_ARGUMENT, "null argument");
    FLUSH(s);
    if (nq > 64) return fail(KBWH_ERR_INVALID_ARGUMENT, "more than 64 qubits");
    uint64_t out = 0;
    for (uint32_t idx = 0; idx < nq; ++idx) {
        const uint32_t q = qubits[nq - 1 - idx];
        double p1 = 0.0;
        int32_t rc = measure_p1(s, q, &p1);
        if (rc) return rc;
        const double p0 = std::max(1.0 - p1, 0.0);
        const double total = p0 + p1;
        if (!(total > 0.0) || p1 < 0.0)
            return fail(KBWH_ERR_ZERO_WEIGHTS, "measurement weights are all zero");
        const double x = uniforms[idx] * uniform_scale(total) + 0.0;
        const int r = (p0 <= x) ? 1 : 0;
        const double prob = r ? p1 : p0;
        const double p = 1.0 / sqrt(std::max(prob, 2.2250738585072014e-308));
        rc = measure_collapse(s, q, r, p);
        if (rc) return rc;
        out |= (uint64_t)r << idx;
    }
    *result = out;
    return KBWH_OK;
}

int32_t kbwh_read_amplitudes(kbwh_state *s, uint64_t first, uint64_t count, double *out) {
    if (!s || (count && !out)) return fail(KBWH_ERR_INVALID_ARGUMENT, "null argument");
    if (first > s->size || count > s->size - first)
        return fail(KBWH_ERR_INVALID_ARGUMENT, "range out of bounds");
    FLUSH(s);
    HIPCHK(hipMemcpyAsync(out, s->amp + first, count * sizeof(double2), hipMemcpyDeviceToHost,
                          s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return KBWH_OK;
}

int32_t kbwh_write_amplitudes(kbwh_state *s, uint64_t first, uint64_t count, const double *in) {
    if (!s || (count && !in)) return fail(KBWH_ERR_INVALID_ARGUMENT, "null argument");
    if (first > s->size || count > s->size - first)
        return fail(KBWH_ERR_INVALID_ARGUMENT, "range out of bounds");
    FLUSH(s);
    HIPCHK(hipMemcpyAsync(s->amp + first, in, count * sizeof(double2), hipMemcpyHostToDevice,
                          s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return KBWH_OK;
}

int32_t kbwh_qubit_map(const kbwh_state *s, uint32_t *l2p) {
    if (!s || !l2p) return fail(KBWH_ERR_INVALID_ARGUMENT, "null argument");
    memcpy(l2p, s->l2p.data(), s->n * sizeof(uint32_t));
    return KBWH_OK;
}

int32_t kbwh_swap_count(const kbwh_state *s, uint64_t *swaps) {
    if (!s || !swaps) return fail(KBWH_ERR_INVALID_ARGUMENT, "null argument");
    *swaps = s->swaps;
    return KBWH_OK;
}

int32_t kbwh_run_stats(const kbwh_state *s, uint64_t *gates, uint64_t *passes) {
    if (!s || !gates || !passes) return fail(KBWH_ERR_INVALID_ARGUMENT, "null argument");
    *gates = s->last_gates;
    *passes = s->last_passes;
    return KBWH_OK;
}

int32_t kbwh_synchronize(kbwh_state *s) {
    if (!s) return fail(KBWH_ERR_INVALID_ARGUMENT, "null handle");
    FLUSH(s);
    HIPCHK(hipStreamSynchronize(s->stream));
    return KBWH_OK;
}

int32_t kbwh_stream(const kbwh_state *s, void **stream) {
    if (!s || !stream) return fail(KBWH_ERR_INVALID_ARGUMENT, "null argument");
    *stream = (void *)s->stream;
    return KBWH_OK;
}

int32_t kbwh_device_buffer(const kbwh_state *s, void **ptr, uint64_t *bytes) {
    if (!s || !ptr || !bytes) return fail(KBWH_ERR_INVALID_ARGUMENT, "null argument");
    *ptr = (void *)s->amp;
    *bytes = s->size * sizeof(double2);
    return KBWH_OK;
}

int32_t kbwh_build_info(const uint8_t **msg, uintptr_t *size) {
    if (!msg || !size) return fail(KBWH_ERR_INVALID_ARGUMENT, "null argument");
    *msg = (const uint8_t *)kBuildInfo;
    *size = sizeof(kBuildInfo) - 1;
    return KBWH_OK;
}

const char *kbwh_last_error(void) { return g_err.c_str(); }

}  // extern "C"
