// dump, sample and exp_value entry points (include/kbwh.h).
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "state.h"

using namespace kbwh;

namespace kbwh {

int32_t pauli_masks(const kbwh_state *s, const uint8_t *paulis, const uint32_t *qubits,
                    uint32_t lo, uint32_t hi, bool physical, PauliMasks *out) {
    PauliMasks m{0, 0, 0};
    uint64_t seen = 0;
    for (uint32_t k = lo; k < hi; ++k) {
        const uint32_t q = qubits[k];
        if (q >= s->n) return set_error(KBWH_ERR_INVALID_ARGUMENT, "Pauli qubit out of range");
        if (seen >> q & 1ull)
            return set_error(KBWH_ERR_INVALID_ARGUMENT, "qubit repeated inside one Pauli string");
        seen |= 1ull << q;
        const uint64_t bit = 1ull << (physical ? s->l2p[q] : q);
        switch (paulis[k]) {
        case KBWH_PAULI_X: m.x |= bit; break;
        case KBWH_PAULI_Y: m.x |= bit; m.z |= bit; m.ny++; break;
        case KBWH_PAULI_Z: m.z |= bit; break;
        default: return set_error(KBWH_ERR_INVALID_ARGUMENT, "unknown Pauli letter");
        }
    }
    *out = m;
    return KBWH_OK;
}

}  // namespace kbwh

namespace {

int32_t physical_list(const kbwh_state *s, const uint32_t *qubits, uint32_t nq, QubitList *ql) {
    if (nq > 64) return set_error(KBWH_ERR_INVALID_ARGUMENT, "more than 64 qubits");
    ql->nq = nq;
    for (uint32_t j = 0; j < nq; ++j) {
        if (qubits[j] >= s->n) return set_error(KBWH_ERR_INVALID_ARGUMENT, "qubit out of range");
        ql->q[j] = (uint8_t)s->l2p[qubits[j]];
    }
    return KBWH_OK;
}

// Block's exp_value is the trait default, whose basis-change gates go through
// Block::hadamard / Block::phase and so through get_local_qubit
// (quantum_execution.rs:114-121, 144-151; block.rs:268-279): replay those
// lookups so the physical layout (and the LRU stamps) evolve as in the
// reference, without applying the gates.
int32_t replay_basis_touches(kbwh_state *s, const uint8_t *paulis, const uint32_t *qubits,
                             uint32_t lo, uint32_t hi) {
    for (uint32_t k = lo; k < hi; ++k) {
        const int reps = paulis[k] == KBWH_PAULI_X ? 1 : paulis[k] == KBWH_PAULI_Y ? 2 : 0;
        for (int r = 0; r < reps; ++r) {
            uint32_t t = 0;
            int32_t rc = local_qubit(s, qubits[k], &t);
            if (rc) return rc;
        }
    }
    return KBWH_OK;
}

}  // namespace

namespace {

// kbwh_exp_value after the argument checks and the queue flush.  work: device
// memory to use instead of the handle's reduction area and gate buffer
// (kbwh_exp_value_noalloc), or null.
int32_t exp_value_impl(kbwh_state *s, const uint8_t *paulis, const uint32_t *qubits, const uint32_t *offsets,
                       const double *coeffs, uint32_t ns, double *out, const ExpWork *work) {
    double *partial = work ? work->partial : s->partial;
    double *results = work ? work->results : s->result;
    if (!s->block && s->n >= 11 && ns > 0) {
        // grouped read-only sweeps, several strings per read of the state; a
        // string with more than 8 X/Y qubits above qubit 2 is summed alone
        std::vector<PauliMasks> grouped, alone;
        std::vector<double> gc, ac;
        for (uint32_t k = 0; k < ns; ++k) {
            PauliMasks m;
            int32_t rc = pauli_masks(s, paulis, qubits, offsets[k], offsets[k + 1], true, &m);
            if (rc) return rc;
            if (__builtin_popcountll(m.x | 7ull) <= 11) {
                grouped.push_back(m);
                gc.push_back(coeffs[k]);
            } else {
                alone.push_back(m);
                ac.push_back(coeffs[k]);
            }
        }
        double total = 0.0;
        if (!grouped.empty()) {
            int32_t rc = exp_value_tiled(s, grouped, gc.data(), &total, work);
            if (rc) return rc;
        }
        for (size_t k = 0; k < alone.size(); ++k) {
            uint32_t nb = 0;
            hipError_t e = launch_pauli_partial(s->amp, s->n, alone[k].x, alone[k].z, alone[k].ny, partial, &nb,
                                                s->stream);
            if (e == hipSuccess) e = launch_sum_final(partial, nb, results, s->stream);
            if (e != hipSuccess) return hip_status(e, "exp_value");
            double v = 0.0;
            int32_t rc = reduce_from(s, results, &v, 1);
            if (rc) return rc;
            total += v * ac[k];
        }
        *out = total;
        return KBWH_OK;
    }
    std::vector<double> values(ns, 0.0);
    uint32_t pending = 0, first = 0;
    for (uint32_t k = 0; k < ns; ++k) {
        const uint32_t lo = offsets[k], hi = offsets[k + 1];
        int32_t rc = KBWH_OK;
        if (s->block && (rc = replay_basis_touches(s, paulis, qubits, lo, hi))) return rc;
        PauliMasks m;
        if ((rc = pauli_masks(s, paulis, qubits, lo, hi, true, &m))) return rc;
        uint32_t nb = 0;
        hipError_t e = launch_pauli_partial(s->amp, s->n, m.x, m.z, m.ny, partial, &nb, s->stream);
        if (e == hipSuccess) e = launch_sum_final(partial, nb, results + pending, s->stream);
        if (e != hipSuccess) return hip_status(e, "exp_value");
        if (s->block && (rc = replay_basis_touches(s, paulis, qubits, lo, hi))) return rc;
        if (++pending == kResultSlots || k + 1 == ns) {
            if ((rc = reduce_from(s, results, values.data() + first, pending))) return rc;
            first += pending;
            pending = 0;
        }
    }
    // quantum_execution.rs:155-157: sum of result * coefficient, in order
    double total = 0.0;
    for (uint32_t k = 0; k < ns; ++k) total += values[k] * coeffs[k];
    *out = total;
    return KBWH_OK;
}

int32_t exp_value_checks(kbwh_state *s, const uint8_t *paulis, const uint32_t *qubits, const uint32_t *offsets,
                         const double *coeffs, uint32_t ns, double *out) {
    if (!s || !out || (ns && (!offsets || !coeffs)))
        return set_error(KBWH_ERR_INVALID_ARGUMENT, "null argument");
    if (ns && offsets[ns] > offsets[0] && (!paulis || !qubits))
        return set_error(KBWH_ERR_INVALID_ARGUMENT, "null Pauli arrays");
    for (uint32_t k = 0; k < ns; ++k)
        if (offsets[k + 1] < offsets[k])
            return set_error(KBWH_ERR_INVALID_ARGUMENT, "offsets not ascending");
    return KBWH_OK;
}

// ---- the allocation-free variant's work area ------------------------------
// One fixed array per GPU, held until the process ends: block partials, then
// per-group / per-string results, then the string table of at most
// kNoAllocStrings records.  Every kernel writes each slot it later reads
// within the same call, so nothing needs clearing between calls.
constexpr uint32_t kNoAllocStrings = 4096;
constexpr size_t kNoAllocDoubles = kExpPartials + kResultSlots + kNoAllocStrings * kExpStringBytes / sizeof(double);
static __device__ __align__(256) double scratch[kNoAllocDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices];

// The device address of `scratch` on `device` (each GPU holds its own copy),
// looked up once per device with that device current.  The array's name is
// only a host-side stand-in; kernels get this address.
int32_t scratch_on(int device, double **out) {
    if (device < 0 || device >= kMaxDevices) return set_error(KBWH_ERR_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e == hipSuccess && cur != device) e = hipSetDevice(device);
        void *p = nullptr;
        if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) {
            const hipError_t r = hipSetDevice(cur);
            if (e == hipSuccess) e = r;
        }
        if (e != hipSuccess) return hip_status(e, "hipGetSymbolAddress(scratch)");
        if (!p) return set_error(KBWH_ERR_DEVICE, "no device address for the exp_value work area");
        g_scratch_ptr[device] = (double *)p;
    }
    *out = g_scratch_ptr[device];
    return KBWH_OK;
}

}  // namespace

extern "C" {

int32_t kbwh_exp_value(kbwh_state *s, const uint8_t *paulis, const uint32_t *qubits,
                       const uint32_t *offsets, const double *coeffs, uint32_t ns, double *out) {
    int32_t rc = exp_value_checks(s, paulis, qubits, offsets, coeffs, ns, out);
    if (rc) return rc;
    FLUSH(s);
    return exp_value_impl(s, paulis, qubits, offsets, coeffs, ns, out, nullptr);
}

int32_t kbwh_exp_value_noalloc(kbwh_state *s, const uint8_t *paulis, const uint32_t *qubits,
                               const uint32_t *offsets, const double *coeffs, uint32_t ns, double *out) {
    int32_t rc = exp_value_checks(s, paulis, qubits, offsets, coeffs, ns, out);
    if (rc) return rc;
    if (ns > kNoAllocStrings) return set_error(KBWH_ERR_CAPACITY, "kbwh_exp_value_noalloc takes at most 4096 strings");
    // flushing queued gates may grow the gate buffer
    if (queue_pending(s))
        return set_error(KBWH_ERR_INVALID_ARGUMENT, "queued gates pending: call kbwh_synchronize before kbwh_exp_value_noalloc");
    double *base = nullptr;
    if ((rc = scratch_on(s->device, &base))) return rc;
    ExpWork w;
    w.partial = base;
    w.results = base + kExpPartials;
    w.strings = base + kExpPartials + kResultSlots;
    w.strings_cap = kNoAllocStrings;
    return exp_value_impl(s, paulis, qubits, offsets, coeffs, ns, out, &w);
}

int32_t kbwh_dump(kbwh_state *s, const uint32_t *qubits, uint32_t nq, uint64_t capacity,
                  uint64_t *count, uint64_t *keys, double *re, double *im) {
    if (!s || !count || (nq && !qubits)) return set_error(KBWH_ERR_INVALID_ARGUMENT, "null argument");
    if (keys && (!re || !im)) return set_error(KBWH_ERR_INVALID_ARGUMENT, "null output arrays");
    FLUSH(s);
    QubitList ql;
    int32_t rc = physical_list(s, qubits, nq, &ql);
    if (rc) return rc;
    const uint64_t nseg = (s->size + kDumpSeg - 1) / kDumpSeg;
    // scratch: per-segment counts (u32) then offsets (u64)
    if ((rc = ensure_scratch(s, nseg + nseg / 2 + 2))) return rc;
    uint32_t *d_counts = (uint32_t *)s->scratch;
    uint64_t *d_offsets = (uint64_t *)(s->scratch + (nseg + 1) / 2 + 1);
    hipError_t e = launch_dump_count(s->amp, s->size, d_counts, s->stream);
    std::vector<uint32_t> counts(nseg);
    if (e == hipSuccess)
        e = hipMemcpyAsync(counts.data(), d_counts, nseg * 4, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return hip_status(e, "dump count");
    std::vector<uint64_t> offsets(nseg);
    uint64_t total = 0;
    for (uint64_t i = 0; i < nseg; ++i) {
        offsets[i] = total;
        total += counts[i];
    }
    *count = total;
    if (!keys) return KBWH_OK;
    if (total > capacity) return set_error(KBWH_ERR_CAPACITY, "dump capacity too small");
    if (total == 0) return KBWH_OK;
    void *d_out = nullptr;
    e = hipMemcpyAsync(d_offsets, offsets.data(), nseg * 8, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = dev_alloc(&d_out, total * 24);
    if (e != hipSuccess) return hip_status(e, "dump alloc");
    uint64_t *d_keys = (uint64_t *)d_out;
    double *d_re = (double *)(d_keys + total), *d_im = d_re + total;
    e = launch_dump_write(s->amp, s->size, d_offsets, ql, d_keys, d_re, d_im, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(keys, d_keys, total * 8, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(re, d_re, total * 8, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(im, d_im, total * 8, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    hipError_t f = hipFree(d_out);
    if (e == hipSuccess) e = f;
    return hip_status(e, "dump write");
}

int32_t kbwh_sample(kbwh_state *s, const uint32_t *qubits, uint32_t nq, uint64_t shots,
                    const double *uniforms, uint64_t *outcomes) {
    if (!s || !qubits || (shots && (!uniforms || !outcomes)))
        return set_error(KBWH_ERR_INVALID_ARGUMENT, "null argument");
    FLUSH(s);
    if (nq == 0)  // quantum_execution.rs:295 asserts a non-empty qubit list
        return set_error(KBWH_ERR_INVALID_ARGUMENT, "sample needs at least one qubit");
    QubitList ql;
    int32_t rc = physical_list(s, qubits, nq, &ql);
    if (rc) return rc;
    uint64_t level_doubles = 0;
    for (uint64_t c = s->size; c > 64; c /= 64) level_doubles += c / 64;
    const uint64_t chunk = std::min<uint64_t>(std::max<uint64_t>(shots, 1), 1ull << 22);
    if ((rc = ensure_scratch(s, level_doubles + 2 * chunk + 8))) return rc;
    double *lev = s->scratch;
    double *d_u = lev + level_doubles;
    uint64_t *d_out = (uint64_t *)(d_u + chunk);
    Levels lv;
    hipError_t e = launch_hier_levels(s->amp, s->n, lev, &lv, s->stream);
    if (e != hipSuccess) return hip_status(e, "sample levels");
    double meta[2] = {0.0, 0.0};
    for (uint64_t done = 0; done < shots || done == 0; done += chunk) {
        const uint64_t m = std::min(chunk, shots - done);
        if (m) e = hipMemcpyAsync(d_u, uniforms + done, m * 8, hipMemcpyHostToDevice, s->stream);
        if (e == hipSuccess) e = launch_hier_sample(s->amp, lv, d_u, m, ql, d_out, s->result, s->stream);
        if (e == hipSuccess && m)
            e = hipMemcpyAsync(outcomes + done, d_out, m * 8, hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(meta, s->result, 16, hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) return hip_status(e, "sample");
        if (!(meta[0] > 0.0)) return set_error(KBWH_ERR_ZERO_WEIGHTS, "sample weights are all zero");
        if (shots == 0) break;
    }
    return KBWH_OK;
}

}  // extern "C"
