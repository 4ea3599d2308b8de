// Adjoint differentiation of E(theta) = <psi|H|psi>, psi = U_K ... U_1 |0>.
//
// Two state vectors besides the handle's own: phi (the circuit state, run
// forward and then un-applied gate by gate) and lambda = H phi (un-applied in
// step).  For a rotation R_G(theta) = exp(-i theta G / 2) on target t with
// controls C, dU/dtheta = (-i/2) (Pi_C (x) G) U, so before un-applying gate k
//     dE/dtheta_k = 2 Re <lambda| dU_k |phi_{k-1}> = Im <lambda| Pi_C G |phi_k>,
// and for the phase gate P(l) = diag(1, e^{il}): dE/dl = -2 Im <lambda| Pi_{C,t=1} |phi_k>.
// One fused pass per parameterised gate reads the four amplitudes of each
// pair (phi and lambda), adds its term to a deterministic block reduction and
// writes both pairs back un-applied (U_k^dagger = the same gate at -theta).
// No intermediate state is stored, so the memory is three vectors at most.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "state.h"

#pragma clang fp contract(off)

namespace kbwh {

namespace {

constexpr int kB = 256;

__device__ __forceinline__ uint64_t expand(uint64_t k, const Insert &ins) {
    for (uint32_t j = 0; j < ins.m; ++j) {
        const uint64_t lo = k & ins.mlo[j];
        k = ((k ^ lo) << 1) | lo;
    }
    return k;
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.y * b.x + a.x * b.y);
}

// Im(conj(a) * b)
__device__ __forceinline__ double im_conj_mul(double2 a, double2 b) { return a.x * b.y - a.y * b.x; }

__device__ __forceinline__ double block_sum(double v) {
    __shared__ double ws[kB / 64];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kB / 64; ++w) t += ws[w];
    return t;
}

enum { KX = 0, KY = 1, KZ = 2, KH = 3, KP = 4, KRX = 5, KRY = 6, KRZ = 7 };

// lambda (+)= coeff * P phi:  (P phi)_i = s(i^x) phi_{i^x}, s(j) = i^ny (-1)^popcount(j & z)
__global__ __launch_bounds__(kB) void k_pauli_apply(const double2 *__restrict__ phi,
                                                    double2 *__restrict__ lam, uint64_t size,
                                                    uint64_t x, uint64_t z, uint32_t ny,
                                                    double coeff, int assign) {
    const uint64_t i = (uint64_t)blockIdx.x * kB + threadIdx.x;
    if (i >= size) return;
    const uint64_t j = i ^ x;
    const double2 f = phi[j];
    const uint32_t q = (ny + 2u * (uint32_t)(__popcll(j & z) & 1)) & 3u;
    double2 v;  // f * i^q
    if (q == 0) v = f;
    else if (q == 1) v = make_double2(-f.y, f.x);
    else if (q == 2) v = make_double2(-f.x, -f.y);
    else v = make_double2(f.y, -f.x);
    v.x *= coeff;
    v.y *= coeff;
    if (assign) {
        lam[i] = v;
    } else {
        double2 l = lam[i];
        l.x += v.x;
        l.y += v.y;
        lam[i] = l;
    }
}

// Re <phi|lam>
__global__ __launch_bounds__(kB) void k_dot_re(const double2 *__restrict__ a,
                                               const double2 *__restrict__ b, uint64_t size,
                                               double *__restrict__ partial) {
    const uint64_t T = (uint64_t)gridDim.x * kB;
    double acc = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * kB + threadIdx.x; i < size; i += T) {
        const double2 x = a[i], y = b[i];
        acc += x.x * y.x + x.y * y.y;
    }
    const double t = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ __launch_bounds__(kB) void k_sum_into(const double *__restrict__ partial, uint32_t m,
                                                 double *__restrict__ out) {
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < m; i += kB) acc += partial[i];
    const double t = block_sum(acc);
    if (threadIdx.x == 0) *out = t;
}

// The gradient term of one parameterised gate, then both pairs un-applied.
// (c, s) / (ph0, ph1) are the coefficients of the inverse gate.
template <int KIND>
__global__ __launch_bounds__(kB) void k_adj_param(double2 *__restrict__ phi, double2 *__restrict__ lam,
                                                  uint64_t npairs, uint64_t tbit, uint64_t set,
                                                  Insert ins, double c, double s, double2 ph0,
                                                  double2 ph1, double *__restrict__ partial) {
    const uint64_t T = (uint64_t)gridDim.x * kB;
    double acc = 0.0;
    for (uint64_t k = (uint64_t)blockIdx.x * kB + threadIdx.x; k < npairs; k += T) {
        const uint64_t i0 = expand(k, ins) | set, i1 = i0 | tbit;
        double2 f0 = phi[i0], f1 = phi[i1], l0 = lam[i0], l1 = lam[i1];
        // gradient term, with phi still at phi_k
        if constexpr (KIND == KRX) {  // Im <l|X|f>
            acc += im_conj_mul(l0, f1) + im_conj_mul(l1, f0);
        } else if constexpr (KIND == KRY) {  // Im <l|Y|f>: (Yf)_0 = -i f1, (Yf)_1 = i f0
            acc += im_conj_mul(l0, make_double2(f1.y, -f1.x)) + im_conj_mul(l1, make_double2(-f0.y, f0.x));
        } else if constexpr (KIND == KRZ) {  // Im <l|Z|f>
            acc += im_conj_mul(l0, f0) - im_conj_mul(l1, f1);
        } else {  // P: -2 Im <l|Pi_1|f>
            acc += -2.0 * im_conj_mul(l1, f1);
        }
        // un-apply (the reference's closures, dense_v2.rs:216, 229-258, 271-272)
        if constexpr (KIND == KRX) {
            const double2 a0 = f0, a1 = f1, b0 = l0, b1 = l1;
            f0 = make_double2(c * a0.x + s * a1.y, c * a0.y - s * a1.x);
            f1 = make_double2(s * a0.y + c * a1.x, -s * a0.x + c * a1.y);
            l0 = make_double2(c * b0.x + s * b1.y, c * b0.y - s * b1.x);
            l1 = make_double2(s * b0.y + c * b1.x, -s * b0.x + c * b1.y);
        } else if constexpr (KIND == KRY) {
            const double2 a0 = f0, a1 = f1, b0 = l0, b1 = l1;
            f0 = make_double2(c * a0.x - s * a1.x, c * a0.y - s * a1.y);
            f1 = make_double2(s * a0.x + c * a1.x, s * a0.y + c * a1.y);
            l0 = make_double2(c * b0.x - s * b1.x, c * b0.y - s * b1.y);
            l1 = make_double2(s * b0.x + c * b1.x, s * b0.y + c * b1.y);
        } else if constexpr (KIND == KRZ) {
            f0 = cmul(f0, ph0);
            f1 = cmul(f1, ph1);
            l0 = cmul(l0, ph0);
            l1 = cmul(l1, ph1);
        } else {
            f1 = cmul(f1, ph1);
            l1 = cmul(l1, ph1);
        }
        phi[i0] = f0;
        phi[i1] = f1;
        lam[i0] = l0;
        lam[i1] = l1;
    }
    const double t = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// |0...0> in each of nsets states of 2^n amplitudes laid out one after another
__global__ __launch_bounds__(kB) void k_init_sets(double2 *__restrict__ a, uint64_t total, uint64_t set_mask) {
    const uint64_t i = (uint64_t)blockIdx.x * kB + threadIdx.x;
    if (i >= total) return;
    a[i] = make_double2((i & set_mask) == 0 ? 1.0 : 0.0, 0.0);
}

// partial[set * gridDim.x + b] = Re <phi|lam> over chunk b of set blockIdx.y
__global__ __launch_bounds__(kB) void k_dot_sets(const double2 *__restrict__ phi, const double2 *__restrict__ lam,
                                                 uint64_t set_size, double *__restrict__ partial) {
    const uint64_t chunk = set_size / gridDim.x;
    const uint64_t first = (uint64_t)blockIdx.y * set_size + (uint64_t)blockIdx.x * chunk;
    double acc = 0.0;
    for (uint64_t i = threadIdx.x; i < chunk; i += kB) {
        const double2 x = phi[first + i], y = lam[first + i];
        acc += x.x * y.x + x.y * y.y;
    }
    const double t = block_sum(acc);
    if (threadIdx.x == 0) partial[(uint64_t)blockIdx.y * gridDim.x + blockIdx.x] = t;
}

// out[set] = the sum of the set's m partials, in order
__global__ __launch_bounds__(kB) void k_sum_sets(const double *__restrict__ partial, uint32_t m,
                                                 double *__restrict__ out) {
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < m; i += kB) acc += partial[(uint64_t)blockIdx.x * m + i];
    const double t = block_sum(acc);
    if (threadIdx.x == 0) out[blockIdx.x] = t;
}

inline uint32_t grid_for(uint64_t items, uint64_t per) { return (uint32_t)((items + per - 1) / per); }

// phi and lambda in one allocation, lambda right after phi (the tiled
// reverse sweep takes the pair as one state of n + 1 qubits)
int32_t ensure_vecs(kbwh_state *s) {
    if (s->adj_phi) return KBWH_OK;
    const int32_t rc = hip_status(dev_alloc(&s->adj_phi, 2 * s->size * sizeof(double2)), "adjoint vectors");
    if (rc) {
        s->adj_phi = nullptr;
        return rc;
    }
    s->adj_lam = s->adj_phi + s->size;
    return KBWH_OK;
}

}  // namespace

void free_adjoint(kbwh_state *s) {
    if (s->adj_phi) (void)hipFree(s->adj_phi);  // lambda shares the allocation
    if (s->adj_partial) (void)hipFree(s->adj_partial);
    kbwh_set_buffers &sb = s->sets;
    for (void *p : {sb.vecs, sb.gates, sb.tmpl, sb.params, sb.partial})
        if (p) (void)hipFree(p);
    sb = kbwh_set_buffers();
    s->adj_phi = s->adj_lam = nullptr;
    s->adj_partial = nullptr;
    s->adj_partial_cap = 0;
}

}  // namespace kbwh

using namespace kbwh;

extern "C" int32_t kbwh_adjoint_gradient(kbwh_state *s, const kbwh_op *ops, uint64_t num_ops,
                                         const double *params, uint32_t num_params,
                                         const uint8_t *paulis, const uint32_t *qubits,
                                         const uint32_t *offsets, const double *coeffs,
                                         uint32_t num_strings, double *energy, double *grad) {
    if (!s || !energy || (num_ops && !ops) || (num_params && (!params || !grad)) ||
        (num_strings && (!offsets || !coeffs)))
        return set_error(KBWH_ERR_INVALID_ARGUMENT, "null argument");
    FLUSH(s);
    if (s->block) return set_error(KBWH_ERR_INVALID_ARGUMENT, "adjoint gradient needs a flat state");
    const uint32_t n = s->n;
    const uint64_t size = s->size;
    for (uint64_t i = 0; i < num_ops; ++i) {
        const kbwh_op &op = ops[i];
        if (op.kind > KRZ || op.target >= n || (n < 64 && (op.ctrl_mask >> n)))
            return set_error(KBWH_ERR_INVALID_ARGUMENT, "bad op");
        if (op.param >= 0) {
            if ((uint32_t)op.param >= num_params)
                return set_error(KBWH_ERR_INVALID_ARGUMENT, "param index out of range");
            if (!(op.kind == KRX || op.kind == KRY || op.kind == KRZ || op.kind == KP))
                return set_error(KBWH_ERR_INVALID_ARGUMENT, "parameterised ops must be RX, RY, RZ or P");
        }
    }
    std::vector<PauliMasks> masks(num_strings);
    for (uint32_t k = 0; k < num_strings; ++k) {
        int32_t rc = pauli_masks(s, paulis, qubits, offsets[k], offsets[k + 1], false, &masks[k]);
        if (rc) return rc;
    }
    int32_t rc = ensure_vecs(s);
    if (rc) return rc;
    uint64_t nparam_ops = 0;
    for (uint64_t i = 0; i < num_ops; ++i) nparam_ops += ops[i].param >= 0;
    // partial sums: the largest reduction grid below (k_dot_re over all
    // amplitudes), or the tiled reverse sweep's per-pass partials
    const uint32_t nb = (uint32_t)std::max<size_t>(reduce_blocks(size), n >= 11 ? kAdjTiledScratch : 0);
    const size_t need = (size_t)nb + nparam_ops + 2;
    if (s->adj_partial_cap < need) {
        if (s->adj_partial) {
            hipError_t e = hipStreamSynchronize(s->stream);
            if (e == hipSuccess) e = hipFree(s->adj_partial);
            if (e != hipSuccess) return hip_status(e, "adjoint scratch");
        }
        s->adj_partial = nullptr;
        s->adj_partial_cap = 0;
        rc = hip_status(dev_alloc(&s->adj_partial, need * sizeof(double)), "adjoint scratch");
        if (rc) return rc;
        s->adj_partial_cap = need;
    }
    double *partial = s->adj_partial;
    double *slots = partial + nb;  // [0] = energy, [1 + j] = term of the j-th parameterised op
    hipStream_t st = s->stream;

    // forward: phi = U |0>, with tiled passes on the phi buffer
    hipError_t e = launch_init(s->adj_phi, size, st);
    if (e != hipSuccess) return hip_status(e, "adjoint init");
    // dependency-ordered passes (on kernels compiled per pass with kbwh_set_compiled)
    rc = run_tiled_on(s, s->adj_phi, ops, num_ops, params, true, s->compiled_passes);
    if (rc) return rc;
    s->adj_forward = n >= 11 ? s->last_passes : num_ops;
    s->adj_lambda = 0;
    s->adj_reverse = n >= 11 ? 0 : nparam_ops;
    // lambda = H phi and E = Re <phi|lambda>
    const uint32_t g1 = grid_for(size, kB);
    if (num_strings == 0) {
        e = launch_init(s->adj_lam, size, st);  // H = 0: lambda = 0 (then overwritten below)
        if (e == hipSuccess) {
            k_pauli_apply<<<g1, kB, 0, st>>>(s->adj_phi, s->adj_lam, size, 0, 0, 0, 0.0, 1);
            e = hipGetLastError();
        }
    }
    // from 11 qubits, the strings whose X/Y qubits fit one tile (qubits 0..2
    // and 8 more) in grouped sweeps, one read of phi per group; the others
    // one launch each
    std::vector<PauliMasks> narrow;
    std::vector<double> narrow_coeffs;
    std::vector<uint32_t> wide;
    for (uint32_t k = 0; k < num_strings; ++k) {
        if (n >= 11 && __builtin_popcountll(masks[k].x | 7ull) <= 11) {
            narrow.push_back(masks[k]);
            narrow_coeffs.push_back(coeffs[k]);
        } else {
            wide.push_back(k);
        }
    }
    bool assigned = false;
    if (!narrow.empty()) {
        uint32_t sweeps = 0;
        rc = pauli_apply_tiled(s, s->adj_phi, s->adj_lam, narrow, narrow_coeffs.data(), 0, 0, &sweeps);
        if (rc) return rc;
        s->adj_lambda += sweeps;
        assigned = true;
    }
    s->adj_lambda += wide.size();
    for (size_t w = 0; w < wide.size() && e == hipSuccess; ++w) {
        const uint32_t k = wide[w];
        k_pauli_apply<<<g1, kB, 0, st>>>(s->adj_phi, s->adj_lam, size, masks[k].x, masks[k].z, masks[k].ny,
                                         coeffs[k], !assigned);
        e = hipGetLastError();
        assigned = true;
    }
    if (e == hipSuccess) {
        const uint32_t gd = reduce_blocks(size);
        k_dot_re<<<gd, kB, 0, st>>>(s->adj_phi, s->adj_lam, size, partial);
        k_sum_into<<<1, kB, 0, st>>>(partial, gd, slots);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return hip_status(e, "adjoint lambda");
    // reverse sweep: tiled passes from 11 qubits on, gate by gate below
    uint64_t slot = nparam_ops;
    if (n >= 11) {
        rc = adjoint_reverse_tiled(s, ops, num_ops, params, slots + 1, partial, nb);
        if (rc) return rc;
        slot = 0;
    }
    for (uint64_t ii = n >= 11 ? 0 : num_ops; ii-- > 0 && e == hipSuccess;) {
        const kbwh_op &op = ops[ii];
        const double angle = (op.param >= 0) ? params[op.param] : op.angle;
        const uint64_t tbit = 1ull << op.target;
        const uint64_t m = op.ctrl_mask & ~tbit;
        if (op.param < 0) {
            // a fixed gate: un-apply it on both vectors
            const double inv = (op.kind == KRX || op.kind == KRY || op.kind == KRZ || op.kind == KP) ? -angle : angle;
            e = launch_gate(s->adj_phi, n, op.kind, inv, op.target, m, st);
            if (e == hipSuccess) e = launch_gate(s->adj_lam, n, op.kind, inv, op.target, m, st);
            continue;
        }
        --slot;
        const Coef k = make_coef(op.kind, -angle);
        const uint32_t c = (uint32_t)__builtin_popcountll(m);
        const uint64_t npairs = 1ull << (n - 1 - c);
        const Insert ins = make_insert(m | tbit);
        const uint32_t gb = reduce_blocks(npairs);
        switch (op.kind) {
        case KRX: k_adj_param<KRX><<<gb, kB, 0, st>>>(s->adj_phi, s->adj_lam, npairs, tbit, m, ins, k.c, k.s, k.ph0, k.ph1, partial); break;
        case KRY: k_adj_param<KRY><<<gb, kB, 0, st>>>(s->adj_phi, s->adj_lam, npairs, tbit, m, ins, k.c, k.s, k.ph0, k.ph1, partial); break;
        case KRZ: k_adj_param<KRZ><<<gb, kB, 0, st>>>(s->adj_phi, s->adj_lam, npairs, tbit, m, ins, k.c, k.s, k.ph0, k.ph1, partial); break;
        default: k_adj_param<KP><<<gb, kB, 0, st>>>(s->adj_phi, s->adj_lam, npairs, tbit, m, ins, k.c, k.s, k.ph0, k.ph1, partial); break;
        }
        k_sum_into<<<1, kB, 0, st>>>(partial, gb, slots + 1 + slot);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return hip_status(e, "adjoint reverse");
    std::vector<double> host(nparam_ops + 1);
    e = hipMemcpyAsync(host.data(), slots, host.size() * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_status(e, "adjoint read-back");
    *energy = host[0];
    for (uint32_t j = 0; j < num_params; ++j) grad[j] = 0.0;
    uint64_t j = 0;
    for (uint64_t i = 0; i < num_ops; ++i)
        if (ops[i].param >= 0) grad[ops[i].param] += host[1 + j++];
    return KBWH_OK;
}

// Parameter sets in one sweep: the nsets states lie one after another, set
// k's 2^n amplitudes at k 2^n, so together they are one state of
// n + ceil(log2 nsets) qubits whose top bits select the set.  Every gate acts
// below qubit n, so each tiled pass (8 or more tiles per set) runs over all
// sets at once with each set's own gate records; lambda = H phi is the same
// Hamiltonian on every set; energies and gradient terms are reduced per set.
extern "C" int32_t kbwh_adjoint_batch(kbwh_state *s, const kbwh_op *ops, uint64_t num_ops, const double *params,
                                      uint32_t num_params, uint32_t num_sets, const uint8_t *paulis,
                                      const uint32_t *qubits, const uint32_t *offsets, const double *coeffs,
                                      uint32_t num_strings, double *energies, double *grads) {
    if (!s || !energies || num_sets == 0 || (num_ops && !ops) || (num_params && (!params || !grads)) ||
        (num_strings && (!offsets || !coeffs)))
        return set_error(KBWH_ERR_INVALID_ARGUMENT, "null argument");
    FLUSH(s);
    if (s->block) return set_error(KBWH_ERR_INVALID_ARGUMENT, "adjoint gradient needs a flat state");
    const uint32_t n = s->n;
    if (n < 11) {
        // states smaller than a tile: one gradient per set
        for (uint32_t k = 0; k < num_sets; ++k) {
            const int32_t rc = kbwh_adjoint_gradient(s, ops, num_ops, params + (uint64_t)k * num_params, num_params,
                                                     paulis, qubits, offsets, coeffs, num_strings, energies + k,
                                                     grads ? grads + (uint64_t)k * num_params : nullptr);
            if (rc) return rc;
        }
        return KBWH_OK;
    }
    if (num_sets > 65535) return set_error(KBWH_ERR_INVALID_ARGUMENT, "at most 65535 parameter sets");
    for (uint64_t i = 0; i < num_ops; ++i) {
        const kbwh_op &op = ops[i];
        if (op.kind > KRZ || op.target >= n || (n < 64 && (op.ctrl_mask >> n)))
            return set_error(KBWH_ERR_INVALID_ARGUMENT, "bad op");
        if (op.param >= 0) {
            if ((uint32_t)op.param >= num_params)
                return set_error(KBWH_ERR_INVALID_ARGUMENT, "param index out of range");
            if (!(op.kind == KRX || op.kind == KRY || op.kind == KRZ || op.kind == KP))
                return set_error(KBWH_ERR_INVALID_ARGUMENT, "parameterised ops must be RX, RY, RZ or P");
        }
    }
    std::vector<PauliMasks> masks(num_strings);
    for (uint32_t k = 0; k < num_strings; ++k) {
        const int32_t rc = pauli_masks(s, paulis, qubits, offsets[k], offsets[k + 1], false, &masks[k]);
        if (rc) return rc;
    }
    uint32_t b = 0;
    while ((1u << b) < num_sets) ++b;
    if (n + b + 1 > 40) return set_error(KBWH_ERR_INVALID_ARGUMENT, "parameter sets too large for one sweep");
    const uint64_t set_size = s->size, span = 1ull << (n + b);
    const uint64_t total = (uint64_t)num_sets * set_size;
    uint64_t nparam_ops = 0;
    for (uint64_t i = 0; i < num_ops; ++i) nparam_ops += ops[i].param >= 0;
    kbwh_set_buffers &sb = s->sets;
    int32_t rc = grow_device(s, &sb.vecs, &sb.vecs_cap, 2 * span * sizeof(double2));
    if (rc) return rc;
    // device doubles: the parameters, the energies, the gradient terms, the energy partials
    constexpr uint32_t kEB = 16;
    const uint64_t np_all = (uint64_t)num_sets * num_params;
    const uint64_t nslots = (uint64_t)num_sets * nparam_ops;
    if ((rc = grow_device(s, &sb.params, &sb.params_cap,
                          (np_all + num_sets + nslots + (uint64_t)num_sets * kEB + 1) * sizeof(double))))
        return rc;
    double *d_params = (double *)sb.params, *d_energy = d_params + np_all, *d_slots = d_energy + num_sets;
    double *d_epart = d_slots + nslots;
    double2 *phi = (double2 *)sb.vecs, *lam = phi + span;
    hipStream_t st = s->stream;
    hipError_t e = np_all ? hipMemcpyAsync(d_params, params, np_all * sizeof(double), hipMemcpyHostToDevice, st)
                          : hipSuccess;
    if (e == hipSuccess) {
        k_init_sets<<<grid_for(total, kB), kB, 0, st>>>(phi, total, set_size - 1);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return hip_status(e, "adjoint sets init");
    // forward: phi_k = U(params_k) |0>
    if ((rc = run_sets_forward(s, phi, n, num_sets, ops, num_ops, d_params, num_params, sb))) return rc;
    s->adj_forward = s->last_passes;
    s->adj_lambda = 0;
    s->adj_reverse = 0;
    // lambda = H phi on every set, then E_k = Re <phi_k|lambda_k>
    std::vector<PauliMasks> narrow;
    std::vector<double> narrow_coeffs;
    std::vector<uint32_t> wide;
    for (uint32_t k = 0; k < num_strings; ++k) {
        if (__builtin_popcountll(masks[k].x | 7ull) <= 11) {
            narrow.push_back(masks[k]);
            narrow_coeffs.push_back(coeffs[k]);
        } else {
            wide.push_back(k);
        }
    }
    s->adj_lambda = wide.size();
    bool assigned = false;
    if (!narrow.empty()) {
        uint32_t sweeps = 0;
        rc = pauli_apply_tiled(s, phi, lam, narrow, narrow_coeffs.data(), n, (uint64_t)num_sets << (n - 11),
                               &sweeps);
        if (rc) return rc;
        s->adj_lambda += sweeps;
        assigned = true;
    }
    const uint32_t gt = grid_for(total, kB);
    if (num_strings == 0) {
        k_pauli_apply<<<gt, kB, 0, st>>>(phi, lam, total, 0, 0, 0, 0.0, 1);  // H = 0: lambda = 0
        e = hipGetLastError();
    }
    for (size_t w = 0; w < wide.size() && e == hipSuccess; ++w) {
        const uint32_t k = wide[w];
        k_pauli_apply<<<gt, kB, 0, st>>>(phi, lam, total, masks[k].x, masks[k].z, masks[k].ny, coeffs[k], !assigned);
        e = hipGetLastError();
        assigned = true;
    }
    if (e == hipSuccess) {
        k_dot_sets<<<dim3(kEB, num_sets), kB, 0, st>>>(phi, lam, set_size, d_epart);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        k_sum_sets<<<num_sets, kB, 0, st>>>(d_epart, kEB, d_energy);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return hip_status(e, "adjoint sets lambda");
    // reverse sweep over all sets
    if (nparam_ops && (rc = adjoint_reverse_sets(s, phi, n, num_sets, n + b, ops, num_ops, d_params, num_params,
                                                 d_slots)))
        return rc;
    std::vector<double> host(num_sets + nslots);
    e = hipMemcpyAsync(host.data(), d_energy, host.size() * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_status(e, "adjoint sets read-back");
    for (uint32_t k = 0; k < num_sets; ++k) {
        energies[k] = host[k];
        if (!grads) continue;
        double *g = grads + (uint64_t)k * num_params;
        for (uint32_t j = 0; j < num_params; ++j) g[j] = 0.0;
        const double *t = host.data() + num_sets + (uint64_t)k * nparam_ops;
        uint64_t j = 0;
        for (uint64_t i = 0; i < num_ops; ++i)
            if (ops[i].param >= 0) g[ops[i].param] += t[j++];
    }
    return KBWH_OK;
}

extern "C" int32_t kbwh_adjoint_stats(const kbwh_state *s, uint64_t *forward_passes, uint64_t *lambda_sweeps,
                                      uint64_t *reverse_passes) {
    if (!s || !forward_passes || !lambda_sweeps || !reverse_passes)
        return set_error(KBWH_ERR_INVALID_ARGUMENT, "null argument");
    *forward_passes = s->adj_forward;
    *lambda_sweeps = s->adj_lambda;
    *reverse_passes = s->adj_reverse;
    return KBWH_OK;
}
