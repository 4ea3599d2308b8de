// Device side of the tiled reverse sweep of the adjoint gradient: the gate
// record with its gradient slot, the partner fetch and the gradient term.
// Included after tile_dev.h built for the sweep's 12-qubit tile (8 thread
// bits, register bits 8..10 for targets, 11 for lambda's selector qubit).
// Self-contained like tile_dev.h, so the same text also compiles at run time
// into pass-specialised reverse kernels (tile_jit.hip).
#pragma once

#include "tile_dev.h"

namespace kbwh {
namespace tile {
namespace {

constexpr uint32_t kAdjSlots = 24;  // parameterised gates per pass (LDS accumulators of TT / 4 doubles each)
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;

struct AGate {
    TGate inv;       // the gate un-applied (the same gate at -angle)
    uint32_t slot;   // accumulator of the pass, or kNoSlot
    uint32_t gkind;  // the parameterised gate's kind
    uint32_t pad[2];
};

__device__ __forceinline__ double im_conj_mul(dv2 a, dv2 b) { return a.x * b.y - a.y * b.x; }

// Per-parameter term accumulators in LDS.  A term is first folded over lane
// bits 5 and 4 (two permlane row swaps and adds), so lanes 0..15 of each wave hold the
// wave's sum in 16 parts and a slot takes TT / 4 doubles (512 bytes for 256
// threads) instead of one per thread: 24 slots fit in 12 KiB instead of 48,
// and LDS no longer holds the reverse kernels to 2 or 3 workgroups per CU.
// Each owner lane keeps its own entry (zeroed, added to and read by the same
// thread, so no barrier is needed between the three).  Called where every
// lane of the workgroup takes the same path (gate records and tile bits are
// uniform).
constexpr int kAccPerSlot = TT / 4;
__device__ __forceinline__ bool adj_acc_owner(uint32_t tid) { return (tid & 48u) == 0u; }
__device__ __forceinline__ uint32_t adj_acc_index(uint32_t tid) { return ((tid >> 6) << 4) | (tid & 15u); }
__device__ __forceinline__ void adj_acc_zero(double *acc, uint32_t nslots, uint32_t tid) {
    if (adj_acc_owner(tid))
        for (uint32_t s = 0; s < nslots; ++s) acc[s * kAccPerSlot + adj_acc_index(tid)] = 0.0;
}
// t[l] + t[l ^ 32], then the same over lane bit 4, by permlane row swaps of
// two copies (the lower lane's value is the first addend in both lanes)
__device__ __forceinline__ double fold_lane_bits_45(double t) {
    double p = t, q = t;
    swap_rows<5>(p, q);
    t = p + q;
    p = t;
    q = t;
    swap_rows<4>(p, q);
    return p + q;
}
__device__ __forceinline__ void adj_acc_add(double *acc, uint32_t slot, uint32_t tid, double term) {
    term = fold_lane_bits_45(term);
    if (adj_acc_owner(tid)) acc[slot * kAccPerSlot + adj_acc_index(tid)] += term;
}
__device__ __forceinline__ double adj_acc_get(const double *acc, uint32_t slot, uint32_t tid) {
    return adj_acc_owner(tid) ? acc[slot * kAccPerSlot + adj_acc_index(tid)] : 0.0;
}

__device__ __forceinline__ double adj_block_sum(double v, double *ws) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < TT / 64; ++w) t += ws[w];
    __syncthreads();
    return t;
}

constexpr int kHalf = TR / 2;  // slots r < kHalf hold phi, r + kHalf lambda

// o[r]: the phi amplitude paired with slot r (r < kHalf) across register bit RB
template <int RB>
__device__ __forceinline__ void reg_partners(const dv2 (&v)[TR], dv2 (&o)[kHalf]) {
#pragma unroll
    for (int r = 0; r < kHalf; ++r) o[r] = v[r ^ (1 << RB)];
}

// o[r]: the phi amplitude paired with slot r across tile position tp (a
// target, so never the lambda bit).  Register bits are dispatched on their
// bits (no multi-way branch on tp), lane bits go through shuffles and wave
// bits through LDS, kChunk slots at a time.
__device__ __forceinline__ void phi_partners(const dv2 (&v)[TR], dv2 (&o)[kHalf], dv2 *lds, uint32_t tid,
                                             uint32_t tp) {
    if (tp >= (uint32_t)TB) {
        const uint32_t rb = tp - TB;
        if (rb & 2u)
            reg_partners<2>(v, o);
        else if (rb & 1u)
            reg_partners<1>(v, o);
        else
            reg_partners<0>(v, o);
        return;
    }
    if (tp < 6) {
#pragma unroll
        for (int r = 0; r < kHalf; ++r) o[r] = shfl_xor2(v[r], 1 << tp);
        return;
    }
    const uint32_t pt = tid ^ (1u << tp);
#pragma unroll
    for (int q = 0; q < kHalf; q += kChunk) {
        __syncthreads();  // earlier exchange reads are done
#pragma unroll
        for (int r = 0; r < kChunk; ++r) lds[tid + (r << TB)] = v[q + r];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kChunk; ++r) o[q + r] = lds[pt + (r << TB)];
    }
}

// ---- compile-time variants (pass-specialised reverse kernels) ----

// phi_partners with the target position TP a template argument
template <uint32_t TP>
__device__ __forceinline__ void phi_partners_at(const dv2 (&v)[TR], dv2 (&o)[kHalf], dv2 *lds, uint32_t tid) {
    if constexpr (TP >= (uint32_t)TB) {
        reg_partners<(int)(TP - TB)>(v, o);
    } else if constexpr (TP < 6) {
#pragma unroll
        for (int r = 0; r < kHalf; ++r) o[r] = shfl_xor2(v[r], 1 << TP);
    } else {
        const uint32_t pt = tid ^ (1u << TP);
#pragma unroll
        for (int q = 0; q < kHalf; q += kChunk) {
            __syncthreads();
#pragma unroll
            for (int r = 0; r < kChunk; ++r) lds[tid + (r << TB)] = v[q + r];
            __syncthreads();
#pragma unroll
            for (int r = 0; r < kChunk; ++r) o[q + r] = lds[pt + (r << TB)];
        }
    }
}

// The gradient term of a parameterised gate of kind GK (RX, RY, RZ or P) with
// target position TP and in-tile controls CM, as the generic reverse kernel
// computes it: RX / RY: Im conj(lambda_i) (G phi)_i over the control-satisfied
// amplitudes (G = X, or Y: -+i times the partner); RZ: +-Im(conj(lambda) phi)
// by the target bit; P: -2 Im(conj(lambda) phi) on the |1> half.
template <uint32_t GK, uint32_t TP, uint32_t CM>
__device__ __forceinline__ double adj_term_at(const dv2 (&v)[TR], dv2 *lds, uint32_t tid, uint64_t base,
                                              const TGate &G) {
    const bool tok = thread_ctrl_ok(tid, CM);
    constexpr bool reg = TP != kOutside && TP >= (uint32_t)TB;
    constexpr uint32_t rb = reg ? TP - TB : 0u;
    bool one_t;
    if constexpr (TP == kOutside)
        one_t = (base & G.otb) != 0;
    else if constexpr (reg)
        one_t = false;
    else
        one_t = ((tid >> TP) & 1u) != 0;
    double term = 0.0;
    if constexpr (GK == KRX || GK == KRY) {
        dv2 o[kHalf];
        phi_partners_at<TP>(v, o, lds, tid);
#pragma unroll
        for (int r = 0; r < kHalf; ++r) {
            const bool one = reg ? (((uint32_t)r >> rb) & 1u) != 0 : one_t;
            const dv2 q = GK == KRX ? o[r] : one ? mk2(-o[r].y, o[r].x) : mk2(o[r].y, -o[r].x);
            const double w = im_conj_mul(v[r + kHalf], q);
            term += slot_ctrl_ok(tok, r, CM) ? w : 0.0;
        }
    } else {
#pragma unroll
        for (int r = 0; r < kHalf; ++r) {
            const bool one = reg ? (((uint32_t)r >> rb) & 1u) != 0 : one_t;
            const double w = im_conj_mul(v[r + kHalf], v[r]);
            const double c = GK == KRZ ? (one ? -w : w) : (one ? -2.0 * w : 0.0);
            term += slot_ctrl_ok(tok, r, CM) ? c : 0.0;
        }
    }
    return term;
}

}  // namespace
}  // namespace tile
}  // namespace kbwh
