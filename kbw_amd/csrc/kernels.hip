// Per-gate HBM-streaming kernels and FP64 reductions for gfx950.
//
// Every gate kernel walks only the control-satisfied subspace: a compressed
// index k is expanded by inserting the target bit (0) and the control bits
// (set to 1), so a gate with c controls on n qubits launches 2^(n-1-c) pair
// work items and touches exactly the amplitudes the reference touches
// (dense_v2.rs:48-147).  Consecutive lanes take consecutive k, so each wave
// reads and writes contiguous 16-byte amplitudes whenever the target and
// the controls sit above the low bits.
//
// The per-pair arithmetic is the reference's, operand for operand, and the
// file is compiled without FP contraction, so results equal the reference's
// values exactly (the angles' cos/sin come from the host's libm).
#include <math.h>

#include <algorithm>

#include "gate_math.h"
#include "internal.h"

#pragma clang fp contract(off)

namespace kbwh {

namespace {

constexpr int kBlock = 256;
constexpr int kPPT = 1;  // pairs (or amplitudes) per thread: one measured best on gfx950 for in-place streaming

// 16-byte amplitude loads/stores with the nontemporal hint: the state is
// streamed once per pass and is far larger than the L2 and MALL.
typedef double d2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 ldnt(const double2 *p) {
    const d2v v = __builtin_nontemporal_load((const d2v *)p);
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ void stnt(double2 *p, double2 v) {
    d2v w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, (d2v *)p);
}

__device__ __forceinline__ uint64_t expand(uint64_t k, const Insert &ins) {
    for (uint32_t j = 0; j < ins.m; ++j) {
        const uint64_t lo = k & ins.mlo[j];
        k = ((k ^ lo) << 1) | lo;
    }
    return k;
}

__device__ __forceinline__ uint64_t expand1(uint64_t k, uint64_t mlo) {
    const uint64_t lo = k & mlo;
    return ((k ^ lo) << 1) | lo;
}

// The block a workgroup of a per-gate launch takes.  Workgroups are
// dispatched in block order, round-robin over the 8 XCDs; on a power-of-two
// grid of at least 64 blocks the block index is rotated right by 3 bits, so
// each XCD walks its own eighth of the index space in order.  Measured on an
// uncontrolled H pass over 2^30 amplitudes: targets 20-23 and 25 from 0.69 to
// 0.71-0.78 of 8 TB/s, targets 3 and 6 from 0.74 to 0.78, no target lower by
// more than 0.01.  A permutation of the blocks: every amplitude is computed
// as before.
__device__ __forceinline__ uint64_t xcd_block(uint32_t b, uint32_t grid) {
    if (grid < 64u || (grid & (grid - 1u))) return b;
    const uint32_t L = 31u - (uint32_t)__builtin_clz(grid);
    return ((b >> 3) | (b << (L - 3u))) & (grid - 1u);
}

using gm::cmul;
using gm::pair_math;
enum { KX = gm::KX, KY = gm::KY, KZ = gm::KZ, KH = gm::KH, KP = gm::KP, KRX = gm::KRX, KRY = gm::KRY, KRZ = gm::KRZ };

template <int KIND>
__global__ __launch_bounds__(kBlock) void k_pair(double2 *__restrict__ a, uint64_t npairs,
                                                 uint64_t tbit, uint64_t set, Insert ins, double c,
                                                 double s) {
    const uint64_t base = xcd_block(blockIdx.x, gridDim.x) * (kBlock * kPPT) + threadIdx.x;
    uint64_t i0[kPPT];
    double2 v0[kPPT], v1[kPPT];
#pragma unroll
    for (int u = 0; u < kPPT; ++u) {
        const uint64_t k = base + (uint64_t)u * kBlock;
        i0[u] = (k < npairs) ? (expand(k, ins) | set) : ~0ull;
    }
#pragma unroll
    for (int u = 0; u < kPPT; ++u)
        if (i0[u] != ~0ull) {
            v0[u] = ldnt(&a[i0[u]]);
            v1[u] = ldnt(&a[i0[u] | tbit]);
        }
#pragma unroll
    for (int u = 0; u < kPPT; ++u)
        if (i0[u] != ~0ull) {
            pair_math<KIND>(v0[u], v1[u], c, s);
            stnt(&a[i0[u]], v0[u]);
            stnt(&a[i0[u] | tbit], v1[u]);
        }
}

// Diagonal gates: Z and P touch only the |1> amplitudes (dense_v2.rs:103-147);
// RZ multiplies both halves by their phase (dense_v2.rs:269-274).
template <int KIND>
__global__ __launch_bounds__(kBlock) void k_diag(double2 *__restrict__ a, uint64_t count,
                                                 uint64_t tbit, uint64_t set, Insert ins,
                                                 double2 ph0, double2 ph1) {
    const uint64_t base = xcd_block(blockIdx.x, gridDim.x) * (kBlock * kPPT) + threadIdx.x;
    uint64_t idx[kPPT];
    double2 v[kPPT];
#pragma unroll
    for (int u = 0; u < kPPT; ++u) {
        const uint64_t k = base + (uint64_t)u * kBlock;
        idx[u] = (k < count) ? (expand(k, ins) | set) : ~0ull;
    }
#pragma unroll
    for (int u = 0; u < kPPT; ++u)
        if (idx[u] != ~0ull) v[u] = ldnt(&a[idx[u]]);
#pragma unroll
    for (int u = 0; u < kPPT; ++u)
        if (idx[u] != ~0ull) {
            double2 r;
            if constexpr (KIND == KZ)
                r = cmul(v[u], make_double2(-1.0, -0.0));  // dense_v2.rs:191
            else if constexpr (KIND == KP)
                r = cmul(v[u], ph1);  // dense_v2.rs:216
            else
                r = cmul(v[u], (idx[u] & tbit) ? ph1 : ph0);  // dense_v2.rs:271-272
            stnt(&a[idx[u]], r);
        }
}

// Low targets (pair gates below bit 6, Z and P below bit 3; also RZ over
// more than 2^31 amplitudes, which this kernel's grid stride covers): each lane
// takes one amplitude, so every wave reads and writes 64 amplitudes of the
// control-satisfied subspace, and gets the other amplitude of its pair from
// the lane across the target's bit of the compressed index (lm; a cross-lane
// shuffle).  The pair is computed with the reference's arithmetic, and each
// lane keeps its own half, so the results are the same bits as k_pair /
// k_diag.  Controls at bit 3 and above are inserted into the index (only
// control-satisfied runs of at least 8 amplitudes, whole 128-byte lines, are
// visited); controls below bit 3 are tested per lane (lo_ctrl).  Z and P on
// bits 0..2 touch every other amplitude of each 128-byte line, so the whole
// line is read and written back either way; streaming whole lines avoids
// partial-line writes.
template <int KIND>
__global__ __launch_bounds__(kBlock) void k_low(double2 *__restrict__ a, uint64_t count, uint32_t t, uint32_t lm,
                                                uint64_t lo_ctrl, uint64_t set_hi, Insert ins_hi, double c,
                                                double s, double2 ph0, double2 ph1) {
    // count and the stride are multiples of 64, so a wave is either wholly
    // in range or wholly out (the shuffle needs all 64 lanes)
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t k = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x; k < count; k += stride) {
        const uint64_t i = expand(k, ins_hi) | set_hi;
        double2 v = ldnt(&a[i]);
        const bool one = ((i >> t) & 1u) != 0;
        const bool ok = (i & lo_ctrl) == lo_ctrl;
        if constexpr (KIND == KZ || KIND == KP) {
            if (ok && one) v = KIND == KZ ? cmul(v, make_double2(-1.0, -0.0)) : cmul(v, ph1);  // dense_v2.rs:191, 216
        } else if constexpr (KIND == KRZ) {
            if (ok) v = cmul(v, one ? ph1 : ph0);  // dense_v2.rs:271-272
        } else {
            // component-wise selects keep the pair in registers (a select of
            // whole double2 objects is lowered to a select of addresses)
            const double ox = __shfl_xor(v.x, (int)lm, 64), oy = __shfl_xor(v.y, (int)lm, 64);
            double2 a0 = make_double2(one ? ox : v.x, one ? oy : v.y);
            double2 a1 = make_double2(one ? v.x : ox, one ? v.y : oy);
            pair_math<KIND>(a0, a1, c, s);
            if (ok) v = make_double2(one ? a1.x : a0.x, one ? a1.y : a0.y);
        }
        stnt(&a[i], v);
    }
}

__global__ __launch_bounds__(kBlock) void k_init(double2 *__restrict__ a, uint64_t size) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < size; i += stride)
        a[i] = make_double2(i == 0 ? 1.0 : 0.0, 0.0);
}

// Deterministic block reduction: wave butterfly, then the 4 wave sums in order.
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double wsum[kBlock / 64];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) wsum[wave] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kBlock / 64; ++w) t += wsum[w];
    return t;
}

// measure_p1 (dense_v2.rs:279-294): sum of |a|^2 over amplitudes with bit t = 1.
__global__ __launch_bounds__(kBlock) void k_p1_partial(const double2 *__restrict__ a,
                                                       uint64_t count, uint64_t tbit, uint64_t mlo,
                                                       double *__restrict__ partial) {
    const uint64_t T = (uint64_t)gridDim.x * kBlock;
    double acc = 0.0;
    for (uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x; k < count; k += 4 * T) {
        double2 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint64_t kk = k + u * T;
            v[u] = kk < count ? ldnt(&a[expand1(kk, mlo) | tbit])
                              : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) acc += v[u].x * v[u].x + v[u].y * v[u].y;
    }
    const double t = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ __launch_bounds__(kBlock) void k_sum_final(const double *__restrict__ partial,
                                                      uint32_t m, double *__restrict__ out) {
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < m; i += kBlock) acc += partial[i];
    const double t = block_sum(acc);
    if (threadIdx.x == 0) *out = t;
}

// measure_collapse (dense_v2.rs:296-306): the kept half is scaled by p, the
// other half is zeroed (read 16 B + write 32 B per pair).
__global__ __launch_bounds__(kBlock) void k_collapse(double2 *__restrict__ a, uint64_t count,
                                                     uint64_t tbit, uint64_t mlo, int result,
                                                     double p) {
    const uint64_t base = (uint64_t)blockIdx.x * (kBlock * kPPT) + threadIdx.x;
#pragma unroll
    for (int u = 0; u < kPPT; ++u) {
        const uint64_t k = base + (uint64_t)u * kBlock;
        if (k >= count) break;
        const uint64_t i0 = expand1(k, mlo), i1 = i0 | tbit;
        const uint64_t keep = result ? i1 : i0, drop = result ? i0 : i1;
        double2 v = ldnt(&a[keep]);
        v.x = v.x * p;
        v.y = v.y * p;
        stnt(&a[keep], v);
        stnt(&a[drop], make_double2(0.0, 0.0));
    }
}

// <psi|P|psi> for one Pauli string with x = X|Y qubits, z = Z|Y qubits and
// ny Y letters: P|i> = i^ny (-1)^popcount(i & z) |i ^ x>.  With x != 0 the
// pair (i, i^x) contributes 2 Re(conj(a_{i^x}) a_i s_i).
__global__ __launch_bounds__(kBlock) void k_pauli_partial(const double2 *__restrict__ a,
                                                          uint64_t count, uint64_t xmask,
                                                          uint64_t zmask, uint32_t ny,
                                                          uint64_t hmlo,
                                                          double *__restrict__ partial) {
    const uint64_t T = (uint64_t)gridDim.x * kBlock;
    double acc = 0.0;
    if (xmask == 0) {
        for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < count; i += T) {
            const double2 v = ldnt(&a[i]);
            const double w = v.x * v.x + v.y * v.y;
            acc += (__popcll(i & zmask) & 1) ? -w : w;
        }
    } else {
        for (uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x; k < count; k += T) {
            const uint64_t i = expand1(k, hmlo), j = i ^ xmask;
            const double2 ai = ldnt(&a[i]);
            const double2 aj = ldnt(&a[j]);
            const double wr = aj.x * ai.x + aj.y * ai.y;
            const double wi = aj.x * ai.y - aj.y * ai.x;
            const uint32_t q = (ny + 2u * (uint32_t)(__popcll(i & zmask) & 1)) & 3u;
            const double r = (q == 0) ? wr : (q == 1) ? -wi : (q == 2) ? -wr : wi;
            acc += 2.0 * r;
        }
    }
    const double t = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// In-place SWAP of two physical bits: the global<->local exchange of
// dense_v2.rs:360-396 done in HBM (32 B moved per swapped pair).
__global__ __launch_bounds__(kBlock) void k_swap_bits(double2 *__restrict__ a, uint64_t count,
                                                      Insert ins, uint64_t blo, uint64_t bhi) {
    const uint64_t base = (uint64_t)blockIdx.x * (kBlock * kPPT) + threadIdx.x;
#pragma unroll
    for (int u = 0; u < kPPT; ++u) {
        const uint64_t k = base + (uint64_t)u * kBlock;
        if (k >= count) break;
        const uint64_t idx = expand(k, ins);
        const uint64_t i01 = idx | blo, i10 = idx | bhi;
        const double2 x = a[i01], y = a[i10];
        a[i01] = y;
        a[i10] = x;
    }
}

// ---- sampling: hierarchical 64-way cumulative order -------------------
__device__ __forceinline__ double weight(double2 v) {
    const double w = v.x * v.x + v.y * v.y;
    return w > 1e-30 ? w : 0.0;
}

__global__ __launch_bounds__(kBlock) void k_hier_leaf(const double2 *__restrict__ a, uint64_t cnt,
                                                      double *__restrict__ out) {
    const uint64_t j = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (j >= cnt) return;
    const double2 *p = a + j * 64;
    double acc = 0.0;
#pragma unroll 8
    for (int e = 0; e < 64; ++e) acc = acc + weight(p[e]);
    out[j] = acc;
}

__global__ __launch_bounds__(kBlock) void k_hier_level(const double *__restrict__ in, uint64_t cnt,
                                                       double *__restrict__ out) {
    const uint64_t j = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (j >= cnt) return;
    const double *p = in + j * 64;
    double acc = 0.0;
#pragma unroll 8
    for (int e = 0; e < 64; ++e) acc = acc + p[e];
    out[j] = acc;
}

__device__ __forceinline__ double level_value(const double2 *a, const Levels &lv, int l,
                                              uint64_t i) {
    return l == 0 ? weight(a[i]) : lv.lev[l][i];
}

__device__ __forceinline__ uint64_t dump_key(uint64_t state, const QubitList &ql) {
    uint64_t key = 0;
    for (uint32_t j = 0; j < ql.nq; ++j) key |= ((state >> ql.q[ql.nq - 1 - j]) & 1ull) << j;
    return key;
}

// meta[0] = total, meta[1] = rand UniformFloat scale for [0, total)
__global__ void k_hier_total(const double2 *__restrict__ a, Levels lv, double *meta) {
    const int top = lv.nlev - 1;
    double total = 0.0;
    for (uint64_t e = 0; e < lv.count[top]; ++e) total = total + level_value(a, lv, top, e);
    double scale = total - 0.0;
    if (total > 0.0) {
        const double max_rand = 1.0 - 2.220446049250313080847e-16;
        for (int it = 0; it < 64 && scale * max_rand + 0.0 >= total; ++it)
            scale = __longlong_as_double(__double_as_longlong(scale) - 1);
    }
    meta[0] = total;
    meta[1] = scale;
}

__global__ __launch_bounds__(kBlock) void k_hier_sample(const double2 *__restrict__ a, Levels lv,
                                                        const double *__restrict__ u,
                                                        uint64_t shots, QubitList ql,
                                                        const double *__restrict__ meta,
                                                        uint64_t *__restrict__ out) {
    const uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= shots) return;
    const double total = meta[0];
    if (!(total > 0.0)) {
        out[t] = 0;
        return;
    }
    const double x = u[t] * meta[1] + 0.0;
    const int top = lv.nlev - 1;
    double base = 0.0;
    uint64_t start = 0, cnt = lv.count[top], pick = 0;
    for (int l = top; l >= 0; --l) {
        // the first child whose running sum (from the parent's base) exceeds
        // x; else the last child of nonzero weight.  The parent summed its
        // children from 0, so rounding can leave x past every running sum
        // here, and a trailing child of weight 0 must not be picked.
        double acc = base, acc_nz = base;
        uint64_t e = 0, nz = cnt;
        bool hit = false;
        for (; e + 1 < cnt; ++e) {
            const double w = level_value(a, lv, l, start + e);
            const double nxt = acc + w;
            if (x < nxt) {
                hit = true;
                break;
            }
            if (w > 0.0) {
                nz = e;
                acc_nz = acc;
            }
            acc = nxt;
        }
        if (!hit && !(level_value(a, lv, l, start + cnt - 1) > 0.0) && nz < cnt) {
            e = nz;
            acc = acc_nz;
        }
        base = acc;
        pick = start + e;
        start = pick * 64;
        cnt = 64;
    }
    out[t] = dump_key(pick, ql);
}

// ---- dump: stream compaction of |a| > 1e-15 (dense_v2.rs:313) --------
__device__ __forceinline__ bool keep(double2 v) { return hypot(v.x, v.y) > 1e-15; }

__global__ __launch_bounds__(kBlock) void k_dump_count(const double2 *__restrict__ a,
                                                       uint64_t size, uint32_t *__restrict__ counts) {
    __shared__ uint32_t ws[kBlock / 64];
    const uint64_t seg = (uint64_t)blockIdx.x * kDumpSeg;
    uint32_t c = 0;
    for (uint32_t r = 0; r < kDumpSeg / kBlock; ++r) {
        const uint64_t i = seg + r * kBlock + threadIdx.x;
        if (i < size && keep(a[i])) ++c;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) c += __shfl_xor(c, off, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

__global__ __launch_bounds__(kBlock) void k_dump_write(const double2 *__restrict__ a, uint64_t size,
                                                       const uint64_t *__restrict__ offsets,
                                                       QubitList ql, uint64_t *__restrict__ keys,
                                                       double *__restrict__ re,
                                                       double *__restrict__ im) {
    __shared__ uint32_t wc[kBlock / 64];
    const uint64_t seg = (uint64_t)blockIdx.x * kDumpSeg;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint64_t out = offsets[blockIdx.x];
    for (uint32_t r = 0; r < kDumpSeg / kBlock; ++r) {
        const uint64_t i = seg + r * kBlock + threadIdx.x;
        double2 v = make_double2(0.0, 0.0);
        bool f = false;
        if (i < size) {
            v = a[i];
            f = keep(v);
        }
        const uint64_t bal = __ballot(f);
        const uint32_t before = __popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) wc[wave] = __popcll(bal);
        __syncthreads();
        uint64_t o = out;
        for (int w = 0; w < wave; ++w) o += wc[w];
        if (f) {
            keys[o + before] = dump_key(i, ql);
            re[o + before] = v.x;
            im[o + before] = v.y;
        }
        out += wc[0] + wc[1] + wc[2] + wc[3];
        __syncthreads();
    }
}

inline uint32_t grid_for(uint64_t items, uint64_t per_block) {
    return (uint32_t)((items + per_block - 1) / per_block);
}

}  // namespace

Insert make_insert(uint64_t positions_mask) {
    Insert ins{};
    ins.m = 0;
    for (uint32_t p = 0; p < 64; ++p)
        if (positions_mask >> p & 1ull) ins.mlo[ins.m++] = (p == 0) ? 0ull : ((1ull << p) - 1ull);
    return ins;
}

Coef make_coef(int kind, double angle) {
    double v[6];
    host_coef(kind, angle, v);
    Coef k;
    k.c = v[0];
    k.s = v[1];
    k.ph0 = make_double2(v[2], v[3]);
    k.ph1 = make_double2(v[4], v[5]);
    return k;
}

hipError_t launch_init(double2 *a, uint64_t size, hipStream_t st) {
    const uint32_t g = (uint32_t)std::min<uint64_t>(grid_for(size, kBlock), 65536);
    k_init<<<g, kBlock, 0, st>>>(a, size);
    return hipGetLastError();
}

hipError_t launch_gate(double2 *a, uint32_t n, int kind, double angle, uint32_t target,
                       uint64_t ctrl_mask, hipStream_t st) {
    const uint64_t tbit = 1ull << target;
    ctrl_mask &= ~tbit;  // a control equal to the target is dropped (dense_v2.rs:56-57)
    const uint32_t c = (uint32_t)__builtin_popcountll(ctrl_mask);
    const Coef k = make_coef(kind, angle);
    // RZ over more than 2^31 amplitudes goes to the grid-striding kernel below
    if (kind == KRZ && n - c <= 31) {
        const uint64_t count = 1ull << (n - c);
        const Insert ins = make_insert(ctrl_mask);
        k_diag<KRZ><<<grid_for(count, kBlock * kPPT), kBlock, 0, st>>>(a, count, tbit, ctrl_mask,
                                                                        ins, k.ph0, k.ph1);
        return hipGetLastError();
    }
    // low targets: one amplitude per lane, the pair's partner from another
    // lane (needs whole 64-amplitude segments of the compressed index: the
    // controls inserted at bit 3 and up leave at least 6 free bits)
    const uint64_t hi_ctrl = ctrl_mask & ~7ull, lo_ctrl = ctrl_mask & 7ull;
    const uint32_t c_hi = (uint32_t)__builtin_popcountll(hi_ctrl);
    const bool diag = kind == KZ || kind == KP;
    if ((kind == KRZ || target < (diag ? 3u : 6u)) && n >= 6 + c_hi) {
        const uint64_t cnt = 1ull << (n - c_hi);
        const Insert ins = make_insert(hi_ctrl);
        // the target's bit in the compressed index (inserted controls below it removed)
        const uint32_t tc = target - (uint32_t)__builtin_popcountll(hi_ctrl & (tbit - 1));
        const uint32_t lm = target < 6u ? 1u << tc : 0u;
        const uint32_t g = (uint32_t)std::min<uint64_t>(grid_for(cnt, kBlock), 1u << 22);  // 2^30 lanes at most
        switch (kind) {
        case KX: k_low<KX><<<g, kBlock, 0, st>>>(a, cnt, target, lm, lo_ctrl, hi_ctrl, ins, k.c, k.s, k.ph0, k.ph1); break;
        case KY: k_low<KY><<<g, kBlock, 0, st>>>(a, cnt, target, lm, lo_ctrl, hi_ctrl, ins, k.c, k.s, k.ph0, k.ph1); break;
        case KZ: k_low<KZ><<<g, kBlock, 0, st>>>(a, cnt, target, lm, lo_ctrl, hi_ctrl, ins, k.c, k.s, k.ph0, k.ph1); break;
        case KH: k_low<KH><<<g, kBlock, 0, st>>>(a, cnt, target, lm, lo_ctrl, hi_ctrl, ins, k.c, k.s, k.ph0, k.ph1); break;
        case KP: k_low<KP><<<g, kBlock, 0, st>>>(a, cnt, target, lm, lo_ctrl, hi_ctrl, ins, k.c, k.s, k.ph0, k.ph1); break;
        case KRX: k_low<KRX><<<g, kBlock, 0, st>>>(a, cnt, target, lm, lo_ctrl, hi_ctrl, ins, k.c, k.s, k.ph0, k.ph1); break;
        case KRY: k_low<KRY><<<g, kBlock, 0, st>>>(a, cnt, target, lm, lo_ctrl, hi_ctrl, ins, k.c, k.s, k.ph0, k.ph1); break;
        case KRZ: k_low<KRZ><<<g, kBlock, 0, st>>>(a, cnt, target, lm, lo_ctrl, hi_ctrl, ins, k.c, k.s, k.ph0, k.ph1); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    const uint64_t count = 1ull << (n - 1 - c);
    const uint32_t grid = grid_for(count, kBlock * kPPT);
    if (kind == KZ || kind == KP) {
        const Insert ins = make_insert(ctrl_mask | tbit);
        if (kind == KZ)
            k_diag<KZ><<<grid, kBlock, 0, st>>>(a, count, tbit, ctrl_mask | tbit, ins, k.ph0, k.ph1);
        else
            k_diag<KP><<<grid, kBlock, 0, st>>>(a, count, tbit, ctrl_mask | tbit, ins, k.ph0, k.ph1);
        return hipGetLastError();
    }
    const Insert ins = make_insert(ctrl_mask | tbit);
    switch (kind) {
    case KX: k_pair<KX><<<grid, kBlock, 0, st>>>(a, count, tbit, ctrl_mask, ins, k.c, k.s); break;
    case KY: k_pair<KY><<<grid, kBlock, 0, st>>>(a, count, tbit, ctrl_mask, ins, k.c, k.s); break;
    case KH: k_pair<KH><<<grid, kBlock, 0, st>>>(a, count, tbit, ctrl_mask, ins, k.c, k.s); break;
    case KRX: k_pair<KRX><<<grid, kBlock, 0, st>>>(a, count, tbit, ctrl_mask, ins, k.c, k.s); break;
    case KRY: k_pair<KRY><<<grid, kBlock, 0, st>>>(a, count, tbit, ctrl_mask, ins, k.c, k.s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

uint32_t reduce_blocks(uint64_t count) {
    // fixed per size, so the reduction tree (and the result bits) are too
    const uint64_t b = (count + kBlock * 16 - 1) / (kBlock * 16);
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, 4096));
}

hipError_t launch_p1_partial(const double2 *a, uint32_t n, uint32_t target, double *partial,
                             uint32_t *nblocks, hipStream_t st) {
    const uint64_t count = 1ull << (n - 1);
    const uint32_t g = reduce_blocks(count);
    const uint64_t tbit = 1ull << target;
    k_p1_partial<<<g, kBlock, 0, st>>>(a, count, tbit, tbit - 1, partial);
    *nblocks = g;
    return hipGetLastError();
}

hipError_t launch_sum_final(const double *partial, uint32_t m, double *out, hipStream_t st) {
    k_sum_final<<<1, kBlock, 0, st>>>(partial, m, out);
    return hipGetLastError();
}

hipError_t launch_collapse(double2 *a, uint32_t n, uint32_t target, int result, double p,
                           hipStream_t st) {
    const uint64_t count = 1ull << (n - 1);
    const uint64_t tbit = 1ull << target;
    k_collapse<<<grid_for(count, kBlock * kPPT), kBlock, 0, st>>>(a, count, tbit, tbit - 1, result, p);
    return hipGetLastError();
}

hipError_t launch_pauli_partial(const double2 *a, uint32_t n, uint64_t xmask, uint64_t zmask,
                                uint32_t ny, double *partial, uint32_t *nblocks, hipStream_t st) {
    uint64_t count = 1ull << n, hmlo = 0;
    if (xmask) {
        count >>= 1;
        const uint32_t h = (uint32_t)__builtin_ctzll(xmask);
        hmlo = (1ull << h) - 1;
    }
    const uint32_t g = reduce_blocks(count);
    k_pauli_partial<<<g, kBlock, 0, st>>>(a, count, xmask, zmask, ny & 3u, hmlo, partial);
    *nblocks = g;
    return hipGetLastError();
}

hipError_t launch_swap_bits(double2 *a, uint32_t n, uint32_t b_lo, uint32_t b_hi, hipStream_t st) {
    const uint64_t count = 1ull << (n - 2);
    const Insert ins = make_insert((1ull << b_lo) | (1ull << b_hi));
    k_swap_bits<<<grid_for(count, kBlock * kPPT), kBlock, 0, st>>>(a, count, ins, 1ull << b_lo,
                                                                   1ull << b_hi);
    return hipGetLastError();
}

hipError_t launch_hier_levels(const double2 *a, uint32_t n, double *scratch, Levels *lv,
                              hipStream_t st) {
    lv->lev[0] = nullptr;
    lv->count[0] = 1ull << n;
    lv->nlev = 1;
    double *w = scratch;
    while (lv->count[lv->nlev - 1] > 64) {
        const int l = lv->nlev;
        const uint64_t cnt = lv->count[l - 1] / 64;
        if (l == 1)
            k_hier_leaf<<<grid_for(cnt, kBlock), kBlock, 0, st>>>(a, cnt, w);
        else
            k_hier_level<<<grid_for(cnt, kBlock), kBlock, 0, st>>>(lv->lev[l - 1], cnt, w);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        lv->lev[l] = w;
        lv->count[l] = cnt;
        lv->nlev = l + 1;
        w += cnt;
    }
    return hipSuccess;
}

hipError_t launch_hier_sample(const double2 *a, const Levels &lv, const double *uniforms,
                              uint64_t shots, const QubitList &ql, uint64_t *out, double *meta,
                              hipStream_t st) {
    k_hier_total<<<1, 1, 0, st>>>(a, lv, meta);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || shots == 0) return e;
    k_hier_sample<<<grid_for(shots, kBlock), kBlock, 0, st>>>(a, lv, uniforms, shots, ql, meta, out);
    return hipGetLastError();
}

hipError_t launch_dump_count(const double2 *a, uint64_t size, uint32_t *counts, hipStream_t st) {
    k_dump_count<<<grid_for(size, kDumpSeg), kBlock, 0, st>>>(a, size, counts);
    return hipGetLastError();
}

hipError_t launch_dump_write(const double2 *a, uint64_t size, const uint64_t *offsets,
                             const QubitList &ql, uint64_t *keys, double *re, double *im,
                             hipStream_t st) {
    k_dump_write<<<grid_for(size, kDumpSeg), kBlock, 0, st>>>(a, size, offsets, ql, keys, re, im);
    return hipGetLastError();
}

}  // namespace kbwh
