// Fused k-qubit gate blocks on FP64 MFMA (KBWH_RUN_FUSED).
//
// Consecutive gates whose qubits fit in k (4 or 5) qubits form a block whose
// 2^k x 2^k unitary U is computed on the host in FP64, by applying the gates,
// with the reference's per-pair formulas (dense_v2.rs:166-277), to every
// basis vector of the block.  Consecutive blocks whose qubits fit in one
// 11-qubit tile share an HBM pass.  A workgroup stages one tile in LDS
// (2^11 amplitudes, 32 KiB) and applies each block as the real product
//     [Re out]   [Re U  -Im U] [Re V]
//     [Im out] = [Im U   Re U] [Im V]
// (2^(k+1) x 2^(k+1) times 2^(k+1) x 2^(11-k)) with v_mfma_f64_16x16x4f64:
// the 4096 outputs are 16 tiles of 16 x 16, four per wave.  Operand lanes
// (measured on MI355X): A[i][k] in lane i + 16k, B[k][j] in lane j + 16k, and
// result register t of lane l is D[l/16 + 4t][l%16].
//
// Products are accumulated in a different order than gate by gate, so the
// amplitudes agree with the reference to rounding (about 1e-16 per block),
// not bit for bit; KBWH_RUN_PER_GATE and KBWH_RUN_TILED stay exact.  Gates
// on more than k qubits (counting controls) run as single-gate passes.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <complex>
#include <vector>

#include "internal.h"
#include "state.h"
// the block kernel is written for 11-qubit tiles of 256 threads (4 waves)
#undef KBWH_TILE_BITS
#undef KBWH_TILE_TBITS
#define KBWH_TILE_BITS 11
#define KBWH_TILE_TBITS 8
#include "tile_core.h"

#pragma clang fp contract(off)

namespace kbwh {

namespace {

using namespace tile;

typedef double d4 __attribute__((ext_vector_type(4)));

// One block of a pass: the tile index of each block row (the block qubits'
// bits) and of each column (the other tile qubits' bits).
struct FBlock {
    uint16_t row_off[32];
    uint16_t col_off[128];
};
static_assert(sizeof(FBlock) == 320, "FBlock layout");

template <int K>
__global__ __launch_bounds__(TT) void k_fused(dv2 *__restrict__ a, TileSpec ts, const FBlock *__restrict__ blocks,
                                              const double *__restrict__ amat, uint32_t nb, uint32_t ntiles) {
    constexpr int R = 2 << K;   // real dimension of the block: 64 (k = 5) or 32 (k = 4)
    constexpr int KS = R / 4;   // MFMA k-steps
    constexpr int HALF = R / 2; // complex rows
    __shared__ double tl[2 << TL];  // the tile as (re, im) pairs
    const uint32_t tid = threadIdx.x;
    const uint32_t tile = blockIdx.x;
    if (tile >= ntiles) return;
    uint64_t tmask = 0;
#pragma unroll
    for (int b = 0; b < 8; ++b)
        if ((tid >> b) & 1u) tmask |= ts.qbit[b];
    const uint64_t dt = tile_base(tile, ts) | tmask;
#pragma unroll
    for (int r = 0; r < TR; ++r) {
        const dv2 x = __builtin_nontemporal_load(a + slot_offset(dt, r, ts));
        tl[2 * (tid + 256 * r)] = x.x;
        tl[2 * (tid + 256 * r) + 1] = x.y;
    }
    __syncthreads();
    const uint32_t w = tid >> 6, l = tid & 63;
    const uint32_t rt = (K == 5) ? w : (w >> 1);        // output row tile of this wave
    const uint32_t ct0 = (K == 5) ? 0u : 4u * (w & 1u);  // its first of four column tiles
    for (uint32_t bi = 0; bi < nb; ++bi) {
        const FBlock &B = blocks[bi];
        const double *A = amat + (size_t)bi * R * R + (size_t)rt * KS * 64;
        double af[KS];
#pragma unroll
        for (int s = 0; s < KS; ++s) af[s] = A[s * 64 + l];
        uint32_t coff[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) coff[c] = 2u * B.col_off[16 * (ct0 + c) + (l & 15)];
        d4 acc[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[c] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const uint32_t k = 4 * s + (l >> 4);
            const uint32_t rb = 2u * B.row_off[k % HALF] + k / HALF;
#pragma unroll
            for (int c = 0; c < 4; ++c)
                acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[s], tl[rb + coff[c]], acc[c], 0, 0, 0);
        }
        __syncthreads();  // every wave has read the block's input
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const uint32_t i = 16 * rt + (l >> 4) + 4 * t;
            const uint32_t ob = 2u * B.row_off[i % HALF] + i / HALF;
#pragma unroll
            for (int c = 0; c < 4; ++c) tl[ob + coff[c]] = acc[c][t];
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < TR; ++r)
        __builtin_nontemporal_store(mk2(tl[2 * (tid + 256 * r)], tl[2 * (tid + 256 * r) + 1]),
                                    a + slot_offset(dt, r, ts));
}

using cplx = std::complex<double>;

// One gate applied to a 2^k host vector whose bit b is qubit q[b], with the
// reference's per-pair arithmetic (dense_v2.rs:166-277).
void apply_local(std::vector<cplx> &v, int k, int kind, double angle, int tb, uint32_t cmask) {
    const Coef cf = make_coef(kind, angle);
    const size_t n = (size_t)1 << k;
    for (size_t i = 0; i < n; ++i) {
        if ((i >> tb) & 1u) continue;
        if ((i & cmask) != cmask) continue;
        const size_t j = i | ((size_t)1 << tb);
        const double r0 = v[i].real(), i0 = v[i].imag(), r1 = v[j].real(), i1 = v[j].imag();
        double o0r = r0, o0i = i0, o1r = r1, o1i = i1;
        switch (kind) {
        case 0: o0r = r1; o0i = i1; o1r = r0; o1i = i0; break;                     // X
        case 1: o0r = i1; o0i = -r1; o1r = -i0; o1i = r0; break;                   // Y
        case 2: o1r = -r1; o1i = -i1; break;                                       // Z
        case 3: {                                                                  // H
            const double h = 0.70710678118654752440;
            o0r = (r0 + r1) * h; o0i = (i0 + i1) * h; o1r = (r0 - r1) * h; o1i = (i0 - i1) * h;
            break;
        }
        case 4:                                                                    // P
            o1r = r1 * cf.ph1.x - i1 * cf.ph1.y; o1i = i1 * cf.ph1.x + r1 * cf.ph1.y;
            break;
        case 5:                                                                    // RX
            o0r = cf.c * r0 + cf.s * i1; o0i = cf.c * i0 - cf.s * r1;
            o1r = cf.s * i0 + cf.c * r1; o1i = -cf.s * r0 + cf.c * i1;
            break;
        case 6:                                                                    // RY
            o0r = cf.c * r0 - cf.s * r1; o0i = cf.c * i0 - cf.s * i1;
            o1r = cf.s * r0 + cf.c * r1; o1i = cf.s * i0 + cf.c * i1;
            break;
        default:                                                                   // RZ
            o0r = r0 * cf.ph0.x - i0 * cf.ph0.y; o0i = i0 * cf.ph0.x + r0 * cf.ph0.y;
            o1r = r1 * cf.ph1.x - i1 * cf.ph1.y; o1i = i1 * cf.ph1.x + r1 * cf.ph1.y;
            break;
        }
        v[i] = cplx(o0r, o0i);
        v[j] = cplx(o1r, o1i);
    }
}

struct HostOp {
    int kind;
    uint32_t target;
    uint64_t ctrl;  // without the target bit
    double angle;
};

struct Item {
    bool single = false;      // a gate on more than k qubits: its own per-gate pass
    HostOp op{};              // when single
    uint64_t qmask = 0;       // block qubits (exactly k of them)
    std::vector<HostOp> ops;  // the block's gates
};

// Block qubits padded to exactly k with the lowest qubits not yet used.
uint64_t pad_mask(uint64_t m, int k, uint32_t n) {
    for (uint32_t q = 0; q < n && __builtin_popcountll(m) < k; ++q) m |= 1ull << q;
    return m;
}

// 2^(k+1) x 2^(k+1) real form of the block, in the MFMA A-operand lane order:
// out[(rt * KS + s) * 64 + l] = A[16 rt + l % 16][4 s + l / 16].
void block_matrix(const Item &it, int k, double *out) {
    const int half = 1 << k, R = 2 * half, KS = R / 4;
    int qs[5], nq = 0;
    for (uint32_t q = 0; q < 64; ++q)
        if (it.qmask >> q & 1ull) qs[nq++] = (int)q;
    auto local_bit = [&](uint32_t q) {
        for (int b = 0; b < nq; ++b)
            if (qs[b] == (int)q) return b;
        return -1;
    };
    std::vector<cplx> U((size_t)half * half);
    std::vector<cplx> v(half);
    for (int col = 0; col < half; ++col) {
        std::fill(v.begin(), v.end(), cplx(0.0, 0.0));
        v[col] = cplx(1.0, 0.0);
        for (const HostOp &o : it.ops) {
            uint32_t cm = 0;
            for (uint32_t q = 0; q < 64; ++q)
                if (o.ctrl >> q & 1ull) cm |= 1u << local_bit(q);
            apply_local(v, k, o.kind, o.angle, local_bit(o.target), cm);
        }
        for (int row = 0; row < half; ++row) U[(size_t)row * half + col] = v[row];
    }
    std::vector<double> A((size_t)R * R);
    for (int i = 0; i < half; ++i)
        for (int j = 0; j < half; ++j) {
            const cplx u = U[(size_t)i * half + j];
            A[(size_t)i * R + j] = u.real();
            A[(size_t)i * R + j + half] = -u.imag();
            A[(size_t)(i + half) * R + j] = u.imag();
            A[(size_t)(i + half) * R + j + half] = u.real();
        }
    for (int rt = 0; rt < R / 16; ++rt)
        for (int s = 0; s < KS; ++s)
            for (int l = 0; l < 64; ++l)
                out[((size_t)rt * KS + s) * 64 + l] = A[(size_t)(16 * rt + l % 16) * R + 4 * s + l / 16];
}

}  // namespace

int32_t run_fused(kbwh_state *s, double2 *amp, const kbwh_op *ops, uint64_t num_ops, const double *params,
                  int k) {
    const uint32_t n = s->n;
    if (k != 4 && k != 5) return set_error(KBWH_ERR_INVALID_ARGUMENT, "fusion width must be 4 or 5");
    if (s->block) return set_error(KBWH_ERR_INVALID_ARGUMENT, "fused passes need a state without blocks");
    for (uint64_t i = 0; i < num_ops; ++i) {
        const kbwh_op &op = ops[i];
        if (op.kind > 7) return set_error(KBWH_ERR_INVALID_ARGUMENT, "unknown gate kind");
        if (op.target >= n) return set_error(KBWH_ERR_INVALID_ARGUMENT, "target out of range");
        if (n < 64 && (op.ctrl_mask >> n)) return set_error(KBWH_ERR_INVALID_ARGUMENT, "control out of range");
    }
    if (n < (uint32_t)TL) {
        // the whole state is smaller than one tile: one pass per gate
        for (uint64_t i = 0; i < num_ops; ++i) {
            const double angle = (params && ops[i].param >= 0) ? params[ops[i].param] : ops[i].angle;
            hipError_t e = launch_gate(amp, n, ops[i].kind, angle, ops[i].target,
                                       ops[i].ctrl_mask & ~(1ull << ops[i].target), s->stream);
            if (e != hipSuccess) return hip_status(e, "gate launch");
        }
        s->last_passes = num_ops;
        s->last_gates = num_ops;
        return KBWH_OK;
    }
    // 1. blocks of at most k qubits, in program order
    std::vector<Item> items;
    Item cur;
    auto close_block = [&]() {
        if (cur.ops.empty()) return;
        cur.qmask = pad_mask(cur.qmask, k, n);
        items.push_back(cur);
        cur = Item();
    };
    for (uint64_t i = 0; i < num_ops; ++i) {
        const kbwh_op &op = ops[i];
        const double angle = (params && op.param >= 0) ? params[op.param] : op.angle;
        const HostOp h{(int)op.kind, op.target, op.ctrl_mask & ~(1ull << op.target), angle};
        const uint64_t supp = h.ctrl | (1ull << h.target);
        if (__builtin_popcountll(supp) > k) {
            close_block();
            Item single;
            single.single = true;
            single.op = h;
            items.push_back(single);
            continue;
        }
        if (__builtin_popcountll(cur.qmask | supp) > k) close_block();
        cur.qmask |= supp;
        cur.ops.push_back(h);
    }
    close_block();
    // 2. passes: consecutive blocks whose qubits fit in one tile
    struct FPass {
        TileSpec ts;
        uint32_t first, count;  // blocks
    };
    std::vector<FPass> passes;
    std::vector<FBlock> fblocks;
    const int R = 2 << k;
    std::vector<double> amat;
    std::vector<std::pair<size_t, size_t>> order;  // (kind 0 = pass index, 1 = single item index)
    size_t i = 0;
    while (i < items.size()) {
        if (items[i].single) {
            order.push_back({1, i});
            ++i;
            continue;
        }
        uint64_t u = 0;
        size_t j = i;
        while (j < items.size() && !items[j].single && __builtin_popcountll(u | items[j].qmask) <= TL) u |= items[j++].qmask;
        const uint64_t qmask = pad_mask(u, TL, n);
        uint32_t q_at[TL];
        int pos_of[64];
        for (int b = 0; b < 64; ++b) pos_of[b] = -1;
        int b = 0;
        for (uint32_t q = 0; q < n; ++q)
            if (qmask >> q & 1ull) {
                pos_of[q] = b;
                q_at[b++] = q;
            }
        FPass P;
        P.ts = make_tile_spec(q_at);
        P.first = (uint32_t)fblocks.size();
        P.count = (uint32_t)(j - i);
        for (size_t m = i; m < j; ++m) {
            FBlock fb;
            memset(&fb, 0, sizeof fb);
            uint32_t bpos[5], opos[TL];
            int nb = 0, no = 0;
            for (int t = 0; t < TL; ++t) {
                if (items[m].qmask >> q_at[t] & 1ull)
                    bpos[nb++] = (uint32_t)t;
                else
                    opos[no++] = (uint32_t)t;
            }
            for (int rho = 0; rho < (1 << k); ++rho) {
                uint32_t off = 0;
                for (int x = 0; x < k; ++x)
                    if (rho >> x & 1) off |= 1u << bpos[x];
                fb.row_off[rho] = (uint16_t)off;
            }
            for (int c = 0; c < (1 << (TL - k)); ++c) {
                uint32_t off = 0;
                for (int x = 0; x < TL - k; ++x)
                    if (c >> x & 1) off |= 1u << opos[x];
                fb.col_off[c] = (uint16_t)off;
            }
            fblocks.push_back(fb);
            const size_t at = amat.size();
            amat.resize(at + (size_t)R * R);
            block_matrix(items[m], k, amat.data() + at);
        }
        order.push_back({0, passes.size()});
        passes.push_back(P);
        i = j;
    }
    // 3. upload (synchronously: the host vectors are local) and launch
    hipError_t e = hipStreamSynchronize(s->stream);
    FBlock *dblocks = nullptr;
    double *damat = nullptr;
    if (e == hipSuccess && !fblocks.empty()) e = dev_alloc(&dblocks, fblocks.size() * sizeof(FBlock));
    if (e == hipSuccess && !amat.empty()) e = dev_alloc(&damat, amat.size() * sizeof(double));
    if (e == hipSuccess && dblocks) e = hipMemcpy(dblocks, fblocks.data(), fblocks.size() * sizeof(FBlock), hipMemcpyHostToDevice);
    if (e == hipSuccess && damat) e = hipMemcpy(damat, amat.data(), amat.size() * sizeof(double), hipMemcpyHostToDevice);
    const uint32_t ntiles = (uint32_t)(1ull << (n - TL));
    s->last_passes = 0;
    for (size_t o = 0; o < order.size() && e == hipSuccess; ++o) {
        if (order[o].first == 1) {
            const HostOp &h = items[order[o].second].op;
            e = launch_gate(amp, n, h.kind, h.angle, h.target, h.ctrl, s->stream);
        } else {
            const FPass &P = passes[order[o].second];
            const FBlock *pb = dblocks + P.first;
            const double *pa = damat + (size_t)P.first * R * R;
            if (k == 5)
                k_fused<5><<<ntiles, TT, 0, s->stream>>>((dv2 *)amp, P.ts, pb, pa, P.count, ntiles);
            else
                k_fused<4><<<ntiles, TT, 0, s->stream>>>((dv2 *)amp, P.ts, pb, pa, P.count, ntiles);
            e = hipGetLastError();
        }
        ++s->last_passes;
    }
    // the block tables live until the passes have run
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (dblocks) (void)hipFree(dblocks);
    if (damat) (void)hipFree(damat);
    s->last_gates = num_ops;
    return hip_status(e, "fused passes");
}

}  // namespace kbwh
