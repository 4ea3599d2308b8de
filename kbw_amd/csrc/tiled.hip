// Multi-gate tiled passes (KBWH_RUN_TILED).
//
// A pass applies a run of consecutive gates to the whole state in one HBM
// read and one HBM write.  The state is cut into tiles of 2^12 amplitudes:
// a tile is the set of indices that vary over 12 chosen qubits Q (the tile
// positions) with every other bit fixed by the tile number.  Q always holds
// qubits 0..2, so every 128-byte line of the state belongs to one tile and
// is read whole.  A workgroup of 256 threads holds one tile in registers, 16
// amplitudes per thread:
//   tile position 0..5   -> lane bits      (exchange with cross-lane shuffles)
//   tile position 6..7   -> wave bits      (exchange through LDS)
//   tile position 8..11  -> register index (no exchange)
// A non-diagonal gate (X, Y, H, RX, RY) needs its target in Q; diagonal gates
// (Z, P, RZ) and every control may sit anywhere (bits outside Q are constant
// over a tile).  Gates are applied in program order with the reference's
// per-pair arithmetic (dense_v2.rs:166-277), so a tiled run produces the same
// amplitude bits as one pass per gate; only the number of HBM passes changes.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "state.h"

#pragma clang fp contract(off)

namespace kbwh {

namespace {

constexpr int TL = 12;           // tile qubits
constexpr int TT = 256;          // threads per tile
constexpr int TR = 16;           // amplitudes per thread
constexpr uint32_t kOutside = 255;

struct TGate {
    uint32_t kind;
    uint32_t tpos;  // tile position of the target, kOutside if the target is not in Q
    uint32_t cm;    // controls inside the tile, as tile positions
    uint32_t pad;
    uint64_t ocm;   // controls outside the tile, as global bits
    uint64_t otb;   // the target's global bit (diagonal gates with target outside Q)
    double c, s;
    double ph0x, ph0y, ph1x, ph1y;
};
static_assert(sizeof(TGate) == 80, "TGate layout");

struct TileSpec {
    uint64_t qbit[TL];  // qbit[b] = 1 << Q[b]
    Insert ins;         // the positions of Q, to spread the tile number over the other bits
};

enum { KX = 0, KY = 1, KZ = 2, KH = 3, KP = 4, KRX = 5, KRY = 6, KRZ = 7 };

__device__ __forceinline__ uint64_t expand(uint64_t k, const Insert &ins) {
    for (uint32_t j = 0; j < ins.m; ++j) {
        const uint64_t lo = k & ins.mlo[j];
        k = ((k ^ lo) << 1) | lo;
    }
    return k;
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    double2 r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.y * b.x + a.x * b.y;
    return r;
}

// The pair closures of dense_v2.rs:166-263 (the non-diagonal gates).
template <int KIND>
__device__ __forceinline__ void pair_math(double2 &k0, double2 &k1, double c, double s) {
    const double2 t0 = k0, t1 = k1;
    if constexpr (KIND == KX) {
        k0 = t1;
        k1 = t0;
    } else if constexpr (KIND == KY) {
        k0 = cmul(t1, make_double2(-0.0, -1.0));
        k1 = cmul(t0, make_double2(0.0, 1.0));
    } else if constexpr (KIND == KH) {
        const double f = 0.70710678118654752440;
        k0.x = (t0.x + t1.x) * f;
        k0.y = (t0.y + t1.y) * f;
        k1.x = (t0.x - t1.x) * f;
        k1.y = (t0.y - t1.y) * f;
    } else if constexpr (KIND == KRX) {
        k0.x = c * t0.x + s * t1.y;
        k0.y = c * t0.y - s * t1.x;
        k1.x = s * t0.y + c * t1.x;
        k1.y = -s * t0.x + c * t1.y;
    } else if constexpr (KIND == KRY) {
        k0.x = c * t0.x - s * t1.x;
        k0.y = c * t0.y - s * t1.y;
        k1.x = s * t0.x + c * t1.x;
        k1.y = s * t0.y + c * t1.y;
    }
}

__device__ __forceinline__ double2 shfl_xor2(double2 v, int mask) {
    double2 r;
    r.x = __shfl_xor(v.x, mask, 64);
    r.y = __shfl_xor(v.y, mask, 64);
    return r;
}

// Target on register bit RB: both amplitudes of a pair live in one thread.
template <int KIND, int RB>
__device__ __forceinline__ void reg_pair(double2 (&v)[TR], uint32_t tid, uint32_t cm, double c,
                                         double s) {
#pragma unroll
    for (int r = 0; r < TR; ++r) {
        if ((r >> RB) & 1) continue;
        const uint32_t j0 = tid | ((uint32_t)r << 8);
        if ((j0 & cm) == cm) pair_math<KIND>(v[r], v[r | (1 << RB)], c, s);
    }
}

// Target on a thread bit: each thread computes the output of its own
// amplitude from (its value, the partner's value) with the same pair closure.
template <int KIND>
__device__ __forceinline__ void exchanged_pair(double2 (&v)[TR], const double2 (&p)[TR], uint32_t tid,
                                               uint32_t tpos, uint32_t cm, double c, double s) {
    const bool one = (tid >> tpos) & 1u;
#pragma unroll
    for (int r = 0; r < TR; ++r) {
        const uint32_t j = tid | ((uint32_t)r << 8);
        if ((j & cm) != cm) continue;
        double2 k0 = one ? p[r] : v[r];
        double2 k1 = one ? v[r] : p[r];
        pair_math<KIND>(k0, k1, c, s);
        v[r] = one ? k1 : k0;
    }
}

template <int KIND>
__device__ __forceinline__ void apply_pair(double2 (&v)[TR], double2 *lds, uint32_t tid,
                                           const TGate &G) {
    const uint32_t tp = G.tpos, cm = G.cm;
    const double c = G.c, s = G.s;
    if (tp >= 8) {
        switch (tp - 8) {
        case 0: reg_pair<KIND, 0>(v, tid, cm, c, s); break;
        case 1: reg_pair<KIND, 1>(v, tid, cm, c, s); break;
        case 2: reg_pair<KIND, 2>(v, tid, cm, c, s); break;
        default: reg_pair<KIND, 3>(v, tid, cm, c, s); break;
        }
        return;
    }
    double2 p[TR];
    if (tp < 6) {
#pragma unroll
        for (int r = 0; r < TR; ++r) p[r] = shfl_xor2(v[r], 1 << tp);
    } else {
        __syncthreads();  // earlier exchange reads are done
#pragma unroll
        for (int r = 0; r < TR; ++r) lds[tid | (r << 8)] = v[r];
        __syncthreads();
        const uint32_t pt = tid ^ (1u << tp);
#pragma unroll
        for (int r = 0; r < TR; ++r) p[r] = lds[pt | (r << 8)];
    }
    exchanged_pair<KIND>(v, p, tid, tp, cm, c, s);
}

// Z, P and RZ: each amplitude is multiplied on its own (dense_v2.rs:190-191,
// 215-216, 271-272).
template <int KIND>
__device__ __forceinline__ void apply_diag(double2 (&v)[TR], uint64_t base, uint32_t tid,
                                           const TGate &G) {
    const double2 ph0 = make_double2(G.ph0x, G.ph0y), ph1 = make_double2(G.ph1x, G.ph1y);
    const bool outside = G.tpos == kOutside;
    const bool obit = (base & G.otb) != 0;
#pragma unroll
    for (int r = 0; r < TR; ++r) {
        const uint32_t j = tid | ((uint32_t)r << 8);
        if ((j & G.cm) != G.cm) continue;
        const bool one = outside ? obit : ((j >> G.tpos) & 1u) != 0;
        if constexpr (KIND == KZ) {
            if (one) v[r] = cmul(v[r], make_double2(-1.0, -0.0));
        } else if constexpr (KIND == KP) {
            if (one) v[r] = cmul(v[r], ph1);
        } else {
            v[r] = cmul(v[r], one ? ph1 : ph0);
        }
    }
}

__global__ __launch_bounds__(TT) void k_tile(double2 *__restrict__ a, TileSpec ts,
                                             const TGate *__restrict__ gates, uint32_t ng) {
    __shared__ double2 lds[TT * TR];
    const uint32_t tid = threadIdx.x;
    const uint64_t base = expand(blockIdx.x, ts.ins);
    uint64_t dt = base;
#pragma unroll
    for (int b = 0; b < 8; ++b)
        if ((tid >> b) & 1u) dt |= ts.qbit[b];
    double2 v[TR];
    typedef double d2v __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int r = 0; r < TR; ++r) {
        uint64_t o = dt;
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if ((r >> b) & 1) o |= ts.qbit[8 + b];
        const d2v x = __builtin_nontemporal_load((const d2v *)(a + o));
        v[r] = make_double2(x.x, x.y);
    }
    for (uint32_t i = 0; i < ng; ++i) {
        const TGate G = gates[i];
        if ((base & G.ocm) != G.ocm) continue;  // a control outside the tile is 0 here
        switch (G.kind) {
        case KX: apply_pair<KX>(v, lds, tid, G); break;
        case KY: apply_pair<KY>(v, lds, tid, G); break;
        case KH: apply_pair<KH>(v, lds, tid, G); break;
        case KRX: apply_pair<KRX>(v, lds, tid, G); break;
        case KRY: apply_pair<KRY>(v, lds, tid, G); break;
        case KZ: apply_diag<KZ>(v, base, tid, G); break;
        case KP: apply_diag<KP>(v, base, tid, G); break;
        default: apply_diag<KRZ>(v, base, tid, G); break;
        }
    }
#pragma unroll
    for (int r = 0; r < TR; ++r) {
        uint64_t o = dt;
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if ((r >> b) & 1) o |= ts.qbit[8 + b];
        d2v x;
        x.x = v[r].x;
        x.y = v[r].y;
        __builtin_nontemporal_store(x, (d2v *)(a + o));
    }
}

inline bool is_diagonal(int kind) { return kind == KZ || kind == KP || kind == KRZ; }

// A gate with its physical target and controls, ready for a pass.
struct PhysOp {
    int kind;
    uint32_t target;
    uint64_t ctrl;  // physical control mask without the target bit
    double angle;
};

struct Pass {
    std::vector<PhysOp> ops;
    std::vector<uint32_t> targets;  // distinct non-diagonal targets, in first-use order
};

// Tile positions for a pass: qubits 0..2 on lane bits 0..2; the most used
// targets on register bits, then lane bits 3..5, then wave bits 6..7; free
// positions take the lowest qubits not used yet.
void assign_positions(uint32_t n, const Pass &p, int pos_of[64], uint32_t q_at[TL]) {
    for (int i = 0; i < 64; ++i) pos_of[i] = -1;
    for (int b = 0; b < TL; ++b) q_at[b] = 0xFFFFFFFFu;
    for (uint32_t q = 0; q < 3; ++q) {
        q_at[q] = q;
        pos_of[q] = (int)q;
    }
    std::vector<std::pair<int, uint32_t>> freq;
    for (uint32_t t : p.targets) {
        if (t < 3) continue;
        int f = 0;
        for (const PhysOp &o : p.ops)
            if (!is_diagonal(o.kind) && o.target == t) ++f;
        freq.push_back({-f, t});
    }
    std::stable_sort(freq.begin(), freq.end(),
                     [](const std::pair<int, uint32_t> &x, const std::pair<int, uint32_t> &y) {
                         return x.first < y.first;
                     });
    const int order[9] = {8, 9, 10, 11, 3, 4, 5, 6, 7};
    size_t k = 0;
    for (; k < freq.size(); ++k) {
        q_at[order[k]] = freq[k].second;
        pos_of[freq[k].second] = order[k];
    }
    uint32_t filler = 3;
    const int fill_order[9] = {3, 4, 5, 6, 7, 8, 9, 10, 11};
    for (int i = 0; i < 9; ++i) {
        const int b = fill_order[i];
        if (q_at[b] != 0xFFFFFFFFu) continue;
        while (filler < n && pos_of[filler] >= 0) ++filler;
        q_at[b] = filler;
        pos_of[filler] = b;
        ++filler;
    }
}

}  // namespace

int32_t run_tiled(kbwh_state *s, const kbwh_op *ops, uint64_t num_ops, const double *params) {
    return run_tiled_on(s, s->amp, ops, num_ops, params);
}

int32_t run_tiled_on(kbwh_state *s, double2 *amp, const kbwh_op *ops, uint64_t num_ops,
                     const double *params) {
    if (s->n < TL) {
        // the whole state is smaller than one tile: one pass per gate
        if (amp == s->amp) return kbwh_run(s, ops, num_ops, params, KBWH_RUN_PER_GATE);
        for (uint64_t i = 0; i < num_ops; ++i) {
            const double angle = (params && ops[i].param >= 0) ? params[ops[i].param] : ops[i].angle;
            hipError_t e = launch_gate(amp, s->n, ops[i].kind, angle, ops[i].target,
                                       ops[i].ctrl_mask & ~(1ull << ops[i].target), s->stream);
            if (e != hipSuccess) return hip_status(e, "gate launch");
        }
        return KBWH_OK;
    }
    if (s->block && amp != s->amp) return set_error(KBWH_ERR_INVALID_ARGUMENT, "Block state on a side buffer");
    const uint32_t n = s->n;
    s->last_passes = 0;
    for (uint64_t i = 0; i < num_ops; ++i) {
        const kbwh_op &op = ops[i];
        if (op.kind > KRZ) return set_error(KBWH_ERR_INVALID_ARGUMENT, "unknown gate kind");
        if (op.target >= n) return set_error(KBWH_ERR_INVALID_ARGUMENT, "target out of range");
        if (n < 64 && (op.ctrl_mask >> n)) return set_error(KBWH_ERR_INVALID_ARGUMENT, "control out of range");
    }
    Pass cur;
    // Execute: build passes lazily, flushing before any Block swap.
    std::vector<TGate> gates;
    struct Launch {
        TileSpec spec;
        uint32_t first, count;
    };
    std::vector<Launch> launches;
    auto flush = [&]() -> int32_t {
        if (cur.ops.empty()) return KBWH_OK;
        int pos_of[64];
        uint32_t q_at[TL];
        assign_positions(n, cur, pos_of, q_at);
        Launch L;
        uint64_t qmask = 0;
        for (int b = 0; b < TL; ++b) {
            L.spec.qbit[b] = 1ull << q_at[b];
            qmask |= 1ull << q_at[b];
        }
        L.spec.ins = make_insert(qmask);
        L.first = (uint32_t)gates.size();
        for (const PhysOp &o : cur.ops) {
            TGate g;
            memset(&g, 0, sizeof g);
            g.kind = (uint32_t)o.kind;
            g.tpos = pos_of[o.target] >= 0 ? (uint32_t)pos_of[o.target] : kOutside;
            g.otb = 1ull << o.target;
            for (uint32_t q = 0; q < n; ++q) {
                if (!(o.ctrl >> q & 1ull)) continue;
                if (pos_of[q] >= 0)
                    g.cm |= 1u << pos_of[q];
                else
                    g.ocm |= 1ull << q;
            }
            const Coef k = make_coef(o.kind, o.angle);
            g.c = k.c;
            g.s = k.s;
            g.ph0x = k.ph0.x;
            g.ph0y = k.ph0.y;
            g.ph1x = k.ph1.x;
            g.ph1y = k.ph1.y;
            gates.push_back(g);
        }
        L.count = (uint32_t)cur.ops.size();
        launches.push_back(L);
        cur = Pass();
        return KBWH_OK;
    };
    auto launch_all = [&]() -> int32_t {
        if (launches.empty()) return KBWH_OK;
        const size_t bytes = gates.size() * sizeof(TGate);
        if (s->tile_gates_cap < bytes) {
            hipError_t e = hipStreamSynchronize(s->stream);
            if (e == hipSuccess && s->tile_gates) e = hipFree(s->tile_gates);
            s->tile_gates = nullptr;
            s->tile_gates_cap = 0;
            if (e == hipSuccess) e = hipMalloc(&s->tile_gates, bytes);
            if (e != hipSuccess) return hip_status(e, "tile gate buffer");
            s->tile_gates_cap = bytes;
        }
        hipError_t e = hipMemcpyAsync(s->tile_gates, gates.data(), bytes, hipMemcpyHostToDevice, s->stream);
        if (e != hipSuccess) return hip_status(e, "tile gate upload");
        s->last_passes += launches.size();
        const uint32_t grid = (uint32_t)(1ull << (n - TL));
        const TGate *dg = (const TGate *)s->tile_gates;
        for (const Launch &L : launches) {
            k_tile<<<grid, TT, 0, s->stream>>>(amp, L.spec, dg + L.first, L.count);
            e = hipGetLastError();
            if (e != hipSuccess) return hip_status(e, "k_tile launch");
        }
        // the gate buffer is reused by the next run: the upload above is
        // stream-ordered, and the next upload waits for these kernels
        launches.clear();
        gates.clear();
        return KBWH_OK;
    };
    for (uint64_t i = 0; i < num_ops; ++i) {
        const kbwh_op &op = ops[i];
        const double angle = (params && op.param >= 0) ? params[op.param] : op.angle;
        uint32_t t = op.target;
        if (s->block && s->l2p[op.target] >= s->nlocal) {
            // a global target: finish the pending passes, then swap (block.rs:132-150)
            int32_t rc = flush();
            if (rc == KBWH_OK) rc = launch_all();
            if (rc) return rc;
        }
        int32_t rc = local_qubit(s, op.target, &t);
        if (rc) return rc;
        uint64_t m = physical_mask(s, op.ctrl_mask) & ~(1ull << t);
        PhysOp po{(int)op.kind, t, m, angle};
        if (!is_diagonal(op.kind)) {
            const bool have = t < 3 || std::find(cur.targets.begin(), cur.targets.end(), t) != cur.targets.end();
            if (!have) {
                if (cur.targets.size() >= (size_t)(TL - 3)) {
                    rc = flush();
                    if (rc) return rc;
                }
                cur.targets.push_back(t);
            }
        }
        cur.ops.push_back(po);
    }
    int32_t rc = flush();
    if (rc == KBWH_OK) rc = launch_all();
    s->last_gates = num_ops;
    return rc;
}

}  // namespace kbwh
