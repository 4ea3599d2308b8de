// Tiled reverse sweep of the adjoint gradient (states of 11 or more qubits).
//
// The reverse sweep un-applies the gates from last to first on phi and lambda
// and, before un-applying a parameterised gate k, adds its term
//     RX, RY, RZ:  Im <lambda| Pi_C G |phi>,   G = X, Y, Z on the target
//     P:           -2 Im <lambda| Pi_{C,t=1} |phi>
// (see adjoint.hip).  phi and lambda lie back to back in one buffer, which is
// taken as a state of n + 1 qubits whose top qubit n selects lambda.  Gates
// of the reversed list share tiled passes over that state, planned in
// dependency order as in KBWH_RUN_REORDER (a term is unchanged when both
// vectors first go through gates on other qubits, which commute with G),
// with qubit n pinned to the last register bit of a 12-bit tile: a thread
// holds amplitude i of phi in slot r and amplitude i of lambda in slot r + 8,
// every pass reads and writes the two vectors once, and a gate un-applied on
// both vectors is one uncontrolled gate of the pass.  For RX and RY the term
// is Im sum_{C(i)} conj(lambda_i) (G phi)_i, where (G phi)_i is the phi
// amplitude paired with i across the target (times -+i for Y), fetched like
// a gate's partner (registers, lane shuffles or LDS).  Each term is folded
// over lane bits 4 and 5 and added to the lane's LDS entry for that
// parameterised gate (16 entries per wave, adj_acc_add in adj_dev.h); a
// workgroup reduces its entries at the end and the workgroup sums are added
// in a fixed order.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "internal.h"
#include "state.h"
// a 12-bit tile: 8 thread bits, register bits 8..10 for targets, 11 for qubit n
#undef KBWH_TILE_BITS
#undef KBWH_TILE_TBITS
#define KBWH_TILE_BITS 12
#define KBWH_TILE_TBITS 8
#define KBWH_TILE_GATE_RBITS 3
// contracted (products fused into the following adds), as the compiled
// reverse kernels are: the gradient equals the program-order adjoint to
// rounding either way
#define KBWH_TILE_CONTRACT 1
#include "tile_core.h"
#include "adj_dev.h"

#pragma clang fp contract(off)

namespace kbwh {
namespace {

using namespace tile;

constexpr uint32_t kMaxSlots = kAdjSlots;
static_assert(kAdjSlots == kAdjTiledSlots, "adjoint slots per pass");
constexpr int kTargets = TL - 4;  // free positions for targets (qubit n takes one)

// the parameter index (among the parameterised ops) of each slot of a pass
struct SlotMap {
    uint32_t global[kMaxSlots];
};

// v: the tile of the state with lambda's selector qubit, phi in slots
// r < kHalf and lambda in r + kHalf.  Parameter sets: workgroup w of set
// (blockIdx = set * wps + w) walks tiles set * tps + w, + wps, ... with the
// set's own gate records gs + set * ng (one set: wps = gridDim.x).
__global__ __launch_bounds__(TT) void k_adj_tile(dv2 *__restrict__ a, TileSpec ts, const AGate *__restrict__ gs,
                                                 uint32_t ng, uint32_t tps, uint32_t wps, uint32_t nslots,
                                                 double *__restrict__ partial) {
    __shared__ dv2 lds[TT * kChunk];
    __shared__ double acc[kMaxSlots * kAccPerSlot];
    __shared__ double ws[TT / 64];
    const uint32_t tid = threadIdx.x;
    const uint32_t set = blockIdx.x / wps, wk = blockIdx.x % wps;
    gs += (size_t)set * ng;
    adj_acc_zero(acc, nslots, tid);
    uint64_t tmask = 0;
#pragma unroll
    for (int b = 0; b < TB; ++b)
        if ((tid >> b) & 1u) tmask |= ts.qbit[b];
    for (uint32_t tl = wk; tl < tps; tl += wps) {
        const uint32_t t = set * tps + tl;
        const uint64_t base = tile_base(t, ts);
        const uint64_t dt = base | tmask;
        dv2 v[TR];
#pragma unroll
        for (int r = 0; r < TR; ++r) v[r] = a[slot_offset(dt, r, ts)];
        for (uint32_t g = 0; g < ng; ++g) {
            const AGate &A = gs[g];
            const TGate G = A.inv;
            if ((base & G.ocm) != G.ocm) continue;  // a control outside the tile is 0 here
            if (A.slot != kNoSlot) {
                const bool tok = thread_ctrl_ok(tid, G.cm);
                const uint32_t tp = G.tpos;
                // the target bit of slot r: fixed per tile (outside), per thread (lane/wave) or per slot
                const bool reg = tp - (uint32_t)TB < 8u;
                const uint32_t rb = (tp - TB) & 7u;
                const bool one_t = tp == kOutside ? (base & G.otb) != 0 : ((tid >> (tp & 31u)) & 1u) != 0;
                double term = 0.0;
                if (A.gkind == KRX || A.gkind == KRY) {
                    // Im conj(lambda_i) (G phi)_i, G = X or Y: (X phi)_i = phi_p, (Y phi)_i = -+i phi_p
                    dv2 o[kHalf];
                    phi_partners(v, o, lds, tid, tp);
                    const bool y = A.gkind == KRY;
#pragma unroll
                    for (int r = 0; r < kHalf; ++r) {
                        const bool one = reg ? (((uint32_t)r >> rb) & 1u) != 0 : one_t;
                        const dv2 q = !y ? o[r] : one ? mk2(-o[r].y, o[r].x) : mk2(o[r].y, -o[r].x);
                        const double w = im_conj_mul(v[r + kHalf], q);
                        term += slot_ctrl_ok(tok, r, G.cm) ? w : 0.0;
                    }
                } else {
                    // RZ: +-Im(conj(l) f) by the target bit; P: -2 Im(conj(l) f) on the |1> half
                    const bool rz = A.gkind == KRZ;
#pragma unroll
                    for (int r = 0; r < kHalf; ++r) {
                        const bool one = reg ? (((uint32_t)r >> rb) & 1u) != 0 : one_t;
                        const double w = im_conj_mul(v[r + kHalf], v[r]);
                        const double c = rz ? (one ? -w : w) : (one ? -2.0 * w : 0.0);
                        term += slot_ctrl_ok(tok, r, G.cm) ? c : 0.0;
                    }
                }
                adj_acc_add(acc, A.slot, tid, term);
            }
            if (G.kind == KZ || G.kind == KP || G.kind == KRZ)
                apply_diag(v, base, tid, G);
            else
                apply_pair(v, lds, tid, G);
        }
#pragma unroll
        for (int r = 0; r < TR; ++r) a[slot_offset(dt, r, ts)] = v[r];
    }
    for (uint32_t s = 0; s < nslots; ++s) {
        const double v = adj_block_sum(adj_acc_get(acc, s, tid), ws);
        if (tid == 0) partial[(size_t)s * gridDim.x + blockIdx.x] = v;
    }
}

// out[set * ostride + map.global[s]] = the sum, in order, of slot s's
// partials from the wps workgroups of set (blockIdx.x = s, blockIdx.y = set)
__global__ __launch_bounds__(256) void k_sum_slots(const double *__restrict__ partial, uint32_t grid, uint32_t wps,
                                                  uint64_t ostride, double *__restrict__ out, SlotMap map) {
    __shared__ double ws[4];
    const uint32_t s = blockIdx.x, set = blockIdx.y;
    const double *p = partial + (size_t)s * grid + (size_t)set * wps;
    double a = 0.0;
    // a loop with the same trip count in every lane (units built with
    // uniform regions unstructurized keep per-lane loop exits out)
    for (uint32_t i0 = 0; i0 < wps; i0 += 256) {
        const uint32_t i = i0 + threadIdx.x;
        if (i < wps) a += p[i];
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) a += __shfl_xor(a, off, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) out[set * ostride + map.global[s]] = ((ws[0] + ws[1]) + ws[2]) + ws[3];
}

// AGate records of S parameter sets from one set's templates (as
// k_expand_sets of the forward passes): the inverse gate's coefficients from
// -params[set][pidx[r]] when pidx[r] >= 0
__global__ __launch_bounds__(256) void k_expand_asets(const AGate *__restrict__ tmpl, const int32_t *__restrict__ pidx,
                                                      const uint32_t *__restrict__ dst,
                                                      const uint32_t *__restrict__ stride, uint32_t nrec,
                                                      uint32_t nsets, const double *__restrict__ params,
                                                      uint32_t nparams, AGate *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (uint64_t)nrec * nsets) return;
    const uint32_t set = (uint32_t)(i / nrec), r = (uint32_t)(i % nrec);
    AGate g = tmpl[r];
    if (pidx[r] >= 0) set_coefficients_dev(g.inv, -params[(uint64_t)set * nparams + (uint32_t)pidx[r]]);
    out[dst[r] + (uint64_t)set * stride[r]] = g;
}

// Tile positions of a pass: qubits 0..2 on lane bits 0..2, lambda's selector
// qubit (n for one state) on register bit 11, the most used targets on register bits 8..10, then lane
// bits 3..5, then wave bits 6..7; free positions take the lowest unused qubits.
void assign_adj(uint32_t n, uint32_t lam_qubit, const Pass &p, int pos_of[64], uint32_t q_at[TL]) {
    for (int i = 0; i < 64; ++i) pos_of[i] = -1;
    for (int b = 0; b < TL; ++b) q_at[b] = 0xFFFFFFFFu;
    for (uint32_t q = 0; q < 3; ++q) {
        q_at[q] = q;
        pos_of[q] = (int)q;
    }
    q_at[TL - 1] = lam_qubit;
    pos_of[lam_qubit] = TL - 1;
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
    const int order[kTargets] = {TB, TB + 1, TB + 2, 3, 4, 5, 6, 7};
    for (size_t k = 0; k < freq.size(); ++k) {
        q_at[order[k]] = freq[k].second;
        pos_of[freq[k].second] = order[k];
    }
    uint32_t filler = 3;
    for (int b = 3; b < TL - 1; ++b) {
        if (q_at[b] != 0xFFFFFFFFu) continue;
        while (pos_of[filler] >= 0) ++filler;
        q_at[b] = filler;
        pos_of[filler] = b;
        ++filler;
    }
}

}  // namespace

namespace {

// The reverse sweep's passes for one set: gate tables and slot maps (the
// coefficients of parameterised gates are those of `params`, or left for the
// device when params is null).
struct RPass {
    TileSpec ts;
    uint32_t first, count, nslots;
    SlotMap map;
};

int32_t plan_reverse(uint32_t n, uint32_t lam_qubit, const kbwh_op *ops, uint64_t num_ops, const double *params,
                     std::vector<RPass> &passes, std::vector<AGate> &table, std::vector<int32_t> &pidx) {
    uint64_t nparam = 0;
    for (uint64_t i = 0; i < num_ops; ++i) nparam += ops[i].param >= 0;
    // the reversed list (each gate at -angle for the rotations); rev[j] is
    // ops[num_ops - 1 - j], param_of[j] its index among the parameterised ops
    std::vector<PhysOp> rev(num_ops);
    std::vector<uint8_t> is_param(num_ops);
    std::vector<uint32_t> param_of(num_ops, kNoSlot);
    std::vector<int32_t> param_idx(num_ops, -1);
    {
        uint64_t global = nparam;
        for (uint64_t j = 0; j < num_ops; ++j) {
            const kbwh_op &op = ops[num_ops - 1 - j];
            const double angle = (params && op.param >= 0) ? params[op.param] : (op.param >= 0 ? 0.0 : op.angle);
            const bool rot = op.kind == KRX || op.kind == KRY || op.kind == KRZ || op.kind == KP;
            rev[j] = PhysOp{(int)op.kind, op.target, op.ctrl_mask & ~(1ull << op.target), rot ? -angle : angle};
            is_param[j] = op.param >= 0;
            param_idx[j] = op.param;
            if (is_param[j]) param_of[j] = (uint32_t)--global;
        }
    }
    // dependency-ordered passes: a gradient term is Im <lambda| G |phi>
    // with G on the gate's own qubits, unchanged when both vectors first go
    // through gates on other qubits (they commute with G), so the reversed
    // list may be scheduled like KBWH_RUN_REORDER, kMaxSlots terms per pass
    std::vector<uint32_t> order, ends;
    schedule_passes(rev, n, kTargets, order, ends, &is_param, kMaxSlots);
    if (order.size() != num_ops) return set_error(KBWH_ERR_DEVICE, "adjoint schedule incomplete");
    size_t k = 0;
    for (const uint32_t end : ends) {
        Pass cur;
        for (size_t i = k; i < end; ++i) add_op(cur, rev[order[i]]);
        int pos_of[64];
        uint32_t q_at[TL];
        assign_adj(n, lam_qubit, cur, pos_of, q_at);
        RPass P;
        memset(&P.map, 0, sizeof P.map);
        P.ts = make_tile_spec(q_at);
        P.first = (uint32_t)table.size();
        P.count = (uint32_t)(end - k);
        P.nslots = 0;
        for (size_t i = k; i < end; ++i) {
            const uint32_t j = order[i];
            AGate A;
            memset(&A, 0, sizeof A);
            A.inv = make_tgate(rev[j], pos_of, n);
            A.gkind = (uint32_t)rev[j].kind;
            A.slot = kNoSlot;
            if (is_param[j]) {
                A.slot = P.nslots;
                P.map.global[P.nslots++] = param_of[j];
            }
            table.push_back(A);
            pidx.push_back(param_idx[j]);
        }
        if (P.nslots > kMaxSlots) return set_error(KBWH_ERR_DEVICE, "adjoint pass over its slots");
        passes.push_back(P);
        k = end;
    }
    return KBWH_OK;
}

// the passes' compiled reverse kernels (null where a pass keeps k_adj_tile)
int32_t compiled_reverse(kbwh_state *s, const std::vector<RPass> &passes, const std::vector<AGate> &table,
                         std::vector<void *> &fns) {
    static_assert(sizeof(AGate) == 160, "AGate layout shared with tile_jit.hip");
    std::vector<JitPass> lst;
    for (const RPass &P : passes) lst.push_back(JitPass{&table[P.first], P.count, kJitAdjoint});
    return jit_prepare(s->device, lst, fns);
}

// Workgroups of a reverse kernel resident on the device at once (occupancy
// per CU times the CU count), at most kAdjTiledGrid.  A reverse kernel walks
// its tiles with a fixed grid; a grid above what is resident runs its last
// workgroups in a second, partly idle round (their tiles are uniform work):
// kernels that fit 3 workgroups per CU (768) took 16 rounds of tiles per
// pass at a grid of 1024 instead of 11.  Cached per kernel and device; fn
// null is the generic k_adj_tile.
uint32_t adj_grid_cap(void *fn, int device) {
    static std::mutex mu;
    static std::map<std::pair<void *, int>, uint32_t> cache;
    std::lock_guard<std::mutex> lk(mu);
    const auto key = std::make_pair(fn, device);
    const auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int per_cu = 0, cus = 0;
    hipError_t e = fn ? hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (hipFunction_t)fn, TT, 0)
                      : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_adj_tile, TT, 0);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    uint32_t cap = kAdjTiledGrid;
    if (e == hipSuccess && per_cu > 0 && cus > 0)
        cap = (uint32_t)std::min<uint64_t>(kAdjTiledGrid, (uint64_t)per_cu * (uint64_t)cus);
    else
        (void)hipGetLastError();  // a failed query leaves the default grid
    cache[key] = cap;
    return cap;
}

}  // namespace

// The reverse sweep over phi and lambda (s->adj_phi, lambda right after it)
// in tiled passes; slots[j] receives the term of the j-th parameterised op
// (forward order).  Needs n >= 11.
int32_t adjoint_reverse_tiled(kbwh_state *s, const kbwh_op *ops, uint64_t num_ops, const double *params,
                              double *slots, double *partial, size_t partial_cap) {
    const uint32_t n = s->n;
    if (n < 11 || n > 62) return set_error(KBWH_ERR_INVALID_ARGUMENT, "tiled reverse sweep needs 11..62 qubits");
    std::vector<RPass> passes;
    std::vector<AGate> table;
    std::vector<int32_t> pidx;
    int32_t prc = plan_reverse(n, n, ops, num_ops, params, passes, table, pidx);
    if (prc) return prc;
    const uint32_t ntiles = (uint32_t)(1ull << (n - 11));
    const uint32_t grid = std::min<uint32_t>(ntiles, kAdjTiledGrid);
    if ((size_t)grid * kMaxSlots > partial_cap) return set_error(KBWH_ERR_DEVICE, "adjoint scratch too small");
    const size_t bytes = table.size() * sizeof(AGate);
    if (s->tile_gates_cap < bytes) {
        hipError_t e = hipStreamSynchronize(s->stream);
        if (e == hipSuccess && s->tile_gates) e = hipFree(s->tile_gates);
        s->tile_gates = nullptr;
        s->tile_gates_cap = 0;
        if (e == hipSuccess) e = dev_alloc(&s->tile_gates, bytes);
        if (e != hipSuccess) return hip_status(e, "adjoint gate table");
        s->tile_gates_cap = bytes;
    }
    hipError_t e = bytes ? hipMemcpyAsync(s->tile_gates, table.data(), bytes, hipMemcpyHostToDevice, s->stream)
                         : hipSuccess;
    const AGate *dg = (const AGate *)s->tile_gates;
    std::vector<void *> fns;
    if (s->compiled_passes && e == hipSuccess) {
        if ((prc = compiled_reverse(s, passes, table, fns))) {
            (void)hipStreamSynchronize(s->stream);
            return prc;
        }
    }
    for (size_t i = 0; i < passes.size(); ++i) {
        const RPass &P = passes[i];
        if (e != hipSuccess) break;
        void *fn = fns.empty() ? nullptr : fns[i];
        const uint32_t g = std::min(grid, adj_grid_cap(fn, s->device));
        if (fn) {
            e = jit_launch_adj(fn, s->adj_phi, &P.ts, sizeof P.ts, dg + P.first, ntiles, g, partial, g, s->stream);
        } else {
            k_adj_tile<<<g, TT, 0, s->stream>>>((dv2 *)s->adj_phi, P.ts, dg + P.first, P.count, ntiles, g, P.nslots,
                                                partial);
            e = hipGetLastError();
        }
        if (e == hipSuccess && P.nslots) {
            k_sum_slots<<<dim3(P.nslots, 1), 256, 0, s->stream>>>(partial, g, g, 0, slots, P.map);
            e = hipGetLastError();
        }
    }
    s->adj_reverse = passes.size();
    // the host gate table is local: the upload must be done before it goes away
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    return hip_status(e, "adjoint reverse passes");
}

int32_t adjoint_reverse_sets(kbwh_state *s, double2 *phi, uint32_t n, uint32_t nsets, uint32_t lam_qubit,
                             const kbwh_op *ops, uint64_t num_ops, const double *params_dev, uint32_t num_params,
                             double *slots) {
    if (n < 11 || n > 40 || lam_qubit < n || lam_qubit > 62)
        return set_error(KBWH_ERR_INVALID_ARGUMENT, "tiled reverse sweep over sets needs 11..40 qubits");
    std::vector<RPass> passes;
    std::vector<AGate> table;
    std::vector<int32_t> pidx;
    int32_t rc = plan_reverse(n, lam_qubit, ops, num_ops, nullptr, passes, table, pidx);
    if (rc) return rc;
    uint64_t nparam = 0;
    for (uint64_t i = 0; i < num_ops; ++i) nparam += ops[i].param >= 0;
    const size_t nrec = table.size();
    if (nrec == 0) return KBWH_OK;
    // wps workgroups per set share the set's tiles (about kAdjTiledGrid in all)
    const uint32_t tps = 1u << (n - 11);
    const uint32_t wps = std::max<uint32_t>(1, std::min<uint32_t>(tps, kAdjTiledGrid / nsets));
    const uint64_t grid = (uint64_t)nsets * wps;
    if (grid > 0x7FFFFFFFull || (uint64_t)nrec * nsets > 0xFFFFFFFFull)
        return set_error(KBWH_ERR_INVALID_ARGUMENT, "too many parameter sets");
    kbwh_set_buffers &sb = s->sets;
    if ((rc = grow_device(s, &sb.partial, &sb.partial_cap, grid * kMaxSlots * sizeof(double)))) return rc;
    if ((rc = grow_device(s, &sb.tmpl, &sb.tmpl_cap, nrec * (sizeof(AGate) + 12)))) return rc;
    if ((rc = grow_device(s, &sb.gates, &sb.gates_cap, nrec * nsets * sizeof(AGate)))) return rc;
    std::vector<uint32_t> dst(nrec), stride(nrec);
    for (const RPass &P : passes)
        for (uint32_t i = 0; i < P.count; ++i) {
            dst[P.first + i] = (uint32_t)((uint64_t)nsets * P.first + i);
            stride[P.first + i] = P.count;
        }
    char *tb = (char *)sb.tmpl;
    AGate *d_tmpl = (AGate *)tb;
    int32_t *d_pidx = (int32_t *)(tb + nrec * sizeof(AGate));
    uint32_t *d_dst = (uint32_t *)(d_pidx + nrec), *d_stride = d_dst + nrec;
    hipStream_t st = s->stream;
    hipError_t e = hipMemcpyAsync(d_tmpl, table.data(), nrec * sizeof(AGate), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_pidx, pidx.data(), nrec * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_dst, dst.data(), nrec * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_stride, stride.data(), nrec * 4, hipMemcpyHostToDevice, st);
    AGate *dg = (AGate *)sb.gates;
    if (e == hipSuccess) {
        const uint64_t total = (uint64_t)nrec * nsets;
        k_expand_asets<<<(uint32_t)((total + 255) / 256), 256, 0, st>>>(d_tmpl, d_pidx, d_dst, d_stride,
                                                                        (uint32_t)nrec, nsets, params_dev,
                                                                        num_params, dg);
        e = hipGetLastError();
    }
    double *partial = (double *)sb.partial;
    std::vector<void *> fns;
    if (s->compiled_passes && e == hipSuccess) {
        if ((rc = compiled_reverse(s, passes, table, fns))) {
            (void)hipStreamSynchronize(st);
            return rc;
        }
    }
    for (size_t i = 0; i < passes.size(); ++i) {
        const RPass &P = passes[i];
        if (e != hipSuccess) break;
        // the pass's workgroups per set within what its kernel keeps resident
        void *fn = fns.empty() ? nullptr : fns[i];
        const uint32_t pw = std::max<uint32_t>(1, std::min<uint32_t>(wps, adj_grid_cap(fn, s->device) / nsets));
        const uint32_t pg = (uint32_t)((uint64_t)nsets * pw);
        if (fn) {
            e = jit_launch_adj(fn, phi, &P.ts, sizeof P.ts, dg + (size_t)nsets * P.first, tps, pw, partial, pg, st);
        } else {
            k_adj_tile<<<pg, TT, 0, st>>>((dv2 *)phi, P.ts, dg + (size_t)nsets * P.first, P.count, tps, pw,
                                          P.nslots, partial);
            e = hipGetLastError();
        }
        if (e == hipSuccess && P.nslots) {
            k_sum_slots<<<dim3(P.nslots, nsets), 256, 0, st>>>(partial, pg, pw, nparam, slots, P.map);
            e = hipGetLastError();
        }
    }
    s->adj_reverse = passes.size();
    // the host tables are local: the uploads must be done before they go away
    const hipError_t f = hipStreamSynchronize(st);
    if (e == hipSuccess) e = f;
    return hip_status(e, "adjoint reverse passes over sets");
}

}  // namespace kbwh

// The code object of the compiled reverse-sweep kernel of pass `pass` of a
// flat state of num_qubits (11..40) qubits, compiled on the host.
extern "C" int32_t kbwh_compiled_reverse_pass_code(const kbwh_op *ops, uint64_t num_ops, uint32_t num_qubits,
                                                   uint32_t pass, void *code, uint64_t *size) {
    using namespace kbwh;
    if (!ops || !size) return set_error(KBWH_ERR_INVALID_ARGUMENT, "null argument");
    if (num_qubits < 11 || num_qubits > 40) return set_error(KBWH_ERR_INVALID_ARGUMENT, "11..40 qubits");
    for (uint64_t i = 0; i < num_ops; ++i)
        if (ops[i].kind > KRZ || ops[i].target >= num_qubits || (ops[i].ctrl_mask >> num_qubits))
            return set_error(KBWH_ERR_INVALID_ARGUMENT, "op out of range");
    int32_t rc;
    std::vector<RPass> passes;
    std::vector<AGate> table;
    std::vector<int32_t> pidx;
    if ((rc = plan_reverse(num_qubits, num_qubits, ops, num_ops, nullptr, passes, table, pidx))) return rc;
    if (pass >= passes.size()) return set_error(KBWH_ERR_INVALID_ARGUMENT, "no such pass");
    std::vector<char> obj;
    if ((rc = jit_code(JitPass{&table[passes[pass].first], passes[pass].count, kJitAdjoint}, obj))) return rc;
    if (code) {
        if (*size < obj.size()) return set_error(KBWH_ERR_INVALID_ARGUMENT, "buffer too small");
        memcpy(code, obj.data(), obj.size());
    }
    *size = obj.size();
    return KBWH_OK;
}
