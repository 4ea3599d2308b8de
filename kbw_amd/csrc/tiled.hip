// Multi-gate tiled passes (KBWH_RUN_TILED, KBWH_RUN_REORDER), Block swaps
// folded into passes, and the grouped Pauli-string sweeps.
//
// A pass applies a run of consecutive gates to the whole state in one HBM
// read and one HBM write.  The state is cut into tiles of 2^TL amplitudes
// (TL = 11 by default): a tile is the set of indices that vary over TL chosen
// qubits Q (the tile positions) with every other bit fixed by the tile
// number.  Q always holds qubits 0..2, so every 128-byte line of the state
// belongs to one tile and is read whole.  A workgroup of 256 threads holds
// one tile in registers, 2^(TL-8) amplitudes per thread:
//   tile position 0..5    -> lane bits      (DPP moves or cross-lane shuffles)
//   tile position 6..7    -> wave bits      (exchange through LDS)
//   tile position 8..TL-1 -> register index (no exchange)
// A non-diagonal gate (X, Y, H, RX, RY) needs its target in Q; diagonal gates
// (Z, P, RZ) and every control may sit anywhere (bits outside Q are constant
// over a tile).  KBWH_RUN_TILED applies gates in program order and every
// output equals the reference's per-pair closure (dense_v2.rs:166-277) in
// value, so a tiled run produces the same amplitudes as one pass per gate;
// only the number of HBM passes changes.  KBWH_RUN_REORDER plans the passes
// in a dependency order of the gates (schedule_ops) and runs them in the
// contracted kernel copy (tiled_fma.hip): the same unitary, equal to rounding.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "state.h"
#include "tile_core.h"

#pragma clang fp contract(off)
namespace kbwh {

namespace {

using namespace tile;

// One Pauli string of a grouped exp_value sweep: P|j> = i^ny (-1)^popcount(j & z) |j ^ x>,
// with x as a mask of tile positions (xt); coeff is the string's weight.
struct SString {
    uint64_t z;
    uint32_t ny, xt;
    double coeff;
};

__device__ __forceinline__ double block_sum(double v, double *ws) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < TT / 64; ++w) t += ws[w];
    __syncthreads();
    return t;
}

// sum_k coeff_k <P_k> over the strings of one group, whose X/Y qubits are all
// tile qubits of ts: each workgroup walks tiles t = blockIdx.x + j gridDim.x,
// reads each tile once (partners through LDS) and writes one partial sum.
__global__ __launch_bounds__(TT) __attribute__((amdgpu_waves_per_eu(KBWH_TILE_WAVES, KBWH_TILE_WAVES))) void k_pauli_tile(
    const dv2 *__restrict__ a, TileSpec ts, const SString *__restrict__ ss, uint32_t ns, uint32_t has_x,
    uint32_t ntiles, double *__restrict__ partial) {
    __shared__ dv2 lds[TT * TR];
    __shared__ double ws[TT / 64];
    const uint32_t tid = threadIdx.x;
    uint64_t tmask = 0;
#pragma unroll
    for (int b = 0; b < TB; ++b)
        if ((tid >> b) & 1u) tmask |= ts.qbit[b];
    double acc = 0.0;
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint64_t dt = tile_base(t, ts) | tmask;
        dv2 v[TR];
#pragma unroll
        for (int r = 0; r < TR; ++r) v[r] = __builtin_nontemporal_load(a + slot_offset(dt, r, ts));
        if (has_x) {
            __syncthreads();  // the previous tile's partner reads are done
#pragma unroll
            for (int r = 0; r < TR; ++r) lds[tid | (r << TB)] = v[r];
            __syncthreads();
        }
        for (uint32_t k = 0; k < ns; ++k) {
            const SString S = ss[k];
            // popcount(i & z) = popcount(dt & z) + popcount(r & zr), zr = z on the register bits
            uint32_t zr = 0;
#pragma unroll
            for (int b = 0; b < TL - TB; ++b)
                if (ts.qbit[TB + b] & S.z) zr |= 1u << b;
            const bool pd = (__popcll(dt & S.z) & 1) != 0;
            double sk = 0.0;
#pragma unroll
            for (int r = 0; r < TR; ++r) {
                bool flip = pd != ((__builtin_popcount((uint32_t)r & zr) & 1) != 0);
                double term;
                if (S.xt == 0) {
                    term = v[r].x * v[r].x + v[r].y * v[r].y;
                } else {
                    const dv2 o = lds[(tid | (r << TB)) ^ S.xt], m = v[r];
                    const double wr = o.x * m.x + o.y * m.y;
                    const double wi = o.x * m.y - o.y * m.x;
                    // Re(i^q (wr + i wi)): wr, -wi, -wr, wi for q = 0..3
                    const uint32_t q = S.ny & 3u;
                    term = (q & 1u) ? wi : wr;
                    flip = flip != (((q + 1u) & 2u) != 0);
                }
                sk += flip ? -term : term;
            }
            acc += S.coeff * sk;
        }
    }
    const double total = block_sum(acc, ws);
    if (tid == 0) partial[blockIdx.x] = total;
}

// lambda (+)= sum_k coeff_k P_k phi over the strings of one group, whose X/Y
// qubits are all tile qubits of ts; one tile per workgroup.  phi's tile is
// read once (partners through LDS) and lambda's tile read (assign = 0) and
// written once.  Per string and amplitude the arithmetic is k_pauli_apply's:
// (P phi)_i = i^q phi_{i^x}, q = ny + 2 popcount((i ^ x) & z), times coeff,
// added to lambda_i; here ss[k].ny holds ny + 2 popcount(x & z) (mod 4), so
// q = ny + 2 popcount(i & z).
__global__ __launch_bounds__(TT) __attribute__((amdgpu_waves_per_eu(KBWH_TILE_WAVES, KBWH_TILE_WAVES))) void k_pauli_apply_tile(
    const dv2 *__restrict__ phi, dv2 *__restrict__ lam, TileSpec ts, const SString *__restrict__ ss, uint32_t ns,
    uint32_t has_x, uint32_t ntiles, uint32_t assign) {
    __shared__ dv2 lds[TT * TR];
    const uint32_t tid = threadIdx.x;
    const uint32_t tile = blockIdx.x;
    if (tile >= ntiles) return;
    uint64_t tmask = 0;
#pragma unroll
    for (int b = 0; b < TB; ++b)
        if ((tid >> b) & 1u) tmask |= ts.qbit[b];
    const uint64_t dt = tile_base(tile, ts) | tmask;
    dv2 v[TR], out[TR];
#pragma unroll
    for (int r = 0; r < TR; ++r) v[r] = phi[slot_offset(dt, r, ts)];
    if (has_x) {
#pragma unroll
        for (int r = 0; r < TR; ++r) lds[tid | (r << TB)] = v[r];
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < TR; ++r) out[r] = assign ? mk2(0.0, 0.0) : lam[slot_offset(dt, r, ts)];
    for (uint32_t k = 0; k < ns; ++k) {
        const SString S = ss[k];
        // popcount(i & z) = popcount(dt & z) + popcount(r & zr), zr = z on the register bits
        uint32_t zr = 0;
#pragma unroll
        for (int b = 0; b < TL - TB; ++b)
            if (ts.qbit[TB + b] & S.z) zr |= 1u << b;
        const uint32_t qd = S.ny + 2u * (uint32_t)(__popcll(dt & S.z) & 1);
#pragma unroll
        for (int r = 0; r < TR; ++r) {
            const dv2 f = S.xt ? lds[(tid | (r << TB)) ^ S.xt] : v[r];
            const uint32_t q = (qd + 2u * (uint32_t)(__builtin_popcount((uint32_t)r & zr) & 1)) & 3u;
            // i^q f: (f.x, f.y), (-f.y, f.x), (-f.x, -f.y), (f.y, -f.x) for q = 0..3
            const double a = (q & 1u) ? f.y : f.x, b = (q & 1u) ? f.x : f.y;
            const double re = ((q + 1u) & 2u) ? -a : a, im = (q & 2u) ? -b : b;
            out[r] = mk2(out[r].x + re * S.coeff, out[r].y + im * S.coeff);
        }
    }
#pragma unroll
    for (int r = 0; r < TR; ++r) lam[slot_offset(dt, r, ts)] = out[r];
}

// Gate records of S parameter sets from one set's templates: record r of
// the templates becomes out[dst[r] + set * stride[r]] (a pass's records for
// set 0, then set 1, ...), with the coefficients of pidx[r]'s angle in that
// set, times sign (-1: the inverse gate), when pidx[r] >= 0.
__global__ __launch_bounds__(256) void k_expand_sets(const TGate *__restrict__ tmpl, const int32_t *__restrict__ pidx,
                                                     const uint32_t *__restrict__ dst,
                                                     const uint32_t *__restrict__ stride, uint32_t nrec,
                                                     uint32_t nsets, const double *__restrict__ params,
                                                     uint32_t nparams, double sign, TGate *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (uint64_t)nrec * nsets) return;
    const uint32_t set = (uint32_t)(i / nrec), r = (uint32_t)(i % nrec);
    TGate g = tmpl[r];
    if (pidx[r] >= 0) set_coefficients_dev(g, sign * params[(uint64_t)set * nparams + (uint32_t)pidx[r]]);
    out[dst[r] + (uint64_t)set * stride[r]] = g;
}

struct SGroup {
    TileSpec ts;
    uint32_t first, count, has_x;
};

// Strings put greedily into groups whose X/Y qubits fit one tile (qubits 0..2
// plus up to TL - 3 more); diagonal strings join the first group.  With
// apply_form, ny also takes the sign of popcount(x & z) (k_pauli_apply_tile).
int32_t group_strings(uint32_t n, const std::vector<PauliMasks> &ms, const double *coeffs, bool apply_form,
                      std::vector<SGroup> &groups, std::vector<SString> &strings) {
    const size_t ns = ms.size();
    std::vector<bool> done(ns, false);
    for (;;) {
        uint64_t qset = 7;
        bool any = false;
        for (size_t k = 0; k < ns; ++k) {
            if (done[k]) continue;
            any = true;
            if (__builtin_popcountll(qset | ms[k].x) <= TL) qset |= ms[k].x;
        }
        if (!any) break;
        for (uint32_t q = 3; q < n && __builtin_popcountll(qset) < TL; ++q) qset |= 1ull << q;
        uint32_t q_at[TL];
        int b = 0;
        for (uint32_t q = 0; q < n; ++q)
            if (qset >> q & 1ull) q_at[b++] = q;
        SGroup g;
        g.ts = make_tile_spec(q_at);
        g.first = (uint32_t)strings.size();
        g.has_x = 0;
        for (size_t k = 0; k < ns; ++k) {
            if (done[k] || (ms[k].x & ~qset)) continue;
            uint32_t ny = ms[k].ny;
            if (apply_form) ny = (ny + 2u * (uint32_t)(__builtin_popcountll(ms[k].x & ms[k].z) & 1)) & 3u;
            SString S{ms[k].z, ny, 0, coeffs[k]};
            for (int p = 0; p < TL; ++p)
                if (ms[k].x & g.ts.qbit[p]) S.xt |= 1u << p;
            if (S.xt) g.has_x = 1;
            strings.push_back(S);
            done[k] = true;
        }
        g.count = (uint32_t)strings.size() - g.first;
        if (g.count == 0) return set_error(KBWH_ERR_INVALID_ARGUMENT, "Pauli string with more than 8 X/Y qubits above qubit 2");
        groups.push_back(g);
    }
    return KBWH_OK;
}

// the string table into the tile gate buffer (stream-ordered reuse)
int32_t upload_strings(kbwh_state *s, const std::vector<SString> &strings) {
    const size_t bytes = strings.size() * sizeof(SString);
    if (s->tile_gates_cap < bytes) {
        hipError_t e = hipStreamSynchronize(s->stream);
        if (e == hipSuccess && s->tile_gates) e = hipFree(s->tile_gates);
        s->tile_gates = nullptr;
        s->tile_gates_cap = 0;
        if (e == hipSuccess) e = dev_alloc(&s->tile_gates, bytes);
        if (e != hipSuccess) return hip_status(e, "string table");
        s->tile_gates_cap = bytes;
    }
    if (bytes == 0) return KBWH_OK;
    return hip_status(hipMemcpyAsync(s->tile_gates, strings.data(), bytes, hipMemcpyHostToDevice, s->stream),
                      "string table");
}

}  // namespace

// exp_value of a flat state (n >= TL) in grouped sweeps: the strings are put
// greedily into groups whose X/Y qubits fit one tile (qubits 0..2 plus up to
// TL - 3 more); every group is one read-only sweep of the state.  Diagonal
// strings join the first group.
int32_t exp_value_tiled(kbwh_state *s, const std::vector<PauliMasks> &ms, const double *coeffs, double *out,
                        const ExpWork *work) {
    static_assert(sizeof(SString) == kExpStringBytes, "string records shared with observables.hip");
    const uint32_t n = s->n;
    if (s->block || n < (uint32_t)TL) return set_error(KBWH_ERR_INVALID_ARGUMENT, "grouped exp_value needs a flat state of at least 11 qubits");
    std::vector<SGroup> groups;
    std::vector<SString> strings;
    int32_t rc = group_strings(n, ms, coeffs, false, groups, strings);
    if (rc) return rc;
    if (groups.empty()) {
        *out = 0.0;
        return KBWH_OK;
    }
    if (groups.size() > kResultSlots) return set_error(KBWH_ERR_CAPACITY, "too many string groups");
    const SString *ds = nullptr;
    double *partial = s->partial, *results = s->result;
    hipError_t e = hipSuccess;
    if (work) {
        // the caller's work area: the table goes in with one stream-ordered
        // copy, and the host copy lives until reduce_from synchronizes
        if (strings.size() > work->strings_cap) return set_error(KBWH_ERR_CAPACITY, "string table larger than the work area");
        e = hipMemcpyAsync(work->strings, strings.data(), strings.size() * sizeof(SString), hipMemcpyHostToDevice,
                           s->stream);
        ds = (const SString *)work->strings;
        partial = work->partial;
        results = work->results;
    } else {
        rc = upload_strings(s, strings);
        if (rc) return rc;
        ds = (const SString *)s->tile_gates;
    }
    const uint32_t ntiles = (uint32_t)(1ull << (n - TL));
    const uint32_t grid = std::min<uint32_t>(ntiles, 2048);
    // with kbwh_set_compiled, each group's sweep runs a kernel compiled for
    // its strings (tile_jit.hip); a work-area call (no device allocation)
    // keeps the generic kernel, since loading a module allocates
    std::vector<void *> fns;
    if (s->compiled_passes && !work && e == hipSuccess) {
        std::vector<std::vector<SweepKey>> keys(groups.size());
        std::vector<JitPass> lst;
        for (size_t gi = 0; gi < groups.size(); ++gi) {
            const SGroup &g = groups[gi];
            for (uint32_t k = 0; k < g.count; ++k) {
                const SString &S = strings[g.first + k];
                SweepKey key{S.xt, 0u, S.ny & 3u, 0u};
                for (int b = 0; b < TL - TB; ++b)
                    if (g.ts.qbit[TB + b] & S.z) key.zr |= 1u << b;
                keys[gi].push_back(key);
            }
            lst.push_back(JitPass{keys[gi].data(), g.count, kJitSweep});
        }
        rc = jit_prepare(s->device, lst, fns);
        if (rc) {
            (void)hipStreamSynchronize(s->stream);  // the string upload reads the host table
            return rc;
        }
    }
    for (size_t gi = 0; gi < groups.size() && e == hipSuccess; ++gi) {
        const SGroup &g = groups[gi];
        if (gi < fns.size() && fns[gi]) {
            e = jit_launch_sweep(fns[gi], s->amp, &g.ts, ds + g.first, ntiles, partial, grid, s->stream);
        } else {
            k_pauli_tile<<<grid, TT, 0, s->stream>>>((const dv2 *)s->amp, g.ts, ds + g.first, g.count, g.has_x,
                                                      ntiles, partial);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = launch_sum_final(partial, grid, results + gi, s->stream);
    }
    if (e != hipSuccess) {
        // the string table's host copy must outlive an enqueued upload
        (void)hipStreamSynchronize(s->stream);
        return hip_status(e, "exp_value sweeps");
    }
    std::vector<double> vals(groups.size());
    rc = reduce_from(s, results, vals.data(), (uint32_t)groups.size());
    if (rc) return rc;
    double total = 0.0;
    for (double v : vals) total += v;
    *out = total;
    return KBWH_OK;
}

// lam = sum_k coeffs[k] P_k phi for vectors of s->n >= TL qubits, in grouped
// sweeps (one read of phi and one read-modify-write of lam per group).
int32_t pauli_apply_tiled(kbwh_state *s, const double2 *phi, double2 *lam, const std::vector<PauliMasks> &ms,
                          const double *coeffs, uint32_t n, uint64_t ntiles_total, uint32_t *sweeps) {
    if (n == 0) n = s->n;
    if (n < (uint32_t)TL) return set_error(KBWH_ERR_INVALID_ARGUMENT, "grouped Pauli sums need at least 11 qubits");
    std::vector<SGroup> groups;
    std::vector<SString> strings;
    int32_t rc = group_strings(n, ms, coeffs, true, groups, strings);
    if (rc) return rc;
    if (groups.empty()) return set_error(KBWH_ERR_INVALID_ARGUMENT, "no Pauli strings");
    if (sweeps) *sweeps = (uint32_t)groups.size();
    rc = upload_strings(s, strings);
    if (rc) return rc;
    const uint32_t ntiles = ntiles_total ? (uint32_t)ntiles_total : (uint32_t)(1ull << (n - TL));
    const SString *ds = (const SString *)s->tile_gates;
    hipError_t e = hipSuccess;
    for (size_t gi = 0; gi < groups.size() && e == hipSuccess; ++gi) {
        const SGroup &g = groups[gi];
        k_pauli_apply_tile<<<ntiles, TT, 0, s->stream>>>((const dv2 *)phi, (dv2 *)lam, g.ts, ds + g.first, g.count,
                                                          g.has_x, ntiles, gi == 0);
        e = hipGetLastError();
    }
    // the host string table is local: the upload must be done before it goes away
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    return hip_status(e, "Pauli sum sweeps");
}

}  // namespace kbwh

// The dependency-ordered pass plan of KBWH_RUN_REORDER on the host only.
extern "C" int32_t kbwh_reorder_plan(const kbwh_op *ops, uint64_t num_ops, uint32_t num_qubits,
                                     uint32_t tile_targets, uint32_t *order, uint32_t *pass_end,
                                     uint64_t *num_passes) {
    using namespace kbwh;
    if ((num_ops && (!ops || !order || !pass_end)) || !num_passes)
        return set_error(KBWH_ERR_INVALID_ARGUMENT, "null argument");
    if (num_qubits < 3 || num_qubits > 64) return set_error(KBWH_ERR_INVALID_ARGUMENT, "3..64 qubits");
    if (tile_targets < 1 || tile_targets > 64) return set_error(KBWH_ERR_INVALID_ARGUMENT, "tile targets must be 1..64");
    if (num_ops > 0xFFFFFFFFull) return set_error(KBWH_ERR_INVALID_ARGUMENT, "too many ops");
    std::vector<PhysOp> phys(num_ops);
    for (uint64_t i = 0; i < num_ops; ++i) {
        const kbwh_op &op = ops[i];
        if (op.kind > KRZ || op.target >= num_qubits || (num_qubits < 64 && (op.ctrl_mask >> num_qubits)))
            return set_error(KBWH_ERR_INVALID_ARGUMENT, "op out of range");
        phys[i] = PhysOp{(int)op.kind, op.target, op.ctrl_mask & ~(1ull << op.target), op.angle};
    }
    std::vector<uint32_t> ord, ends;
    schedule_passes(phys, num_qubits, std::min<uint32_t>(tile_targets, TL - 3), ord, ends);
    std::copy(ord.begin(), ord.end(), order);
    std::copy(ends.begin(), ends.end(), pass_end);
    *num_passes = ends.size();
    return KBWH_OK;
}

namespace kbwh {

namespace {

// Builds tiled passes from physical ops and Block swaps and launches them.
// A pass is planned in the qubits of the state as its tiles are loaded; ld[p]
// is the load-time qubit whose amplitudes now sit at physical qubit p (after
// the swaps folded into the pass so far) and ph is its inverse.  Each swap
// keeps both of its qubits in the tile; the pass stores the tile through the
// permuted map.
struct PassBuilder {
    static_assert(sizeof(TileSpec) == kTileSpecBytes && sizeof(TGate) == kTGateBytes,
                  "tile records shared with tiled_fma.hip");
    struct Launch {
        TileSpec spec, store;
        uint32_t first, count;
        bool perm;
    };
    kbwh_state *s;
    double2 *amp;
    uint32_t n, cap;
    bool contract;  // flat dependency-ordered runs: the contracted kernel copy
    bool compiled = false;  // ... with pass-specialised kernels where they compile
    Pass cur;
    std::vector<uint32_t> ld, ph;
    bool swapped = false;
    std::vector<TGate> gates;
    std::vector<Launch> launches;

    PassBuilder(kbwh_state *st, double2 *a, bool contracted)
        : s(st), amp(a), n(st->n), cap(std::min<uint32_t>(st->tile_targets, TL - 3)), contract(contracted),
          ld(st->n), ph(st->n) {
        for (uint32_t q = 0; q < n; ++q) ld[q] = ph[q] = q;
    }

    // closes the pending pass into a launch record
    int32_t flush() {
        if (cur.ops.empty() && !swapped) return KBWH_OK;
        int pos_of[64];
        uint32_t q_at[TL];
        assign_positions(n, cur, pos_of, q_at);
        Launch L;
        L.spec = make_tile_spec(q_at);
        L.store = L.spec;
        L.perm = swapped;
        if (swapped) {
            uint32_t q_st[TL];
            uint64_t in = 0, out = 0;
            for (int b = 0; b < TL; ++b) {
                q_st[b] = ph[q_at[b]];
                in |= 1ull << q_at[b];
                out |= 1ull << q_st[b];
            }
            if (in != out) return set_error(KBWH_ERR_DEVICE, "swapped qubit outside the tile");
            L.store = make_tile_spec(q_st);
        }
        L.first = (uint32_t)gates.size();
        for (const PhysOp &o : cur.ops) gates.push_back(make_tgate(o, pos_of, n));
        if (cur.ops.empty()) {
            // swaps only: one record that no tile applies (its control is
            // qubit 0, always a tile qubit, so never set in a tile base)
            TGate g;
            memset(&g, 0, sizeof g);
            g.kind = KZ;
            g.ocm = 1;
            gates.push_back(g);
        }
        L.count = (uint32_t)gates.size() - L.first;
        launches.push_back(L);
        cur = Pass();
        for (uint32_t q = 0; q < n; ++q) ld[q] = ph[q] = q;
        swapped = false;
        return KBWH_OK;
    }

    // uploads the gate records and launches every closed pass
    int32_t launch_all() {
        if (launches.empty()) return KBWH_OK;
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
        std::vector<void *> fns;
        if (compiled) {
            std::vector<JitPass> lst;
            lst.reserve(launches.size());
            for (const Launch &L : launches)
                lst.push_back(JitPass{&gates[L.first], L.count, L.perm ? kJitPermStore : 0u});
            const int32_t rc = jit_prepare(s->device, lst, fns);
            if (rc) return rc;
        }
        hipError_t e = hipMemcpyAsync(s->tile_gates, gates.data(), bytes, hipMemcpyHostToDevice, s->stream);
        if (e != hipSuccess) return hip_status(e, "tile gate upload");
        s->last_passes += launches.size();
        const uint32_t ntiles = (uint32_t)(1ull << (n - TL));
        const TGate *dg = (const TGate *)s->tile_gates;
        // compiled passes tune their tile order on their first runs (order_pick)
        std::vector<OrderTrial> trials;
        bool timing = false;
        for (size_t i = 0; i < launches.size() && e == hipSuccess; ++i) {
            const Launch &L = launches[i];
            if (!fns.empty() && fns[i]) {
                unsigned char key[2 * sizeof(TileSpec)];
                size_t kb = sizeof(TileSpec);
                memcpy(key, &L.spec, kb);
                if (L.perm) {
                    memcpy(key + kb, &L.store, sizeof(TileSpec));
                    kb += sizeof(TileSpec);
                }
                trials.push_back(order_pick(s->device, fns[i], key, kb, ntiles));
                OrderTrial &t = trials.back();
                timing = timing || t.slot;
                if (t.slot) e = hipEventRecord(t.ev[0], s->stream);
                if (e == hipSuccess)
                    e = jit_launch(fns[i], amp, &L.spec, dg + L.first, ntiles, s->stream, 32, ~0ull,
                                   L.perm ? &L.store : nullptr, t.order);
                if (e == hipSuccess && t.slot) e = hipEventRecord(t.ev[1], s->stream);
            } else if (contract && !L.perm) {
                e = launch_tile_contracted(amp, &L.spec, dg + L.first, L.count, ntiles, s->stream);
            } else {
                e = launch_tile_exact(amp, &L.spec, L.perm ? &L.store : nullptr, dg + L.first, L.count, ntiles,
                                      s->stream);
            }
        }
        if (timing) {
            // a run with order trials waits for them and reports their times
            const hipError_t se = hipStreamSynchronize(s->stream);
            for (OrderTrial &t : trials) {
                if (e == hipSuccess && se == hipSuccess) {
                    order_report(t);
                } else if (t.slot) {
                    (void)hipEventDestroy(t.ev[0]);
                    (void)hipEventDestroy(t.ev[1]);
                }
            }
            if (e == hipSuccess) e = se;
        }
        if (e != hipSuccess) return hip_status(e, "k_tile launch");
        // the gate buffer is reused by the next run: the upload above is
        // stream-ordered, and the next upload waits for these kernels
        launches.clear();
        gates.clear();
        return KBWH_OK;
    }

    // tile positions a swap of physical qubits a and b still needs in the pending pass
    uint32_t swap_needs(uint32_t a, uint32_t b) const {
        uint32_t need = 0;
        for (uint32_t q : {ld[a], ld[b]})
            if (q >= 3 && std::find(cur.targets.begin(), cur.targets.end(), q) == cur.targets.end()) ++need;
        return need;
    }
    void require(uint32_t q) {
        if (q >= 3 && std::find(cur.targets.begin(), cur.targets.end(), q) == cur.targets.end())
            cur.targets.push_back(q);
    }
    // folds the exchange of physical qubits a and b into the pending pass
    void fold_swap(uint32_t a, uint32_t b) {
        require(ld[a]);
        require(ld[b]);
        std::swap(ld[a], ld[b]);
        ph[ld[a]] = a;
        ph[ld[b]] = b;
        swapped = true;
    }
    // a gate on physical qubits, written against the load-time qubits
    PhysOp load_time(const PhysOp &o) const {
        uint64_t lm = 0;
        for (uint64_t r = o.ctrl; r; r &= r - 1) lm |= 1ull << ld[__builtin_ctzll(r)];
        PhysOp p = o;
        p.target = ld[o.target];
        p.ctrl = lm;
        return p;
    }

    // program order: an op joins the pending pass while its target fits
    int32_t program_order(const std::vector<kbwh_phys_event> &ev) {
        for (const kbwh_phys_event &E : ev) {
            if (E.swap_local != kNoSwap) {
                // a global target exchanged with a local qubit (block.rs:132-150)
                const uint32_t sl = E.swap_local, sg = E.swap_global;
                int32_t rc = KBWH_OK;
                if (cur.targets.size() + swap_needs(sl, sg) > cap) rc = flush();
                if (rc) return rc;
                if (swap_needs(sl, sg) > cap) {
                    // cannot share a tile under this target cap: a swap pass of its own
                    rc = launch_all();
                    if (rc) return rc;
                    hipError_t e = launch_swap_bits(amp, n, sl, sg, s->stream);
                    if (e != hipSuccess) return hip_status(e, "swap launch");
                    ++s->last_passes;
                } else {
                    fold_swap(sl, sg);
                }
                continue;
            }
            const PhysOp op{(int)E.kind, E.target, E.ctrl, E.angle};
            PhysOp po = load_time(op);
            if (!fits(cur, po, cap)) {
                int32_t rc = flush();
                if (rc) return rc;
                po = load_time(op);
            }
            add_op(cur, po);
        }
        int32_t rc = flush();
        if (rc == KBWH_OK) rc = launch_all();
        return rc;
    }
};

int32_t check_ops(uint32_t n, const kbwh_op *ops, uint64_t num_ops) {
    for (uint64_t i = 0; i < num_ops; ++i) {
        const kbwh_op &op = ops[i];
        if (op.kind > KRZ) return set_error(KBWH_ERR_INVALID_ARGUMENT, "unknown gate kind");
        if (op.target >= n) return set_error(KBWH_ERR_INVALID_ARGUMENT, "target out of range");
        if (n < 64 && (op.ctrl_mask >> n)) return set_error(KBWH_ERR_INVALID_ARGUMENT, "control out of range");
    }
    return KBWH_OK;
}

}  // namespace

int32_t run_tiled(kbwh_state *s, const kbwh_op *ops, uint64_t num_ops, const double *params) {
    return run_tiled_on(s, s->amp, ops, num_ops, params, false);
}

int32_t run_sets_forward(kbwh_state *s, double2 *amp, uint32_t n, uint32_t nsets, const kbwh_op *ops,
                         uint64_t num_ops, const double *params_dev, uint32_t num_params, kbwh_set_buffers &sb) {
    if (n < (uint32_t)TL || n > 40) return set_error(KBWH_ERR_INVALID_ARGUMENT, "parameter sets need 11..40 qubits");
    if (nsets == 0 || ((uint64_t)nsets << (n - TL)) > 0x7FFFFFFFull)
        return set_error(KBWH_ERR_INVALID_ARGUMENT, "parameter set count out of range");
    int32_t rc = check_ops(n, ops, num_ops);
    if (rc) return rc;
    // the structure is the same in every set: parameterised gates get their
    // coefficients on the device (k_expand_sets); no diagonal fusion, whose
    // choices depend on the angles
    std::vector<PhysOp> phys(num_ops);
    for (uint64_t i = 0; i < num_ops; ++i) {
        const kbwh_op &op = ops[i];
        phys[i] = PhysOp{(int)op.kind, op.target, op.ctrl_mask & ~(1ull << op.target), op.param >= 0 ? 0.0 : op.angle};
    }
    std::vector<uint32_t> order, ends;
    schedule_passes(phys, n, std::min<uint32_t>(s->tile_targets, TL - 3), order, ends);
    if (order.size() != num_ops) return set_error(KBWH_ERR_DEVICE, "set schedule incomplete");
    struct SL {
        TileSpec ts;
        uint32_t first, count;
    };
    std::vector<SL> launches;
    std::vector<TGate> tmpl;
    std::vector<int32_t> pidx;
    std::vector<uint32_t> dst, stride;
    size_t k = 0;
    for (const uint32_t end : ends) {
        Pass cur;
        for (size_t i = k; i < end; ++i) add_op(cur, phys[order[i]]);
        int pos_of[64];
        uint32_t q_at[TL];
        assign_positions(n, cur, pos_of, q_at);
        SL L;
        L.ts = make_tile_spec(q_at);
        L.first = (uint32_t)tmpl.size();
        L.count = (uint32_t)(end - k);
        for (size_t i = k; i < end; ++i) {
            tmpl.push_back(make_tgate(phys[order[i]], pos_of, n));
            pidx.push_back(ops[order[i]].param);
            dst.push_back((uint32_t)((uint64_t)nsets * L.first + (i - k)));
            stride.push_back(L.count);
        }
        launches.push_back(L);
        k = end;
    }
    const size_t nrec = tmpl.size();
    if (nrec == 0) return KBWH_OK;
    if ((uint64_t)nrec * nsets > 0xFFFFFFFFull) return set_error(KBWH_ERR_INVALID_ARGUMENT, "too many gate records");
    const size_t tbytes = nrec * (sizeof(TGate) + 12);
    if ((rc = grow_device(s, &sb.tmpl, &sb.tmpl_cap, tbytes))) return rc;
    if ((rc = grow_device(s, &sb.gates, &sb.gates_cap, nrec * nsets * sizeof(TGate)))) return rc;
    char *tb = (char *)sb.tmpl;
    TGate *d_tmpl = (TGate *)tb;
    int32_t *d_pidx = (int32_t *)(tb + nrec * sizeof(TGate));
    uint32_t *d_dst = (uint32_t *)(d_pidx + nrec), *d_stride = d_dst + nrec;
    hipError_t e = hipMemcpyAsync(d_tmpl, tmpl.data(), nrec * sizeof(TGate), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_pidx, pidx.data(), nrec * 4, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_dst, dst.data(), nrec * 4, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_stride, stride.data(), nrec * 4, hipMemcpyHostToDevice, s->stream);
    TGate *d_gates = (TGate *)sb.gates;
    if (e == hipSuccess) {
        const uint64_t total = (uint64_t)nrec * nsets;
        k_expand_sets<<<(uint32_t)((total + 255) / 256), 256, 0, s->stream>>>(
            d_tmpl, d_pidx, d_dst, d_stride, (uint32_t)nrec, nsets, params_dev, num_params, 1.0, d_gates);
        e = hipGetLastError();
    }
    const uint32_t ntiles = nsets << (n - TL);
    std::vector<void *> fns;
    if (s->compiled_passes && e == hipSuccess) {
        std::vector<JitPass> lst;
        for (const SL &L : launches) lst.push_back(JitPass{&tmpl[L.first], L.count, 0u});
        if ((rc = jit_prepare(s->device, lst, fns))) {
            (void)hipStreamSynchronize(s->stream);
            return rc;
        }
    }
    for (size_t i = 0; i < launches.size() && e == hipSuccess; ++i) {
        const SL &L = launches[i];
        if (!fns.empty() && fns[i])
            e = jit_launch(fns[i], amp, &L.ts, d_gates + (size_t)nsets * L.first, ntiles, s->stream, n - TL);
        else
            e = launch_tile_contracted(amp, &L.ts, d_gates + (size_t)nsets * L.first, L.count, ntiles, s->stream,
                                       n - TL);
    }
    s->last_passes = launches.size();
    s->last_gates = num_ops;
    // the host tables are local: the uploads must be done before they go away
    const hipError_t f = hipStreamSynchronize(s->stream);
    if (e == hipSuccess) e = f;
    return hip_status(e, "parameter-set passes");
}

int32_t run_events(kbwh_state *s, double2 *amp, const std::vector<kbwh_phys_event> &ev) {
    s->last_passes = 0;
    if (s->n < TL) {
        // the whole state is smaller than one tile: one launch per op
        for (const kbwh_phys_event &E : ev) {
            const hipError_t e = E.swap_local != kNoSwap
                                     ? launch_swap_bits(amp, s->n, E.swap_local, E.swap_global, s->stream)
                                     : launch_gate(amp, s->n, E.kind, E.angle, E.target, E.ctrl, s->stream);
            if (e != hipSuccess) return hip_status(e, "gate launch");
            ++s->last_passes;
        }
        return KBWH_OK;
    }
    PassBuilder pb(s, amp, false);
    return pb.program_order(ev);
}

int32_t run_tiled_on(kbwh_state *s, double2 *amp, const kbwh_op *ops, uint64_t num_ops,
                     const double *params, bool reorder, bool compiled) {
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
    int32_t rc = check_ops(n, ops, num_ops);
    if (rc) return rc;
    if (compiled && reorder && !s->block && n >= kWideTileQubits) {
        bool done = false;
        rc = run_compiled_wide(s, amp, ops, num_ops, params, &done);
        if (rc || done) return rc;
    }
    PassBuilder pb(s, amp, reorder && !s->block);
    pb.compiled = compiled && reorder;
    // the map replayed in program order, exactly as gate by gate: a Block
    // state's swaps become events before the gates that need them
    auto to_events = [&](std::vector<kbwh_phys_event> &ev) {
        ev.reserve(num_ops + num_ops / 4);
        for (uint64_t i = 0; i < num_ops; ++i) {
            const kbwh_op &op = ops[i];
            const double angle = (params && op.param >= 0) ? params[op.param] : op.angle;
            uint32_t t = op.target, sl = kNoSwap, sg = kNoSwap;
            if (s->block) map_qubit(s, op.target, &t, &sl, &sg);
            if (sl != kNoSwap) ev.push_back(kbwh_phys_event{KX, 0, 0, 0.0, sl, sg});
            const uint64_t m = physical_mask(s, op.ctrl_mask) & ~(1ull << t);
            ev.push_back(kbwh_phys_event{(int32_t)op.kind, t, m, angle, kNoSwap, kNoSwap});
        }
    };
    if (reorder && s->block && pb.cap >= 2) {
        // A Block state: the LRU map and its swaps follow program order (the
        // map is replayed first, exactly as gate by gate), then swaps and
        // gates are scheduled as ops on physical qubits.  A swap needs both of
        // its qubits in the tile and is folded into its pass.  The ops
        // touching each physical qubit keep their order, so the relabelling
        // of every qubit is the same as in program order.
        std::vector<kbwh_phys_event> ev;
        to_events(ev);
        std::vector<uint64_t> qm(ev.size()), req(ev.size());
        for (size_t i = 0; i < ev.size(); ++i) {
            const kbwh_phys_event &E = ev[i];
            if (E.swap_local != kNoSwap) {
                qm[i] = req[i] = (1ull << E.swap_local) | (1ull << E.swap_global);
            } else {
                qm[i] = E.ctrl | (1ull << E.target);
                req[i] = is_diagonal(E.kind) ? 0ull : (1ull << E.target);
            }
        }
        std::vector<uint32_t> order, ends;
        schedule_ops(qm, req, n, pb.cap, order, ends);
        if (order.size() != ev.size()) return set_error(KBWH_ERR_DEVICE, "Block schedule incomplete");
        size_t k = 0;
        for (const uint32_t end : ends) {
            for (; k < end; ++k) {
                const kbwh_phys_event &E = ev[order[k]];
                if (E.swap_local != kNoSwap)
                    pb.fold_swap(E.swap_local, E.swap_global);
                else
                    add_op(pb.cur, pb.load_time(PhysOp{(int)E.kind, E.target, E.ctrl, E.angle}));
            }
            if ((rc = pb.flush())) return rc;
        }
        rc = pb.launch_all();
        s->last_gates = num_ops;
        return rc;
    }
    if (reorder && !s->block) {
        // dependency-ordered passes (a Block state under a target cap of one
        // keeps program order: its swaps cannot share a tile)
        std::vector<PhysOp> phys(num_ops);
        for (uint64_t i = 0; i < num_ops; ++i) {
            const kbwh_op &op = ops[i];
            const double angle = (params && op.param >= 0) ? params[op.param] : op.angle;
            phys[i] = PhysOp{(int)op.kind, op.target, op.ctrl_mask & ~(1ull << op.target), angle};
        }
        fuse_diagonals(phys, n);
        std::vector<uint32_t> order, ends;
        // the plan search: by default for compiled runs only (their first
        // call compiles kernels anyway); plans are kept per structure
        const uint32_t trials = s->plan_trials >= 0 ? (uint32_t)s->plan_trials : (pb.compiled ? 256u : 0u);
        schedule_passes_cached(phys, n, pb.cap, trials, order, ends);
        size_t k = 0;
        for (const uint32_t end : ends) {
            for (; k < end; ++k) add_op(pb.cur, phys[order[k]]);
            if ((rc = pb.flush())) return rc;
        }
        rc = pb.launch_all();
        s->last_gates = num_ops;
        return rc;
    }
    std::vector<kbwh_phys_event> ev;
    to_events(ev);
    rc = pb.program_order(ev);
    s->last_gates = num_ops;
    return rc;
}

}  // namespace kbwh
